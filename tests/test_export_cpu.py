"""Factor export (Solver.getP/getQ/getL/getU/get_rearranged_matrix; reference
src/solver.cpp:1541-1725, test apps/tests/main.cpp:170-195): a global sparse LU
with KLU's export settings (btf = 0, COLAMD, threshold pivoting tol 0.001),
built on the host from set_matrix alone (no analyze, no device).

Checks: the reference's own known case (rearranged == L U); the invariants that
fix the meaning of each export (P and Q permutations with P A Q == rearranged,
L unit lower with |L| <= 1/tol, U upper, |PAQ - LU| within the LU backward-error
bound); row scaling (KLU scale = 2); error paths; and, for a matrix whose BTF is
the identity, P and Q against the oracle's KLU ordering of that one block.
The exact pivot sequence and values of SuiteSparse's KLU cannot be compared
here (no SuiteSparse build): those stay parity unpinned beyond the oracle."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp


def _csr(entries, n):
    a = np.array(entries, dtype=np.float64).reshape(-1, 3)
    return sp.csr_matrix((a[:, 2], (a[:, 0].astype(np.int64), a[:, 1].astype(np.int64))), shape=(n, n))


def _exports(s, n):
    out = {}
    for name in ("get_rearranged_matrix", "getP", "getQ", "getL", "getU"):
        e = getattr(s, name)()
        assert s.last_error() == 0, (name, s.last_error_message())
        out[name] = _csr(e, n)
    return out


def _check_invariants(n, r, c, v, ex, tol=0.001, scaled=None):
    A = sp.csr_matrix((v, (r, c)), shape=(n, n))
    P, Q, R, L, U = ex["getP"], ex["getQ"], ex["get_rearranged_matrix"], ex["getL"], ex["getU"]
    for M in (P, Q):  # permutation matrices
        assert M.nnz == n and np.all(M.data == 1.0)
        assert np.array_equal(np.sort(M.indices), np.arange(n))
        assert np.all(np.diff(M.indptr) == 1)
    As = A if scaled is None else sp.diags(1.0 / scaled) @ A
    assert abs(P @ As @ Q - R).max() <= 1e-15 * abs(As).max()
    assert sp.triu(L, 1).nnz == 0 and np.allclose(L.diagonal(), 1.0)
    assert sp.tril(U, -1).nnz == 0 and np.all(U.diagonal() != 0)
    assert abs(L).max() <= 1.0 / tol * (1 + 1e-12)
    D = abs(R - L @ U)
    bound = (abs(L) @ abs(U)).multiply(4 * n * np.finfo(float).eps)
    excess = (D - bound).tocoo()
    assert excess.nnz == 0 or excess.data.max() <= 0.0


def test_reference_known_case(bd):
    """apps/tests/main.cpp:170-195: rearranged == L * U on the 4x4 case."""
    m = [(0, 0, 100.0), (0, 1, 20.0), (0, 2, 7.0), (1, 0, 30.0), (1, 1, 80.0), (1, 3, 5.0),
         (2, 2, 0.25), (2, 3, 0.05), (3, 2, 0.10), (3, 3, 0.50)]
    r, c, v = (np.array([e[k] for e in m]) for k in range(3))
    s = bd.Solver()
    assert s.set_matrix(r.astype(np.int32), c.astype(np.int32), v, 4)
    ex = _exports(s, 4)
    assert len(s.get_rearranged_matrix()) > 0 and len(s.getL()) > 0 and len(s.getU()) > 0
    np.testing.assert_allclose(ex["get_rearranged_matrix"].toarray(), (ex["getL"] @ ex["getU"]).toarray(),
                               rtol=0, atol=1e-12)
    _check_invariants(4, r.astype(np.int64), c.astype(np.int64), v, ex)
    s.close()


def test_invariants_on_generated_workload(bd):
    sp_ = bd.config_spec(0)
    sp_.n_small_blocks, sp_.n_small_total = 40, 800
    n, r, c, v, _ = bd.generate(sp_)
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n)
    ex = _exports(s, n)
    _check_invariants(n, r, c, v, ex)
    s.close()


def test_row_scaling(bd, monkeypatch):
    """KLU scale = 2 (row maxima): rearranged = P (R \\ A) Q with R by original
    row, every row of R \\ A has maximum magnitude 1, and it still equals L U."""
    monkeypatch.setenv("USERALGO_KLU_SCALE", "2")
    sp_ = bd.config_spec(3)
    sp_.n_small_blocks, sp_.n_small_total = 30, 600
    sp_.n_large_blocks = 0
    n, r, c, v, _ = bd.generate(sp_)
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n)
    ex = _exports(s, n)
    A = sp.csr_matrix((v, (r, c)), shape=(n, n))
    rmax = abs(A).max(axis=1).toarray().ravel()
    _check_invariants(n, r, c, v, ex, scaled=rmax)
    Rm = ex["get_rearranged_matrix"]
    np.testing.assert_allclose(abs(Rm).max(axis=1).toarray().ravel(), 1.0, rtol=1e-15)
    s.close()


@pytest.mark.parametrize("name,scaled", [("USERALGO_KLU_SCALE", True), ("BLOCKDAGLU__KLU_SCALE", True),
                                         ("BLOCKDAGLU_KLU_SCALE", False)])
def test_env_alias_rule(bd, monkeypatch, name, scaled):
    """The reference reads USERALGO_X, then BLOCKDAGLU_ + X with X keeping its
    leading underscore (include/blockdaglu/env.h:19-21 compares 8 characters),
    so BLOCKDAGLU__KLU_SCALE is an alias and BLOCKDAGLU_KLU_SCALE is ignored."""
    for k in ("USERALGO_KLU_SCALE", "BLOCKDAGLU__KLU_SCALE", "BLOCKDAGLU_KLU_SCALE"):
        monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv(name, "2")
    n = 6
    r = np.array([0, 1, 2, 3, 4, 5, 0, 2, 4], np.int32)
    c = np.array([0, 1, 2, 3, 4, 5, 1, 3, 5], np.int32)
    v = np.array([4.0, 8.0, 2.0, 5.0, 3.0, 7.0, 1.0, 1.0, 1.0])
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n)
    Rm = s.get_rearranged_matrix()
    peak = max(abs(x[2]) for x in Rm)
    # Row maxima scaling makes every rearranged entry at most 1 in magnitude.
    assert (peak <= 1.0) == scaled
    s.close()


def test_errors(bd):
    s = bd.Solver()
    assert s.getL() == [] and s.last_error() == bd.SolverError.NotPrepared
    # Structurally singular: column 1 is empty.
    assert s.set_matrix(np.array([0, 1], np.int32), np.array([0, 0], np.int32), np.array([1.0, 2.0]), 2)
    assert s.getU() == [] and s.last_error() == bd.SolverError.ExportFailed
    # Invalid index: the export validates like analyze.
    assert s.set_matrix(np.array([0, 5], np.int32), np.array([0, 1], np.int32), np.array([1.0, 2.0]), 2)
    assert s.getP() == [] and s.last_error() == bd.SolverError.InvalidIndex
    s.close()


def test_orders_match_oracle_on_one_block(bd, orc):
    """A zero-free diagonal makes the maximum transversal the identity, and a
    cycle through every node makes the matrix one strongly connected block, so
    the BLOCK route's sparse factor of that block (the oracle) and the export
    see the same matrix with the same KLU settings: same Q and P."""
    rng = np.random.default_rng(17)
    n = 400
    rows = list(range(n)) + list(range(n - 1)) + [n - 1]
    cols = list(range(n)) + list(range(1, n)) + [0]
    extra = 3 * n
    rows += rng.integers(0, n, extra).tolist()
    cols += rng.integers(0, n, extra).tolist()
    r = np.array(rows, np.int32)
    c = np.array(cols, np.int32)
    v = rng.normal(size=len(r))
    v[:n] += 0.5 * np.sign(v[:n])
    o = orc.OracleSolver(r, c, v)
    assert o.analyze() == 0 and o.factor() == 0
    P0, Q0, R0 = o.permutations()
    assert len(R0) == 2 and np.array_equal(P0, np.arange(n)) and np.array_equal(Q0, np.arange(n))
    q, prow = o.sparse_order(0, n)
    assert q is not None
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n)
    Pe = np.array([e[1] for e in sorted(s.getP())], np.int32)
    Qe = np.array([e[0] for e in sorted(s.getQ(), key=lambda e: e[1])], np.int32)
    assert np.array_equal(Qe, q)
    assert np.array_equal(Pe, prow)
    s.close()
