"""Pins the oracle (the CPU restatement of the reference BLOCK route under
oracle/) against the reference's own known answers and against what mathematics
and independent implementations fix (exact sums, scipy's matching and strongly
connected components, LAPACK getrf, SuperLU). CPU only; the fixtures under
tests/golden/ come from tests/golden/make_golden.py."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

from conftest import backward_error, check_btf, rel_residual

GOLD = Path(__file__).resolve().parent / "golden"


def _csc(n, r, c):
    """Pattern CSC, rows ascending within a column (reference sparse_formats.h:23-53)."""
    A = sp.csc_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32)


def _cache_matrix(n, changed):
    # reference apps/tests/main.cpp:130-143
    m = []
    for row in range(n):
        m.append((row, row, 4.0))
        if row + 2 < n:
            m.append((row, changed if row == n // 2 else row + 1, 0.25))
        else:
            m.append((row, row, 0.5))
    r, c, v = (np.array(t) for t in zip(*m))
    return r, c, v.astype(np.float64)


def test_known_answer_duplicate_3x3(orc):
    # reference apps/tests/main.cpp:111-128 (expected x at :117)
    r = np.array([0, 0, 1, 1, 1, 2, 2, 0, 0])
    c = np.array([0, 1, 0, 1, 2, 1, 2, 0, 0])
    v = np.array([4.0, 1.0, 1.0, 3.0, 1.0, 1.0, 2.0, 1.0, -1.0])
    o = orc.OracleSolver(r, c, v, 3)
    assert o.analyze() == 0 and o.factor() == 0
    rc, x = o.solve(np.array([1.0, 2.0, 3.0]))
    assert rc == 0
    np.testing.assert_allclose(x, [0.2222222222222222, 0.1111111111111111, 1.4444444444444444], rtol=1e-12)


@pytest.mark.parametrize("changed,scale", [(551, 1.0), (552, 1.0), (551, 2.0)])
def test_known_answer_cache_matrices_all_ones(orc, changed, scale):
    # reference apps/tests/main.cpp:145-168: x = 1 for b = A 1
    n = 1100
    r, c, v = _cache_matrix(n, changed)
    v = v * scale
    b = np.zeros(n)
    np.add.at(b, r, v)
    o = orc.OracleSolver(r, c, v, n)
    assert o.analyze() == 0 and o.factor() == 0
    rc, x = o.solve(b)
    assert rc == 0
    np.testing.assert_allclose(x, np.ones(n), rtol=1e-12)


@pytest.mark.parametrize("rows,cols,vals,dim,code", [
    ([-1], [0], [1.0], -1, 3),               # negative index  (tests:94-97)  InvalidIndex
    ([0], [0], [np.inf], -1, 4),             # non-finite value (tests:99-102) NonFiniteValue
    ([0], [0], [np.nan], -1, 4),
    ([2], [2], [1.0], 2, 3),                 # outside declared dimension (tests:104-107)
])
def test_invalid_inputs(orc, rows, cols, vals, dim, code):
    o = orc.OracleSolver(np.array(rows), np.array(cols), np.array(vals, np.float64), dim)
    assert o.analyze() == code


def test_canonicalize_exact():
    """Sorted by (row, col), duplicates summed, exact zeros dropped
    (reference src/solver.cpp:642-722), against exact sums of small integers."""
    from oracle import oracle as orc
    g = np.load(GOLD / "canon.npz")
    for k in np.unique(g["case"]):
        sel = g["case"] == k
        o = orc.OracleSolver(g["rows"][sel], g["cols"][sel], g["vals"][sel])
        o.analyze()
        r, c, v = o.canonical()
        es = g["exp_case"] == k
        np.testing.assert_array_equal(r, g["exp_rows"][es])
        np.testing.assert_array_equal(c, g["exp_cols"][es])
        np.testing.assert_array_equal(v, g["exp_vals"][es])


def test_btf_rank_and_fine_blocks(orc):
    g = np.load(GOLD / "btf.npz")
    for k in range(int(g["ncases"])):
        n, r, c = int(g[f"n{k}"]), g[f"r{k}"], g[f"c{k}"]
        Ap, Ai = _csc(n, r, c)
        P, Q, R, nmatch = orc.btf_order(n, Ap, Ai)
        check_btf(n, r, c, P, Q, R, nmatch, g, k)


def _merge_invariants(R, M, n, merge_max):
    small = max(1, min(merge_max, 32))
    R, M = list(R), list(M)
    if len(R) - 1 <= 1:
        assert M == [0, n]
        return
    assert M[0] == 0 and M[-1] == n and set(M) <= set(R) and M == sorted(set(M))
    kinds = []
    for a, b in zip(M[:-1], M[1:]):
        parts = [(x, y) for x, y in zip(R[:-1], R[1:]) if a <= x and y <= b]
        sizes = [y - x for x, y in parts]
        if len(parts) == 1 and sizes[0] > small:
            kinds.append(("large", sizes))
        else:
            assert all(s <= small for s in sizes) and sum(sizes) <= merge_max
            kinds.append(("run", sizes))
    for (k1, s1), (k2, s2) in zip(kinds[:-1], kinds[1:]):
        if k1 == "run" and k2 == "run":
            assert sum(s1) + s2[0] > merge_max, "greedy run stopped early"


@pytest.mark.parametrize("merge_max", [64, 40, 16, 1])
def test_merge_invariants(orc, merge_max):
    # reference ubtf_block_solver.cpp:393-476 (greedy runs of small blocks)
    rng = np.random.default_rng(merge_max)
    for trial in range(40):
        nb = int(rng.integers(1, 60))
        sizes = np.where(rng.random(nb) < 0.8, rng.integers(1, 33, nb), rng.integers(33, 120, nb))
        R = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        n = int(R[-1])
        M = orc.merge(R, n, merge_max)
        _merge_invariants(R, M, n, merge_max)


def test_dense_lu_matches_lapack(orc):
    """Unblocked right-looking LU with partial pivoting (first largest |a_ik|),
    the reference dense branch (ubtf:1465-1494, :2102-2135): pivots equal
    LAPACK getrf's, values to rounding."""
    g = np.load(GOLD / "dense.npz")
    for k in range(int(g["ncases"])):
        a = g[f"a{k}"]
        s = a.shape[0]
        r, c = np.nonzero(np.ones((s, s)))
        o = orc.OracleSolver(r, c, a[r, c], s)
        assert o.analyze() == 0 and o.factor() == 0
        P, Q, R = o.permutations()
        assert len(R) == 2  # a dense matrix is one strongly connected block
        lu, piv = o.dense_factor(0, s)
        ref_lu, ref_piv = sla.lu_factor(a[np.ix_(P, Q)])
        np.testing.assert_array_equal(piv, ref_piv)
        np.testing.assert_allclose(lu, ref_lu, rtol=1e-9, atol=1e-12 * np.abs(ref_lu).max())


@pytest.mark.parametrize("a,piv", [
    # |a_00| == |a_10|: the first row wins, no swap (strict '>' in the reference
    # pivot search, ubtf:2110-2116; LAPACK idamax also returns the first).
    ([[1.0, 2.0], [-1.0, 3.0]], [0, 1]),
    ([[-2.0, 1.0, 0.0], [2.0, 0.0, 1.0], [1.0, 1.0, 1.0]], [0, 2, 2]),
    # tie after one elimination step (exact in floating point)
    ([[4.0, 2.0, 1.0], [2.0, 0.0, 3.0], [2.0, 2.0, 1.0]], [0, 1, 2]),
])
def test_dense_lu_tie_rule(orc, a, piv):
    a = np.array(a)
    s = a.shape[0]
    r, c = np.nonzero(a)
    o = orc.OracleSolver(r, c, a[r, c], s)
    assert o.analyze() == 0 and o.factor() == 0
    P, Q, R = o.permutations()
    if len(R) != 2:
        pytest.skip("not one strongly connected block under this BTF order")
    lu, p = o.dense_factor(0, s)
    ref_lu, ref_piv = sla.lu_factor(a[np.ix_(P, Q)])
    np.testing.assert_array_equal(p, ref_piv)
    if np.array_equal(P, np.arange(s)) and np.array_equal(Q, np.arange(s)):
        np.testing.assert_array_equal(p, piv)


def test_solve_matches_superlu(orc):
    g = np.load(GOLD / "solve.npz")
    for k in range(int(g["ncases"])):
        r, c, v, b, x_ref = g[f"r{k}"], g[f"c{k}"], g[f"v{k}"], g[f"b{k}"], g[f"x{k}"]
        n = len(b)
        o = orc.OracleSolver(r, c, v, n)
        assert o.analyze() == 0 and o.factor() == 0
        rc, x = o.solve(b)
        assert rc == 0
        assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-10
        assert rel_residual(n, r, c, v, x, b) < 1e-12


def test_klu_blocks_config0(bd, orc):
    """Blocks above 64 go through the KLU-rule sparse LU; whole-system residual,
    agreement with SuperLU, and the refinement stop rule (src/solver.cpp:1508-1521)."""
    n, r, c, v, B = bd.generate(0)
    b = B[:, 0].copy()
    o = orc.OracleSolver(r, c, v)
    assert o.analyze() == 0 and o.factor() == 0
    info = o.info()
    P, Q, R = o.permutations()
    assert np.max(np.diff(R)) > 64  # the case exercises the sparse branch
    rc, x = o.solve(b)
    assert rc == 0
    A = sp.csc_matrix((v, (r, c)), shape=(n, n))
    x_ref = spsolve(A, b)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-8
    assert rel_residual(n, r, c, v, x, b) < 1e-9
    rc, x2 = o.solve(b, refine_iters=3)
    assert rc == 0
    assert rel_residual(n, r, c, v, x2, b) < 1e-12
    assert backward_error(n, r, c, v, x2, b) < 1e-14
    assert info["nmatch"] == n


def test_structurally_singular_fails_at_factor(orc):
    # nmatch < n passes analyze (ubtf:680-699 only records it); factor fails.
    r = np.array([0, 1])
    c = np.array([1, 1])
    o = orc.OracleSolver(r, c, np.array([1.0, 2.0]))
    assert o.analyze() == 0
    assert o.info()["nmatch"] == 1
    assert o.factor() != 0
