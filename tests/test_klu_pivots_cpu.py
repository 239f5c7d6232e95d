"""First-factorization pivot rows of the KLU branch (klu_factor on one block,
ubtf_block_solver.cpp:1806: threshold partial pivoting, tol 0.001, diagonal
preference with the diagonal of column k initially its COLAMD row Q[k], and
KLU's diagonal tracking: when an off-diagonal pivot row r is taken at column
k, the row that was k's diagonal becomes the diagonal of the later column whose
diagonal r was).

Pins:
  * the oracle against SuperLU (scipy splu: COLAMD column order, no
    elimination-tree postorder, threshold 0.001 with diagonal preference, no
    equilibration): the same pivot rows up to the first column whose diagonal
    KLU's tracking has moved (SuperLU has no tracking), on blocks with 0 to 15
    off-diagonal pivots;
  * the product host analysis against the oracle: identical column order, and
    identical pivot rows over the whole block when the dense Schur tail is
    analysed with KLU's rule (BLOCKDAGLU_GPU_TAIL_TOL=0.001; the default
    partial pivoting in the tail keeps |L| <= 1 for the device's explicit-
    inverse panels, and then only the sparse part's rows are KLU's). CPU only."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def _block(s, seed, f):
    """Strongly connected s x s block: a cycle, Poisson(4) random entries per
    row, diagonal f * (row |sum|) + U(0.1, 1) (smaller f: more off-diagonal pivots)."""
    rng = np.random.default_rng(seed)
    R, Cc, V = [], [], []
    for i in range(s):
        R.append(i)
        Cc.append((i + 1) % s)
        V.append(rng.normal())
        for j in rng.choice(s, min(int(rng.poisson(4)), s - 2), replace=False):
            if j != i and j != (i + 1) % s:
                R.append(i)
                Cc.append(int(j))
                V.append(rng.normal())
    A = sp.csr_matrix((V, (R, Cc)), shape=(s, s))
    A.sum_duplicates()
    A = (A + sp.diags(f * np.asarray(abs(A).sum(axis=1)).ravel() + rng.uniform(0.1, 1, s))).tocoo()
    return A.row.astype(np.int32), A.col.astype(np.int32), A.data


def _first_tracked(q, prow):
    """First column whose diagonal row differs from Q[k] because of tracking."""
    s = len(q)
    diag = [int(x) for x in q]
    diag_col = {int(q[k]): k for k in range(s)}
    for k in range(s):
        if diag[k] != q[k]:
            return k
        r, d = int(prow[k]), diag[k]
        kbar = diag_col.get(r)
        if r != d and kbar is not None and kbar > k:
            diag[kbar] = d
            diag_col[d] = kbar
    return s


CASES = [(int(np.random.default_rng(t).integers(70, 400)), t, (0.5, 0.1, 0.01, 0.0)[t % 4]) for t in range(24)]


@pytest.mark.parametrize("s,seed,f", CASES)
def test_oracle_pivots_agree_with_superlu_until_tracking(orc, s, seed, f):
    r, c, v = _block(s, seed, f)
    o = orc.OracleSolver(r, c, v, s)
    assert o.analyze() == 0 and o.factor() == 0
    q, prow = o.sparse_order(0, s)
    lu = sla.splu(sp.csc_matrix((v, (r, c)), shape=(s, s)), permc_spec="COLAMD", diag_pivot_thresh=0.001,
                  options=dict(SymmetricMode=True, Equil=False))
    q_slu = np.empty(s, np.int64)
    q_slu[lu.perm_c] = np.arange(s)
    prow_slu = np.empty(s, np.int64)
    prow_slu[lu.perm_r] = np.arange(s)
    assert np.array_equal(q, q_slu)
    upto = _first_tracked(q, prow)
    assert np.array_equal(prow[:upto], prow_slu[:upto])


@pytest.mark.parametrize("s,seed,f", CASES)
def test_product_pivots_equal_oracle(bd, orc, monkeypatch, s, seed, f):
    monkeypatch.setenv("BLOCKDAGLU_GPU_TAIL_TOL", "0.001")
    r, c, v = _block(s, seed, f)
    h = bd.HostAnalysis(r, c, v, s)
    assert h.error == bd.SolverError.None_
    kinds, _ = h.block_kinds()
    assert list(kinds) == [1]
    q, prow = h.sparse_order(0)
    o = orc.OracleSolver(r, c, v, s)
    assert o.analyze() == 0 and o.factor() == 0
    qo, prowo = o.sparse_order(0, s)
    assert np.array_equal(q, qo)
    assert np.array_equal(prow, prowo)


def test_product_default_tail_partial_pivoting(bd, orc, monkeypatch):
    """Default analysis: KLU's pivot rows in the sparse part, partial pivoting
    (|L| <= 1) in the dense tail."""
    monkeypatch.delenv("BLOCKDAGLU_GPU_TAIL_TOL", raising=False)
    s, seed, f = CASES[2]
    r, c, v = _block(s, seed, f)
    h = bd.HostAnalysis(r, c, v, s)
    st = h.sparse_stats(0)
    q, prow = h.sparse_order(0)
    o = orc.OracleSolver(r, c, v, s)
    assert o.analyze() == 0 and o.factor() == 0
    qo, prowo = o.sparse_order(0, s)
    assert np.array_equal(q, qo) and np.array_equal(prow[: st["k0"]], prowo[: st["k0"]])
    assert st["m"] > 0


def test_scaled_block_pivots_match_oracle(bd, orc):
    """A block of 4,500 (KLU scale=1 row sums apply from 4,096 on) in a system
    whose rows and columns are scaled by 10^U(-4,4), as in the badly scaled
    benchmark system: the product's permutations, column order and sparse-part
    pivot rows equal the oracle's. (Both form each row sum in KLU's order,
    ascending local column; on the full-size scaled system a sum in another
    order flips a near-tie threshold decision, which the GPU fixture test of
    that system catches.)"""
    sp_ = bd.config_spec(3)
    sp_.n_small_blocks = 300
    sp_.n_small_total = 6000
    sp_.n_large_blocks = 1
    sp_.large_size_lo = sp_.large_size_hi = 4500
    n, r, c, v, _ = bd.generate(sp_)
    h = bd.HostAnalysis(r, c, v, n)
    kinds, sizes = h.block_kinds()
    big = [int(b) for b in np.nonzero(sizes >= 4096)[0]]
    assert len(big) == 1
    o = orc.OracleSolver(r, c, v, n)
    assert o.analyze() == 0 and o.factor() == 0
    P, Q, R = h.permutations()
    Po, Qo, Ro = o.permutations()
    assert np.array_equal(P, Po) and np.array_equal(Q, Qo) and np.array_equal(R, Ro)
    b = big[0]
    st = h.sparse_stats(b)
    assert st["scale"] == 1
    q, prow = h.sparse_order(b)
    qo, prowo = o.sparse_order(b, int(sizes[b]))
    assert np.array_equal(q, qo)
    k0 = st["k0"]
    assert np.array_equal(prow[:k0], prowo[:k0])
