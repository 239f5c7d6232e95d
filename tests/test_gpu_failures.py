"""Failure handling of the device factorization, on inputs built on the CPU so
that every kernel still exits normally:
  * an exactly singular dense block fails like the reference's dense branch
    (zero pivot -> FactorizationFailed, ubtf_block_solver.cpp:1475-1477 and
    src/solver.cpp:1259-1262), in FP64 and in FP32, and the oracle agrees;
  * a dense block whose FP32 pivot loses more than half its bits to
    cancellation is refactored in FP64 (exactly one fallback) and FP32+IR still
    reaches ||b - A x|| / ||b|| < 1e-12;
  * new values on a cached pattern that make the fixed pivot order of a sparse
    block grow |L| beyond 1e8 (FP64) are re-analysed (REPIVOT) and solved to
    rounding level."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import rel_residual

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _with_singular_2x2(n_id=8):
    """Unit diagonal of size n_id followed by the strongly connected 2x2 block
    [[1, 2], [2, 4]] (rank 1). With partial pivoting, column 0 pivots on the
    second row (|2| > |1|), l = 1/2, and the second pivot is 2 - 0.5 * 4 = 0
    exactly, in FP64 and in FP32."""
    r = list(range(n_id)) + [n_id, n_id, n_id + 1, n_id + 1]
    c = list(range(n_id)) + [n_id, n_id + 1, n_id, n_id + 1]
    v = [1.0] * n_id + [1.0, 2.0, 2.0, 4.0]
    return n_id + 2, np.array(r, np.int32), np.array(c, np.int32), np.array(v)


@pytest.mark.parametrize("precision", [0, 1])
def test_singular_dense_block_fails_factor(gpu, orc, precision):
    bd = gpu
    _say(f"case: exactly singular dense block, precision {precision}")
    n, r, c, v = _with_singular_2x2()
    s = bd.Solver()
    s.set_precision(bd.Precision(precision))
    assert s.set_matrix(r, c, v, n)
    assert s.analyze(), s.last_error_message()
    assert not s.factor()
    assert s.last_error() == bd.SolverError.FactorizationFailed
    assert "block factorization failed" in s.last_error_message()
    x = np.ones(n)
    assert not s.solve_in_place(x) and s.last_error() == bd.SolverError.NotPrepared
    o = orc.OracleSolver(r, c, v, n)
    assert o.analyze() == 0 and o.factor() != 0
    s.close()


def test_fp32_small_pivot_falls_back_to_fp64(gpu):
    """Block [[1, 1], [1, 1 + 1e-5]]: its second pivot 1e-5 is computed as
    (1 + 1e-5) - 1, which keeps only ~7 of FP32's 24 bits, so the scale-invariant
    check (pivot < 2^-12 (|L||U|)_kk) sends that block to FP64."""
    bd = gpu
    _say("case: FP32 small-pivot check -> FP64 fallback of one dense block")
    rng = np.random.default_rng(3)
    m = 30  # well-conditioned diagonal part
    r = list(range(m)) + [m, m, m + 1, m + 1]
    c = list(range(m)) + [m, m + 1, m, m + 1]
    v = list(rng.uniform(1.0, 2.0, m)) + [1.0, 1.0, 1.0, 1.0 + 1e-5]
    n = m + 2
    r, c, v = np.array(r, np.int32), np.array(c, np.int32), np.array(v)
    # b = A x with x = O(1): the block is near-singular (condition ~4e5), and a
    # random b would give |x| ~ 1e5, where FP64 cannot resolve ||b - A x|| / ||b||
    # below ~1e-11 whatever the solver.
    b = _dense(n, r, c, v) @ rng.uniform(-1, 1, n)
    s = bd.Solver()
    s.set_precision(bd.Precision.FP32_IR)
    assert s.set_matrix(r, c, v, n)
    assert s.analyze() and s.factor(), s.last_error_message()
    rep = s.report()
    x = b.copy()
    assert s.solve_in_place(x), s.last_error_message()
    rr = rel_residual(n, r, c, v, x, b)
    _say(f"  fp64 fallbacks={rep['fp64_fallback_blocks']} rel_residual={rr:.2e} "
         f"iterations={s.report()['ir_iterations']}")
    assert rep["fp64_fallback_blocks"] == 1
    assert rr < 1e-12
    s.close()


def _sparse_block(s=120, seed=5):
    """One strongly connected block of size s > 64 (a KLU block in the
    reference): a cycle, ~4 random entries per row, strong diagonal."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(s):
        rows += [i, i]
        cols += [i, (i + 1) % s]
        vals += [10.0, rng.normal()]
        for j in rng.choice(s, 4, replace=False):
            if j != i and j != (i + 1) % s:
                rows.append(i)
                cols.append(int(j))
                vals.append(rng.normal())
    return s, np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals)


def test_pivot_growth_forces_reanalysis(gpu):
    bd = gpu
    _say("case: FP64 pivot growth on a cached pattern -> re-analysis")
    n, r, c, v0 = _sparse_block()
    rng = np.random.default_rng(8)
    b = rng.uniform(-1, 1, n)
    s = bd.Solver()
    assert s.set_matrix(r, c, v0, n) and s.analyze() and s.factor(), s.last_error_message()
    assert s.report()["sparse_blocks"] == 1
    # Same pattern, diagonal shrunk to 1e-12: on the first analysis' pivot order
    # (diagonal pivots) |L| would reach ~1e12.
    v1 = np.where(r == c, 1e-12, v0)
    assert s.set_matrix(r, c, v1, n)
    assert s.analyze(), s.last_error_message()
    assert s.report()["plan_reused"] == 1
    assert s.factor(), s.last_error_message()
    rep = s.report()
    x = b.copy()
    assert s.solve_in_place(x), s.last_error_message()
    rr = rel_residual(n, r, c, v1, x, b)
    xr = np.linalg.solve(_dense(n, r, c, v1), b)
    fwd = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    _say(f"  repivot blocks={rep['repivot_blocks']} rel_residual={rr:.2e} forward error vs LAPACK={fwd:.2e}")
    assert rep["repivot_blocks"] == 1
    assert rr < 1e-12 and fwd < 1e-9
    s.close()


def _dense(n, r, c, v):
    A = np.zeros((n, n))
    np.add.at(A, (r, c), v)
    return A
