"""GPU parity: the HIP path through the C-ABI against the oracle (CPU restatement
of the reference BLOCK route) and the reference's own known answers."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from conftest import backward_error, rel_residual

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_cfg1(bd, giant=2400, nsmall=600, seed=11):
    """The circuit-like workload (one power-law block + small blocks), scaled down."""
    sp = bd.config_spec(1)
    sp.seed = seed
    sp.n_small_blocks = nsmall
    sp.n_small_total = nsmall * 20
    sp.large_size_lo = sp.large_size_hi = giant
    return sp


def solve_gpu(bd, n, r, c, v, b, precision=0):
    s = bd.Solver()
    s.set_precision(bd.Precision(precision))
    assert s.set_matrix(r, c, v, n)
    assert s.analyze(), s.last_error_message()
    assert s.factor(), s.last_error_message()
    x = np.ascontiguousarray(b, np.float64).copy()
    assert s.solve_in_place(x), s.last_error_message()
    return s, x


def test_known_answer_3x3_with_duplicates(gpu):
    # reference apps/tests/main.cpp:111-128 (known answer at :117)
    _say("case: 3x3 known answer")
    bd = gpu
    coo = [(0, 0, 4.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 3.0), (1, 2, 1.0), (2, 1, 1.0), (2, 2, 2.0),
           (0, 0, 1.0), (0, 0, -1.0)]
    s = bd.Solver()
    assert s.set_matrix(coo, dimension=3)
    assert s.analyze() and s.factor(), s.last_error_message()
    x = np.array([1.0, 2.0, 3.0])
    assert s.solve_in_place(x)
    np.testing.assert_allclose(x, [0.2222222222222222, 0.1111111111111111, 1.4444444444444444], rtol=1e-9, atol=1e-9)
    bad = np.array([1.0, 2.0])
    assert not s.solve_in_place(bad)
    assert s.last_error() == bd.SolverError.InvalidRightHandSide
    assert s.selected_mode() == "BLOCK"


def cache_matrix(n, changed):
    # reference apps/tests/main.cpp:130-143
    m = []
    for row in range(n):
        m.append((row, row, 4.0))
        if row + 2 < n:
            m.append((row, changed if row == n // 2 else row + 1, 0.25))
        else:
            m.append((row, row, 0.5))
    return m


@pytest.mark.parametrize("changed,scale", [(551, 1.0), (552, 1.0), (551, 2.0)])
def test_cache_matrices_all_ones(gpu, changed, scale):
    # reference apps/tests/main.cpp:145-168: solution is all ones
    _say(f"case: cache matrix changed={changed} scale={scale}")
    bd = gpu
    n = 1100
    m = [(r, c, v * scale) for r, c, v in cache_matrix(n, changed)]
    r = np.array([t[0] for t in m]); c = np.array([t[1] for t in m]); v = np.array([t[2] for t in m])
    ones = np.ones(n)
    b = np.zeros(n)
    np.add.at(b, r, v * ones[c])
    s, x = solve_gpu(bd, n, r, c, v, b)
    np.testing.assert_allclose(x, ones, rtol=1e-9)


def test_tail_lu_kernels_match_numpy(gpu):
    """The no-pivot blocked LU used for dense Schur tails (MFMA trailing
    updates), against a float64 numpy LU without pivoting on a diagonally
    dominant matrix, in every launch structure: separate diagonal LU launches,
    the diagonal LU fused into each bulk update, and paired panels with K = 128
    bulk updates (with 128 tiles throughout, and with 64 tiles below 300 rows).
    The structures apply the same MFMA sequence to every element, so their
    factors must be identical bit for bit. m = 333 ends on an unpaired full
    panel and a 13-column one; m = 704 is 11 full panels."""
    bd = gpu
    L = bd.lib()
    ctx = C.c_void_p(L.bdlu_solver_device_context(L.bdlu_solver_create()))
    rng = np.random.default_rng(5)
    for m in (333, 704):
        A = rng.standard_normal((m, m)) + m * np.eye(m)
        ref = A.copy()
        for k in range(m - 1):
            ref[k + 1:, k] /= ref[k, k]
            ref[k + 1:, k + 1:] -= np.outer(ref[k + 1:, k], ref[k, k + 1:])
        for prec, tol in ((0, 1e-12), (1, 2e-5)):
            first = None
            for mode, small in ((0, 0), (1, 0), (2, 0), (2, 300)):
                _say(f"case: debug dense LU m={m} prec={prec} structure={mode} small_tile_below={small}")
                a = np.asfortranarray(A.copy())
                assert L.gpu.bdlu_gpu_debug_dense_lu(ctx, m, a.ctypes.data_as(C.POINTER(C.c_double)), prec, mode,
                                                     small) == 0
                err = np.max(np.abs(a - ref)) / np.max(np.abs(ref))
                _say(f"  max rel err={err:.3e}")
                assert err < tol
                if first is None:
                    first = a
                else:
                    assert np.array_equal(a, first), "launch structures differ"


def test_config0_fp64_matches_oracle(gpu, orc):
    _say("case: config 0 fp64 vs oracle")
    bd = gpu
    n, r, c, v, B = bd.generate(0)
    b = B[:, 0].copy()
    o = orc.OracleSolver(r, c, v)
    assert o.analyze() == 0 and o.factor() == 0
    rc, xo = o.solve(b)
    assert rc == 0
    s, x = solve_gpu(bd, n, r, c, v, b)
    P, Q, R = s.permutations()
    Po, Qo, Ro = o.permutations()
    assert np.array_equal(P, Po) and np.array_equal(Q, Qo) and np.array_equal(R, Ro)
    fe = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    rr = rel_residual(n, r, c, v, x, b)
    ro = rel_residual(n, r, c, v, xo, b)
    _say(f"  forward diff vs oracle={fe:.3e} rel_resid gpu={rr:.3e} oracle={ro:.3e}")
    # Plain FP64 (no refinement, the reference default): same accuracy class as
    # the oracle's KLU-rule factorization; both agree to rounding-level error.
    assert fe < 1e-9
    assert rr <= max(20 * ro, 1e-12)


def test_small_cfg1_fp64_and_refactor(gpu, orc):
    _say("case: scaled circuit-like workload, 3 refactors")
    bd = gpu
    n, r, c, v, B = bd.generate(small_cfg1(bd))
    b = B[:, 0].copy()
    s = bd.Solver()
    o = orc.OracleSolver(r, c, v)
    for it in range(3):
        vals = v if it == 0 else bd.perturb(v, seed=100 + it)
        assert s.set_matrix(r, c, vals, n)
        assert s.analyze() and s.factor(), s.last_error_message()
        x = b.copy()
        assert s.solve_in_place(x), s.last_error_message()
        rep = s.report()
        o = orc.OracleSolver(r, c, vals)
        assert o.analyze() == 0 and o.factor() == 0
        _, xo = o.solve(b)
        fe = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        rr = rel_residual(n, r, c, vals, x, b)
        ro = rel_residual(n, r, c, vals, xo, b)
        _say(f"  it={it} plan_reused={rep['plan_reused']} fe={fe:.3e} rr={rr:.3e} oracle_rr={ro:.3e} "
             f"max_block={rep['max_block']}")
        assert fe < 1e-7
        assert rr <= max(20 * ro, 1e-12)
        if it > 0:
            assert rep["plan_reused"] == 1


def test_fp64_refinement_reaches_1e12(gpu, monkeypatch):
    # REFINE_ITERS (reference src/solver.cpp:1461-1464): FP64 refinement with the
    # reference stop rule ||r||_2/||b||_2 < 1e-12 (src/solver.cpp:1508-1521).
    _say("case: FP64 + refinement")
    bd = gpu
    monkeypatch.setenv("USERALGO_REFINE_ITERS", "4")
    n, r, c, v, B = bd.generate(small_cfg1(bd))
    b = B[:, 0].copy()
    s, x = solve_gpu(bd, n, r, c, v, b)
    rep = s.report()
    rr = rel_residual(n, r, c, v, x, b)
    _say(f"  iters={rep['ir_iterations']} rr={rr:.3e} engine_rr={rep['rel_residual']:.3e}")
    assert rr < 1e-12
    assert 1 <= rep["ir_iterations"] <= 4


def test_fp32_ir_reaches_1e12(gpu):
    _say("case: FP32+IR")
    bd = gpu
    for spec in (0, small_cfg1(bd)):
        n, r, c, v, B = bd.generate(spec)
        b = B[:, 0].copy()
        s, x = solve_gpu(bd, n, r, c, v, b, precision=1)
        rep = s.report()
        rr = rel_residual(n, r, c, v, x, b)
        be = backward_error(n, r, c, v, x, b)
        _say(f"  iters={rep['ir_iterations']} rr={rr:.3e} berr={be:.3e} engine_rr={rep['rel_residual']:.3e} "
             f"fallbacks={rep['fp64_fallback_blocks']}")
        assert rr <= 1e-12
        assert be <= 1e-12


@pytest.mark.parametrize("precision", [0, 1])
def test_multi_rhs_matches_single(gpu, precision):
    """k right-hand sides solved together (one pass of the K-column kernels) give
    what k single solves give: bit for bit in plain FP64; with FP32+IR the
    columns refine together, so each meets the residual bound."""
    bd = gpu
    _say(f"case: multi-RHS precision={precision}")
    n, r, c, v, _ = bd.generate(small_cfg1(bd))
    rng = np.random.default_rng(7)
    B = np.asfortranarray(rng.uniform(-1, 1, (n, 5)))
    s = bd.Solver()
    s.set_precision(bd.Precision(precision))
    assert s.set_matrix(r, c, v, n) and s.analyze() and s.factor(), s.last_error_message()
    X = B.copy(order="F")
    assert s.solve_multi_in_place(X, 5), s.last_error_message()
    for j in range(5):
        x = B[:, j].copy()
        assert s.solve_in_place(x)
        rr = rel_residual(n, r, c, v, X[:, j], B[:, j])
        if precision == 0:
            assert np.array_equal(x, X[:, j]), f"column {j} differs"
        else:
            assert rr <= 1e-12, rr


def test_long_rows_residual_and_refinement(gpu):
    """Rows with more than 64 entries (dense-ish blocks) take the wavefront-per-row
    residual. FP32+IR on such a matrix, 4 right-hand sides together (one
    wavefront per column in the block solves, one workgroup per band and column
    in the tail solves), must reach the residual bound measured here, and the
    engine's own residual must agree with it."""
    bd = gpu
    _say("case: long rows, FP32+IR, 4 RHS")
    sp = bd.config_spec(2)
    sp.n_small_blocks = 300
    sp.n_large_blocks = 2
    sp.large_size_lo, sp.large_size_hi = 256, 320
    sp.keep_order = 1
    n, r, c, v, B = bd.generate(sp)
    assert np.bincount(r, minlength=n).max() > 64
    s = bd.Solver()
    s.set_precision(bd.Precision(1))
    assert s.set_matrix(r, c, v, n) and s.analyze() and s.factor(), s.last_error_message()
    X = np.asfortranarray(B[:, :4].copy())
    assert s.solve_multi_in_place(X, 4), s.last_error_message()
    rep = s.report()
    rrs = [rel_residual(n, r, c, v, X[:, j], B[:, j]) for j in range(4)]
    _say(f"  n={n} iters={rep['ir_iterations']} host rr={max(rrs):.3e} engine rr={rep['rel_residual']:.3e}")
    assert max(rrs) <= 1e-12
    assert rep["rel_residual"] <= 1e-12
