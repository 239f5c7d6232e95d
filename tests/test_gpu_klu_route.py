"""The global-KLU route on the GPU (reference src/solver.cpp:878-968 analyze,
:1190-1242 factor, :1264-1316 AUTO fallback, :1472-1479 solve), through the
C-ABI, against the oracle's restatement (oracle analyze_klu):
  * AUTO picks the global-KLU route for the 5,000-unknown system (n < 50,000);
    BTF permutations composed with the small blocks' COLAMD orders, block
    bounds, the row pivots of every block of 2..64 (the device's KLU-rule
    dense kernel) and the column order and sparse-part pivot rows of every
    larger block are identical to the oracle's; FP64 and FP32 factors with FP64
    refinement reach ||b - A x||_2 / ||b||_2 < 1e-12 and agree with the oracle's
    refined solution to 1e-9 (max-norm, relative);
  * KLU's whole-row max scaling (scale = 2) on a system with rows and columns
    scaled by 10^U(-3,3) (componentwise backward error <= 1e-12 there, since
    the 2-norm residual is dominated by rounding in the largest rows), and btf = 0
    (one block for the whole matrix), each against the oracle the same way;
  * AUTO's fallback: a system of 50,000 unknowns (AUTO takes the BLOCK route)
    whose 2x2 block [[1, 1], [3, 3 + 2^-51]] loses its second pivot to exact
    cancellation under partial pivoting fails on the BLOCK route
    (FactorizationFailed) and, in AUTO mode, is refactored on the global-KLU
    route, where KLU's diagonal preference keeps the pivot 2^-51, and solved to
    its exact solution."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import rel_residual


def componentwise_error(n, r, c, v, x, b):
    """max_i |b - A x|_i / (|A| |x| + |b|)_i."""
    y = np.zeros(n)
    np.add.at(y, r, v * x[c])
    den = np.abs(b).copy()
    np.add.at(den, r, np.abs(v * x[c]))
    return float(np.max(np.abs(b - y) / den))

pytestmark = [pytest.mark.gpu, pytest.mark.any_route]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(bd, orc, s, o, n, r, c, v, b, scaled=False):
    P, Q, R = s.permutations()
    Po, Qo, Ro = o.klu_permutations()
    assert np.array_equal(R, Ro) and np.array_equal(P, Po) and np.array_equal(Q, Qo)
    Pb, _, _ = o.permutations()
    sizes = np.diff(R)
    s.set_precision(bd.Precision.FP64)
    assert s.factor(), s.last_error_message()
    piv = s.dense_pivots()
    for blk in range(len(sizes)):
        lo, sz = int(R[blk]), int(sizes[blk])
        qo, prowo = o.sparse_order(blk, sz)
        if sz <= 64:
            # Original row pivoted at each position: device (composed order,
            # gather pivots) against the oracle (BTF order, KLU pivot rows).
            dev_rows = P[lo + piv[lo:lo + sz]]
            ora_rows = Pb[lo + prowo]
            assert np.array_equal(dev_rows, ora_rows), f"block {blk} (size {sz}): pivot rows"
        else:
            q, prow, k0 = s.sparse_order(blk, sz)
            assert np.array_equal(q, qo), f"block {blk}: column order"
            assert np.array_equal(prow[:k0], prowo[:k0]), f"block {blk}: sparse-part pivot rows"
    rc, xo = o.solve(b, 10)
    assert rc == 0
    for prec in (bd.Precision.FP64, bd.Precision.FP32_IR):
        s.set_precision(prec)
        assert s.factor(), s.last_error_message()
        x = b.copy()
        assert s.solve_in_place(x), s.last_error_message()
        rr = rel_residual(n, r, c, v, x, b)
        cw = componentwise_error(n, r, c, v, x, b)
        fwd = float(np.max(np.abs(x - xo)) / np.max(np.abs(xo)))
        _say(f"  {prec.name}: rel_residual {rr:.2e} componentwise {cw:.2e} forward error vs oracle {fwd:.2e} "
             f"iterations {s.report()['ir_iterations']}")
        if scaled:
            assert cw <= 1e-12
        else:
            assert rr < 1e-12
        assert fwd < 1e-9


def test_auto_routes_small_system_to_global_klu(gpu, orc, monkeypatch):
    bd = gpu
    for k in ("USERALGO_MODE", "USERALGO_KLU_BTF", "USERALGO_KLU_SCALE", "USERALGO_KLU_TOL"):
        monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv("USERALGO_REFINE_ITERS", "10")
    n, r, c, v, B = bd.generate(0)
    b = B[:, 0].copy()
    _say(f"case: AUTO on the 5,000-unknown system (n={n})")
    s = bd.Solver()
    assert s.selected_mode() == "BLOCK"  # before analyze, as the reference's default
    assert s.set_matrix(r, c, v, n) and s.analyze(), s.last_error_message()
    assert s.selected_mode() == "KLU"
    o = orc.OracleSolver(r, c, v, n)
    assert o.analyze_klu(1, 0, 0.001) == 0 and o.factor() == 0
    _compare(bd, orc, s, o, n, r, c, v, b)
    s.close()


def test_global_klu_max_scaling(gpu, orc, monkeypatch):
    bd = gpu
    monkeypatch.setenv("USERALGO_MODE", "klu")
    monkeypatch.setenv("USERALGO_KLU_SCALE", "2")
    monkeypatch.delenv("USERALGO_KLU_BTF", raising=False)
    monkeypatch.setenv("USERALGO_REFINE_ITERS", "10")
    n, r, c, v, B = bd.generate(0)
    rng = np.random.default_rng(3)
    dr, dc = 10.0 ** rng.uniform(-3, 3, n), 10.0 ** rng.uniform(-3, 3, n)
    v = v * dr[r] * dc[c]  # rows and columns scaled by 10^U(-3,3)
    b = B[:, 0].copy()
    _say(f"case: global KLU, scale=2, badly scaled values (n={n})")
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n) and s.analyze(), s.last_error_message()
    assert s.selected_mode() == "KLU"
    o = orc.OracleSolver(r, c, v, n)
    assert o.analyze_klu(1, 2, 0.001) == 0 and o.factor() == 0
    monkeypatch.setenv("BLOCKDAGLU_GPU_IR_STOP", "componentwise")
    _compare(bd, orc, s, o, n, r, c, v, b, scaled=True)
    s.close()


def test_global_klu_one_block(gpu, orc, monkeypatch):
    bd = gpu
    monkeypatch.setenv("USERALGO_MODE", "klu")
    monkeypatch.setenv("USERALGO_KLU_BTF", "0")
    monkeypatch.delenv("USERALGO_KLU_SCALE", raising=False)
    monkeypatch.setenv("USERALGO_REFINE_ITERS", "10")
    n, r, c, v, B = bd.generate(bd.config_spec(4, 11))
    b = B[:, 0].copy()
    _say(f"case: global KLU, btf=0 (n={n})")
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n) and s.analyze(), s.last_error_message()
    o = orc.OracleSolver(r, c, v, n)
    assert o.analyze_klu(0, 0, 0.001) == 0 and o.factor() == 0
    _compare(bd, orc, s, o, n, r, c, v, b)
    s.close()


def _cancelling_system(n=50000, at=20000):
    """Identity of size n except rows/cols (at, at+1) = [[1, 1], [3, 3 + 2^-51]].
    Partial pivoting takes row at+1 (|3| > |1|), l = fl(1/3), and
    1 - fl(l * (3 + 2^-51)) = 0 exactly in FP64; KLU keeps the diagonal 1 and
    gets the pivot (3 + 2^-51) - 3 = 2^-51. b = (1, ..., 2, 6, ..., 1) has the
    exact solution x = (1, ..., 2, 0, ..., 1)."""
    r = list(range(n))
    c = list(range(n))
    v = [1.0] * n
    v[at + 1] = 3.0 + 2.0 ** -51
    r += [at, at + 1]
    c += [at + 1, at]
    v += [1.0, 3.0]
    b = np.ones(n)
    b[at], b[at + 1] = 2.0, 6.0
    x = np.ones(n)
    x[at], x[at + 1] = 2.0, 0.0
    return n, np.array(r, np.int32), np.array(c, np.int32), np.array(v), b, x


def test_auto_falls_back_to_global_klu(gpu, monkeypatch):
    bd = gpu
    n, r, c, v, b, xt = _cancelling_system()
    _say(f"case: AUTO fallback, n={n}, cancelling 2x2 block")
    monkeypatch.setenv("USERALGO_MODE", "block")
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n) and s.analyze(), s.last_error_message()
    assert s.selected_mode() == "BLOCK"
    assert not s.factor()
    assert s.last_error() == bd.SolverError.FactorizationFailed
    s.close()
    monkeypatch.delenv("USERALGO_MODE")
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n) and s.analyze(), s.last_error_message()
    assert s.selected_mode() == "BLOCK"
    assert s.factor(), s.last_error_message()
    assert s.selected_mode() == "KLU"
    x = b.copy()
    assert s.solve_in_place(x), s.last_error_message()
    assert np.array_equal(x, xt)
    s.close()
