"""The no-allocation solve (bdlu_gpu_solve_noalloc, include/bdlu_gpu.h) against
bdlu_gpu_solve on the same resident factors, through the C-ABI. Both run the
same kernels with the same arithmetic, so the bar is bit-for-bit equality of
the solutions (integer-exact comparison of the float64 bits) and equal
refinement counts. Two calls in a row share the library's scratch array: the
second is a smaller system with fewer right-hand sides, so stale contents of
the first call would show."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BDLU_MEM_DEVICE = 1


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _factored(bd, spec, precision):
    n, r, c, v, _ = bd.generate(spec)
    s = bd.Solver()
    s.set_precision(bd.Precision(precision))
    assert s.set_matrix(r, c, v, n)
    assert s.analyze(), s.last_error_message()
    assert s.factor(), s.last_error_message()
    return s, n


def _both(bd, s, n, k, ir, seed):
    L = bd.lib()
    rng = np.random.default_rng(seed)
    B = rng.uniform(-1, 1, n * k)  # n x k column-major
    ctx = C.c_void_p(s.device_context())
    plan = s.device_plan()
    assert ctx.value and plan
    d_ref = bd.DeviceArray.from_numpy(B)
    d_new = bd.DeviceArray.from_numpy(B)
    st_ref, st_new = bd.SolveStats(), bd.SolveStats()
    rc = L.gpu.bdlu_gpu_solve(ctx, plan, d_ref.address, k, n, BDLU_MEM_DEVICE, C.byref(ir), C.byref(st_ref))
    assert rc == 0, L.gpu.bdlu_gpu_last_error(ctx)
    rc = L.gpu.bdlu_gpu_solve_noalloc(ctx, plan, d_new.address, k, n, C.byref(ir), C.byref(st_new))
    assert rc == 0, L.gpu.bdlu_gpu_last_error(ctx)
    x_ref, x_new = d_ref.numpy(), d_new.numpy()
    same = np.array_equal(x_ref.view(np.uint64), x_new.view(np.uint64))
    _say(f"  n={n} k={k} iterations {st_ref.iterations}/{st_new.iterations} "
         f"rel_residual {st_ref.rel_residual:.2e}/{st_new.rel_residual:.2e} bitwise_equal={same}")
    assert same, "no-allocation solve differs from bdlu_gpu_solve"
    assert st_ref.iterations == st_new.iterations
    assert st_ref.rel_residual == st_new.rel_residual
    d_ref.free()
    d_new.free()
    return st_new


def test_noalloc_solve_equals_solve_on_two_calls(gpu):
    bd = gpu
    # First call: circuit-like system (one 2,400 power-law block + 600 small
    # blocks), FP32 factors with FP64 refinement, 3 right-hand sides.
    _say("case: no-allocation solve, call 1 (circuit-like, FP32+IR, 3 RHS)")
    sp = bd.config_spec(1)
    sp.seed, sp.n_small_blocks, sp.n_small_total = 11, 600, 12000
    sp.large_size_lo = sp.large_size_hi = 2400
    s1, n1 = _factored(bd, sp, precision=1)
    ir = bd.IrOpts(30, 0, 1e-12, 0)
    st = _both(bd, s1, n1, 3, ir, seed=1)
    assert st.rel_residual < 1e-12
    # Second call: the smaller 5,000-unknown system (200 hidden blocks), plain FP64, 1 RHS.
    _say("case: no-allocation solve, call 2 (n=5000, plain FP64, 1 RHS)")
    s2, n2 = _factored(bd, bd.config_spec(0), precision=0)
    assert n2 < n1
    _both(bd, s2, n2, 1, bd.IrOpts(0, -1, 1e-12, 0), seed=2)
    s1.close()
    s2.close()
