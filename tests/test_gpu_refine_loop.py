"""Iterative refinement run on the device against the host-driven loop, through
the C-ABI.

bdlu_gpu_solve on a plan's own work set (graphs on, the default) runs the
solve, the first residual and the refinement steps as one captured graph whose
stop decisions are taken on the device (k_ir_check), with as many steps as the
plan's last solve took; a loop still running at the graph's end continues on
the host. bdlu_gpu_solve_noalloc runs the same kernels with every stop decision
taken on the host after a copy of the norms. Same factors, same right-hand
sides, same options: the solutions must agree bit for bit and the iteration
counts and reported measures exactly, whether the graph holds exactly the steps
needed, more (the extra steps are gated off), or fewer (the host takes over),
for each stop rule (0: ||r||_2/||b||_2 < tol, 1: normwise backward error
<= tol, 2: componentwise backward error <= tol), a zero right-hand side among
others, and an iteration cap that ends the loop before the rule is met.

The stall rule (a step that fails to halve the worst measure ends a loop on
FP32 factors, which are then refactored in FP64) is checked against its
definition: the residual after m steps, for m = 0, 1, ..., read from host-loop
solves capped at m steps, fixes the step j at which the FP32 loop must stop,
so a fallback solve with an unattainable tolerance must report j plus the cap
of the FP64 loop that follows, and the FP64 refactor."""
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


@pytest.fixture(scope="module")
def fp32_system(gpu):
    bd = gpu
    # Circuit-like: 300 small hidden blocks and one 1,800-unknown power-law block.
    sp = bd.config_spec(1)
    sp.seed, sp.n_small_blocks, sp.n_small_total = 23, 300, 6000
    sp.large_size_lo = sp.large_size_hi = 1800
    n, r, c, v, _ = bd.generate(sp)
    s = bd.Solver()
    s.set_precision(bd.Precision.FP32_IR)
    assert s.set_matrix(r, c, v, n) and s.analyze(), s.last_error_message()
    assert s.factor(), s.last_error_message()
    yield bd, s, n
    s.close()


def _solve(bd, s, n, B, k, ir, noalloc):
    L = bd.lib()
    ctx = C.c_void_p(s.device_context())
    d = bd.DeviceArray.from_numpy(B)
    st = bd.SolveStats()
    if noalloc:
        rc = L.gpu.bdlu_gpu_solve_noalloc(ctx, s.device_plan(), d.address, k, n, C.byref(ir), C.byref(st))
    else:
        rc = L.gpu.bdlu_gpu_solve(ctx, s.device_plan(), d.address, k, n, BDLU_MEM_DEVICE, C.byref(ir), C.byref(st))
    assert rc == 0, L.gpu.bdlu_gpu_last_error(ctx)
    x = d.numpy()
    d.free()
    return x, st


CASES = [
    # (label, right-hand sides, IrOpts(max_iters, stop_rule, tol, fp64_fallback), zero column)
    ("rule 0, 3 RHS, first call", 3, (30, 0, 1e-12, 0), None),
    ("rule 0, 3 RHS, same again", 3, (30, 0, 1e-12, 0), None),
    ("rule 0, loose tol (graph holds extra steps)", 3, (30, 0, 1e-7, 0), None),
    ("rule 0, tight tol (host continues)", 2, (30, 0, 1e-15, 0), None),
    ("rule 1, 8 RHS with a zero column", 8, (30, 1, 1e-15, 0), 5),
    ("rule 2, 1 RHS", 1, (30, 2, 1e-13, 0), None),
    ("cap of 2 steps, unattainable tol", 4, (2, 0, 1e-30, 0), None),
    ("cap of 2 steps, same again", 4, (2, 0, 1e-30, 0), None),
]


@pytest.mark.parametrize("label,k,opts,zero", CASES, ids=[c[0] for c in CASES])
def test_device_loop_equals_host_loop(fp32_system, label, k, opts, zero):
    bd, s, n = fp32_system
    _say(f"case: device refinement loop vs host loop: {label} (n={n}, k={k}, opts={opts})")
    rng = np.random.default_rng(sum(map(ord, label)))
    B = rng.uniform(-1, 1, n * k)
    if zero is not None:
        B[zero * n:(zero + 1) * n] = 0.0
    ir = bd.IrOpts(*opts)
    x_dev, st_dev = _solve(bd, s, n, B, k, ir, noalloc=False)
    x_host, st_host = _solve(bd, s, n, B, k, ir, noalloc=True)
    same = np.array_equal(x_dev.view(np.uint64), x_host.view(np.uint64))
    _say(f"  iterations {st_dev.iterations}/{st_host.iterations} rel {st_dev.rel_residual:.3e}/"
         f"{st_host.rel_residual:.3e} berr {st_dev.backward_error:.3e} cw {st_dev.componentwise_error:.3e} "
         f"bitwise_equal={same}")
    assert same, "device-side refinement differs from the host loop"
    assert st_dev.iterations == st_host.iterations
    assert st_dev.rel_residual == st_host.rel_residual
    assert st_dev.backward_error == st_host.backward_error
    assert st_dev.componentwise_error == st_host.componentwise_error
    if opts[2] >= 1e-13:
        assert st_dev.iterations < opts[0]  # the rule was met, not the cap
    if opts[0] == 2:
        assert st_dev.iterations == 2
    if zero is not None:
        assert np.all(x_dev[zero * n:(zero + 1) * n] == 0.0)


def test_device_loop_stall_rule(fp32_system):
    bd, s, n = fp32_system
    _say(f"case: stall rule of the device loop, FP32 factors then FP64 fallback (n={n})")
    B = np.random.default_rng(77).uniform(-1, 1, n)
    cap = 12
    # Residual after m host-loop steps (rule 0 can never be met at 1e-30).
    res = []
    for m in range(cap + 1):
        opts = bd.IrOpts(m, 0, 1e-30, 0) if m else bd.IrOpts(0, -1, 1e-30, 0)
        _, st = _solve(bd, s, n, B, 1, opts, noalloc=True)
        assert st.iterations == m
        res.append(st.rel_residual)
    j = next((m for m in range(1, cap + 1) if not (res[m] <= 0.5 * res[m - 1])), None)
    _say(f"  residual by step: {' '.join(f'{v:.2e}' for v in res)}; stall at step {j}")
    assert j is not None and j < cap
    x_dev, st_dev = _solve(bd, s, n, B, 1, bd.IrOpts(cap, 0, 1e-30, 1), noalloc=False)
    _say(f"  fallback solve: iterations {st_dev.iterations} fp64_refactor {st_dev.fp64_refactor}")
    assert st_dev.fp64_refactor == 1
    assert st_dev.iterations == j + cap
    # The plan now holds FP64 factors: the device loop and the host loop agree on them too.
    ir = bd.IrOpts(cap, 0, 1e-30, 0)
    x_dev, st_dev = _solve(bd, s, n, B, 1, ir, noalloc=False)
    x_host, st_host = _solve(bd, s, n, B, 1, ir, noalloc=True)
    assert np.array_equal(x_dev.view(np.uint64), x_host.view(np.uint64))
    assert st_dev.iterations == st_host.iterations == cap
