"""Serving loop on the GPU (include/blockdaglu/serve.h): requests on registered
patterns run concurrently in slots; each result must equal the same request run
alone (tolerance: bitwise, since every kernel is deterministic and a request
never shares buffers), and meet the accuracy of its precision."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import rel_residual

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _patterns(bd, count=3):
    out = []
    for p in range(count):
        sp = bd.config_spec(1)
        sp.seed = 500 + p
        sp.n_small_blocks = 60 + 20 * p
        sp.n_small_total = sp.n_small_blocks * 20
        sp.large_size_lo = sp.large_size_hi = 300 + 250 * p
        n, r, c, v, _ = bd.generate(sp)
        out.append((n, r, c, v))
    return out


def _requests(bd, pats, count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        p = int(rng.integers(len(pats)))
        n, r, c, v = pats[p]
        vals = bd.perturb(v, seed=10_000 * seed + i)
        k = int(rng.integers(1, 5))
        B = np.asfortranarray(rng.uniform(-1, 1, (n, k)))
        reqs.append((p, vals, B))
    return reqs


@pytest.mark.parametrize("precision", [0, 1])
def test_serve_alone_equals_mix(gpu, precision):
    bd = gpu
    prec = bd.Precision(precision)
    _say(f"case: serve alone vs mix precision={prec.name}")
    pats = _patterns(bd)
    reqs = _requests(bd, pats, 24, seed=3 + precision)
    mix = bd.Server(slots=4, hbm_budget_gb=8.0, precision=prec)
    pid = [mix.add_pattern(r, c, v, n) for (n, r, c, v) in pats]
    ids = [mix.submit(pid[p], vals, B) for (p, vals, B) in reqs]
    mixed = [mix.wait(i) for i in ids]
    mix.close()
    slots = {res["slot"] for _, res in mixed}
    spans = sorted((res["t_start_ms"], res["t_done_ms"]) for _, res in mixed)
    overlapped = any(b[0] < a[1] for a, b in zip(spans[:-1], spans[1:]))
    _say(f"  slots used={sorted(slots)} overlapped={overlapped}")
    assert len(slots) > 1 and overlapped

    alone = bd.Server(slots=1, hbm_budget_gb=8.0, precision=prec)
    pid1 = [alone.add_pattern(r, c, v, n) for (n, r, c, v) in pats]
    worst_rr = 0.0
    for q, ((p, vals, B), (Xm, rm)) in enumerate(zip(reqs, mixed)):
        Xa, ra = alone.wait(alone.submit(pid1[p], vals, B))
        assert rm["error"] == bd.SolverError.None_ and ra["error"] == bd.SolverError.None_
        assert np.array_equal(Xa, Xm), "result differs between alone and in the mix"
        n, r, c, _ = pats[p]
        rr = max(rel_residual(n, r, c, vals, Xm[:, j], B[:, j]) for j in range(B.shape[1]))
        worst_rr = max(worst_rr, rr)
        if q < 3:
            _say(f"  request {q}: engine rel_residual {rm['rel_residual']:.3e}, recomputed {rr:.3e}")
        # The engine's FP64 residual of the request agrees with a recomputation
        # (different summation order: agreement within a factor of 10 above 1e-16).
        assert rm["rel_residual"] <= 10 * rr + 1e-16 and rr <= 10 * rm["rel_residual"] + 1e-16
        assert rm["ir_iterations"] == ra["ir_iterations"]
    alone.close()
    _say(f"  worst rel residual {worst_rr:.2e}")
    if precision == 1:
        assert worst_rr < 1e-12
    else:
        assert worst_rr < 1e-9


def test_serve_budget_eviction(gpu):
    """A budget that holds one plan per slot still serves every pattern (plans are
    evicted and re-uploaded), with the same results."""
    bd = gpu
    _say("case: serve with a tight HBM budget")
    pats = _patterns(bd)
    reqs = _requests(bd, pats, 9, seed=11)
    big = bd.Server(slots=1, hbm_budget_gb=8.0)
    pb = [big.add_pattern(r, c, v, n) for (n, r, c, v) in pats]
    ref = [big.wait(big.submit(pb[p], vals, B))[0] for (p, vals, B) in reqs]
    big.close()
    tight = bd.Server(slots=1, hbm_budget_gb=0.02)
    pt = [tight.add_pattern(r, c, v, n) for (n, r, c, v) in pats]
    for (p, vals, B), x in zip(reqs, ref):
        X, res = tight.wait(tight.submit(pt[p], vals, B))
        assert res["error"] == bd.SolverError.None_, res
        assert np.array_equal(X, x)
    tight.close()


def test_graph_replays_match_eager(gpu, monkeypatch):
    """Refactor and solve launch sequences captured in HIP graphs and replayed
    (several replays, with other requests and host synchronisation between
    them) give exactly what the same launches issued directly give."""
    bd = gpu
    _say("case: graph replays vs eager launches")
    pats = _patterns(bd, count=2)
    reqs = _requests(bd, pats, 8, seed=21)
    outs = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BLOCKDAGLU_GPU_GRAPH", mode)
        for prec in (bd.Precision.FP64, bd.Precision.FP32_IR):
            srv = bd.Server(slots=1, hbm_budget_gb=8.0, precision=prec)
            pid = [srv.add_pattern(r, c, v, n) for (n, r, c, v) in pats]
            outs[(mode, prec)] = [srv.wait(srv.submit(pid[p], vals, B)) for (p, vals, B) in reqs]
            srv.close()
    for prec in (bd.Precision.FP64, bd.Precision.FP32_IR):
        for (xg, rg), (xe, re) in zip(outs[("1", prec)], outs[("0", prec)]):
            assert rg["error"] == bd.SolverError.None_ and re["error"] == bd.SolverError.None_
            assert np.array_equal(xg, xe), "graph replay differs from eager launches"
            assert rg["ir_iterations"] == re["ir_iterations"]


def test_precision_switch_on_a_cached_plan(gpu):
    """A plan factored in FP64 and then reused by an FP32+IR solver must solve with
    its FP32 factors: same result and iteration count as a fresh context."""
    bd = gpu
    _say("case: precision switch on a cached plan")
    (n, r, c, v), = _patterns(bd, count=1)
    b = np.random.default_rng(5).uniform(-1, 1, n)
    s64 = bd.Solver()
    assert s64.set_matrix(r, c, v, n) and s64.analyze() and s64.factor()
    x64 = b.copy()
    assert s64.solve_in_place(x64)
    s64.close()
    s32 = bd.Solver()
    s32.set_precision(bd.Precision.FP32_IR)
    assert s32.set_matrix(r, c, v, n) and s32.analyze() and s32.factor()
    x32 = b.copy()
    assert s32.solve_in_place(x32)
    rep = s32.report()
    s32.close()
    srv = bd.Server(slots=1, hbm_budget_gb=8.0, precision=bd.Precision.FP32_IR)
    X, res = srv.wait(srv.submit(srv.add_pattern(r, c, v, n), v, b))
    srv.close()
    _say(f"  plan_reused={rep['plan_reused']} iters solver={rep['ir_iterations']} fresh={res['ir_iterations']}")
    assert np.array_equal(X, x32)
    assert rep["ir_iterations"] == res["ir_iterations"]


@pytest.mark.parametrize("large", [4, 1])
def test_side_streams_match_one_stream(gpu, monkeypatch, large):
    """A refactor that spreads the dense-block LU and the dense tails of several
    sparse blocks over side streams, or (one sparse block) runs each next
    diagonal LU of the dense tail on a side stream beside the trailing update
    (lookahead), joined by events and captured in one graph, gives exactly what
    the same work on one stream gives."""
    bd = gpu
    _say(f"case: side streams vs one stream, {large} large block(s)")
    sp = bd.config_spec(1)
    sp.seed = 700 + large
    sp.n_small_blocks = 80
    sp.n_small_total = 1600
    sp.n_large_blocks = large
    sp.large_size_lo, sp.large_size_hi = (200, 700) if large > 1 else (2500, 2500)
    n, r, c, v, _ = bd.generate(sp)
    rng = np.random.default_rng(9)
    reqs = [(bd.perturb(v, seed=90 + i), np.asfortranarray(rng.uniform(-1, 1, (n, 1 + i % 3)))) for i in range(4)]
    outs = {}
    for mode in ("4", "1"):
        monkeypatch.setenv("BLOCKDAGLU_GPU_STREAMS", mode)
        for prec in (bd.Precision.FP64, bd.Precision.FP32_IR):
            srv = bd.Server(slots=1, hbm_budget_gb=8.0, precision=prec)
            pid = srv.add_pattern(r, c, v, n)
            outs[(mode, prec)] = [srv.wait(srv.submit(pid, vals, B)) for vals, B in reqs]
            srv.close()
    for prec in (bd.Precision.FP64, bd.Precision.FP32_IR):
        for (xs, rs), (x1, r1) in zip(outs[("4", prec)], outs[("1", prec)]):
            assert rs["error"] == bd.SolverError.None_ and r1["error"] == bd.SolverError.None_
            assert np.array_equal(xs, x1), "side-stream refactor differs from one stream"
            assert rs["ir_iterations"] == r1["ir_iterations"]
    _say(f"  n={n} requests={len(reqs)} equal")


@pytest.mark.parametrize("precision", [0, 1])
def test_device_resident_requests_equal_host_ones(gpu, precision):
    """Requests whose canonical values, right-hand sides and solution buffers
    are in device memory (submit_device) give exactly the solutions and
    statistics of the same requests handed over as host buffers."""
    bd = gpu
    prec = bd.Precision(precision)
    _say(f"case: device-resident vs host-buffer requests precision={prec.name}")
    pats = _patterns(bd)
    reqs = _requests(bd, pats, 12, seed=21 + precision)
    srv = bd.Server(slots=4, hbm_budget_gb=8.0, precision=prec)
    pid = [srv.add_pattern(r, c, v, n) for (n, r, c, v) in pats]
    host = [srv.wait(srv.submit(pid[p], vals, B)) for (p, vals, B) in reqs]
    keep, ids = [], []
    for p, vals, B in reqs:
        n, r, c, _ = pats[p]
        order = bd.canonical_order(r, c)
        dv = bd.DeviceArray.from_numpy(vals[order])
        dB = bd.DeviceArray.from_numpy(B.ravel(order="F"))
        dX = bd.DeviceArray(B.size)
        keep.append((dv, dB, dX))
        ids.append(srv.submit_device(pid[p], dv.address, dB.address, dX.address, B.shape[1]))
    dev = [srv.wait(i)[1] for i in ids]
    srv.close()
    for (p, vals, B), (xh, rh), rd, (_, dB, dX) in zip(reqs, host, dev, keep):
        xd = dX.numpy().reshape(B.shape, order="F")
        assert rh["error"] == bd.SolverError.None_ and rd["error"] == bd.SolverError.None_
        assert np.array_equal(xd, xh.reshape(B.shape, order="F")), "device-resident request differs"
        assert np.array_equal(dB.numpy(), B.ravel(order="F")), "right-hand sides overwritten"
        assert rd["ir_iterations"] == rh["ir_iterations"] and rd["rel_residual"] == rh["rel_residual"]
    _say(f"  {len(reqs)} requests equal")
