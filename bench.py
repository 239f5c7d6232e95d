#!/usr/bin/env python3
"""Refactor+solve throughput of the MI355X BLOCK engine.

Workload (one GPU): the circuit-like system — n = 200,000, ~1.29M nonzeros,
one 20,000 power-law block plus 9,000 blocks of size 1-40, coupling 0.5/row,
hidden by random row/column permutations (seeded generator). A step is one
refactorization of the resident pattern with new values (10 pre-generated
+-5% perturbations, cycled, already in HBM) followed by one solve of a 1-RHS
system (RHS in HBM). Reported:
  value            FP32 factors + FP64 refinement to ||b-Ax||/||b|| < 1e-12
  fp64             the plain FP64 GPU path (no refinement, the reference's default)
  latency p50/p99  per step, host clock around the synchronous calls
  roofline         the dominant kernel (dense-tail trailing update), per-launch
                   algorithmic bytes/flops over its average HIP-event duration
  cpu_baseline     the CPU restatement of the reference BLOCK route (oracle/),
                   one request = set_matrix -> analyze -> factor -> solve
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "refactor+solve/s, p50/p99 latency; FP32+IR vs FP64 GPU at rel-resid≤1e-12"
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30  # the plain copy that calibrates the HBM byte counters
MFMA_PEAK = {"f64": 78.6, "f32": 157.3}  # TFLOP/s dense


TRAFFIC_FILE = ROOT / "profiles" / "traffic_tail_update_circuit200k.json"
ROOFLINE_KERNEL = "k_tail_update<float, *>"


def is_roofline_kernel(name: str) -> bool:
    """Both tile variants of the FP32 trailing update, by demangled or mangled
    name (not the fused launch k_tail_update_diag of the timed steps)."""
    return "k_tail_update<float," in name or "k_tail_updateIfLi" in name


def _db_view(db, name):
    return name if any(r[0] == name for r in db.execute("select name from sqlite_master")) else None


def profile_summary(db_path: str) -> str:
    """Markdown table of a rocprofv3 --kernel-trace database: per kernel, calls and
    GPU time (all dispatches of the run)."""
    import sqlite3
    db = sqlite3.connect(db_path)
    rows = list(db.execute(
        "select s.kernel_name, count(*), sum(d.end - d.start), min(d.end - d.start), max(d.end - d.start) "
        "from rocpd_kernel_dispatch d join rocpd_info_kernel_symbol s on d.kernel_id = s.id "
        "group by s.kernel_name order by sum(d.end - d.start) desc"))
    total = sum(r[2] for r in rows) or 1
    out = ["| kernel | calls | total ms | avg us | min us | max us | share |", "|---|---:|---:|---:|---:|---:|---:|"]
    fam = [0, 0]
    for name, cnt, tot, mn, mx in rows:
        short = name.split("(")[0].replace("void ", "")
        if is_roofline_kernel(name):
            fam[0] += cnt
            fam[1] += tot
        out.append(f"| `{short}` | {cnt} | {tot / 1e6:.2f} | {tot / cnt / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | "
                   f"{100.0 * tot / total:.1f}% |")
    if fam[0]:
        out.append(f"\nRoofline kernel `{ROOFLINE_KERNEL}` (both tile variants): {fam[0]} launches, "
                   f"average {fam[1] / fam[0] / 1e3:.2f} us.")
    return "\n".join(out)


def traffic_from_counters(fetch_db: str, write_db: str) -> dict:
    """HBM bytes per dispatch of the roofline kernel from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by what the same
    counter reads for the plain 1 GiB copy that --copy-calibration runs."""
    import sqlite3
    out = {}
    for which, path in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        db = sqlite3.connect(path)
        # the calibration copies are the run's first three copy dispatches
        cal = [v for (v,) in db.execute(
            "select value from counters_collection where kernel_name like '%copyBuffer%' and counter_name = ? "
            "order by dispatch_id limit 3", (which,))]
        if not cal:
            raise RuntimeError(f"no {COPY_BYTES}-byte copy dispatch in {path} (run bench.py with --copy-calibration)")
        factor = COPY_BYTES / (float(np.median(cal)) * 1024.0)
        vals = [v for (name, v) in db.execute(
            "select kernel_name, value from counters_collection where counter_name = ?", (which,))
            if is_roofline_kernel(name)]
        out[which] = {"copy_counter_kib": float(np.median(cal)), "factor": factor, "dispatches": len(vals),
                      "mean_kib": float(np.mean(vals)), "mean_bytes_scaled": float(np.mean(vals)) * 1024.0 * factor}
    out["bytes_per_launch"] = out["FETCH_SIZE"]["mean_bytes_scaled"] + out["WRITE_SIZE"]["mean_bytes_scaled"]
    # The same counters without the copy factor: FETCH_SIZE is derived from read
    # requests counted in 64-byte units, which reads a 128-byte request (as the
    # copy issues) as half its size but a 64-byte request in full.
    out["bytes_per_launch_unscaled"] = 1024.0 * (out["FETCH_SIZE"]["mean_kib"] + out["WRITE_SIZE"]["mean_kib"])
    out["kernel"] = ROOFLINE_KERNEL
    out["commands"] = [
        "rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_fetch -o run -- python3 bench.py --steps 2 --warmup 1 "
        "--no-cpu-baseline --copy-calibration",
        "rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_write -o run -- python3 bench.py --steps 2 --warmup 1 "
        "--no-cpu-baseline --copy-calibration",
        "python bench.py --traffic gpurun_out/pmc_fetch/run_results.db gpurun_out/pmc_write/run_results.db"]
    return out


def load_package():
    import importlib.util
    if "blockdaglu_amd" in sys.modules:
        return sys.modules["blockdaglu_amd"]
    spec = importlib.util.spec_from_file_location("blockdaglu_amd", ROOT / "block-daglu_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["blockdaglu_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_residual(n, r, c, v, x, b):
    """||b - A x||_2 / ||b||_2, the worst column when x and b hold several."""
    X = x.reshape(n, -1)
    Bm = b.reshape(n, -1)
    worst = 0.0
    for j in range(X.shape[1]):
        y = np.zeros(n)
        np.add.at(y, r, v * X[c, j])
        worst = max(worst, float(np.linalg.norm(Bm[:, j] - y) / np.linalg.norm(Bm[:, j])))
    return worst


def backward_error(n, r, c, v, x, b):
    """||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), the worst column."""
    X = x.reshape(n, -1)
    Bm = b.reshape(n, -1)
    rowsum = np.zeros(n)
    np.add.at(rowsum, r, np.abs(v))
    worst = 0.0
    for j in range(X.shape[1]):
        y = np.zeros(n)
        np.add.at(y, r, v * X[c, j])
        den = rowsum.max() * np.abs(X[:, j]).max() + np.abs(Bm[:, j]).max()
        worst = max(worst, float(np.abs(Bm[:, j] - y).max() / den))
    return worst


def componentwise_error(n, r, c, v, x, b):
    """max_i |b - A x|_i / (|A| |x| + |b|)_i (Oettli-Prager), the worst column."""
    X = x.reshape(n, -1)
    Bm = b.reshape(n, -1)
    worst = 0.0
    for j in range(X.shape[1]):
        y = np.zeros(n)
        np.add.at(y, r, v * X[c, j])
        den = np.abs(Bm[:, j]).copy()
        np.add.at(den, r, np.abs(v * X[c, j]))
        res = np.abs(Bm[:, j] - y)
        ratio = np.divide(res, den, out=np.where(res > 0, np.inf, 0.0), where=den > 0)
        worst = max(worst, float(ratio.max()))
    return worst


def run_gpu(bd, n, r, c, v_sets, b, precision, steps, warmup, profile):
    """Returns (seconds per step list, final x of the last step, report, prof)."""
    L = bd.lib()
    s = bd.Solver()
    s.set_precision(bd.Precision(precision))
    assert s.set_matrix(r, c, v_sets[0], n), s.last_error_message()
    t0 = time.time()
    assert s.analyze(), s.last_error_message()
    log(f"analyze (plan build + upload) {time.time() - t0:.2f}s precision={precision}")
    assert s.factor(), s.last_error_message()
    order = bd.canonical_order(r, c)
    # Value sets and right-hand sides resident in HBM before the timed region.
    vals_dev = [bd.DeviceArray.from_numpy(vs[order]) for vs in v_sets]
    bk = b.reshape(n, -1)
    nrhs = bk.shape[1]
    b_dev = bd.DeviceArray.from_numpy(bk.ravel(order="F"))  # n x k column-major
    x_dev = bd.DeviceArray(n * nrhs)
    stream = s.device_stream()
    ctx = C.c_void_p(s.device_context())

    def step(i):
        x_dev.copy_from(b_dev, stream)  # on the engine stream, ahead of the refactor
        ok = s.factor_device(vals_dev[i % len(v_sets)].address)
        assert ok, s.last_error_message()
        ok = s.solve_device(x_dev.address, nrhs, n)
        assert ok, s.last_error_message()

    for i in range(warmup):
        t = time.perf_counter()
        step(i)
        log(f"  warmup step {i}: {1e3 * (time.perf_counter() - t):.2f} ms")
    bd.hip().hipDeviceSynchronize()
    ev0, ev1 = bd.Event(), bd.Event()
    times = []
    ev0.record(stream)
    t_start = time.perf_counter()
    for i in range(steps):
        t = time.perf_counter()
        step(i)
        times.append(time.perf_counter() - t)
    ev1.record(stream)
    ev1.synchronize()
    wall = time.perf_counter() - t_start
    gpu_ms = ev0.elapsed_ms(ev1)
    rep = s.report()
    log(f"  precision={precision} refine={os.environ.get('USERALGO_REFINE_ITERS', '0')} steps={steps} "
        f"wall={wall:.3f}s gpu={gpu_ms:.1f}ms p50={1e3 * float(np.median(times)):.2f}ms "
        f"max={1e3 * max(times):.2f}ms ir_iters={rep['ir_iterations']} repivots={rep['repivot_blocks']} "
        f"fp64_fallbacks={rep['fp64_fallback_blocks']} rr_engine={rep['rel_residual']:.2e}")
    x = x_dev.numpy().reshape(nrhs, n).T  # n x k
    last_vals = v_sets[(steps - 1) % len(v_sets)]
    prof = None
    if profile:
        L.gpu.bdlu_gpu_profile(ctx, 1)
        for i in range(steps):
            step(i)
        prof = {}
        for kind, name in ((bd.PROF_TAIL_GEMM, "tail_gemm"), (bd.PROF_TAIL_PANEL, "tail_panel"),
                           (bd.PROF_DENSE_LU, "dense_lu"), (bd.PROF_SPARSE, "sparse_cols"),
                           (bd.PROF_SOLVE, "solve")):
            ms, cnt, fl, by = C.c_double(), C.c_int64(), C.c_double(), C.c_double()
            L.gpu.bdlu_gpu_profile_read(ctx, kind, C.byref(ms), C.byref(cnt), C.byref(fl), C.byref(by))
            prof[name] = {"ms": ms.value, "launches": cnt.value, "flops": fl.value, "bytes": by.value}
        L.gpu.bdlu_gpu_profile(ctx, 0)
    s.close()
    return times, wall, gpu_ms, x, last_vals, rep, prof


CPU_SCRIPT = r"""
import sys, time
sys.path.insert(0, sys.argv[1])
import importlib.util
spec = importlib.util.spec_from_file_location("blockdaglu_amd", sys.argv[1] + "/block-daglu_amd/__init__.py")
bd = importlib.util.module_from_spec(spec); sys.modules["blockdaglu_amd"] = bd; spec.loader.exec_module(bd)
from oracle import oracle as orc
n, r, c, v, B = bd.generate(int(sys.argv[2]), keep_order=int(sys.argv[3]))
o = orc.OracleSolver(r, c, v)
t = time.perf_counter(); assert o.analyze() == 0; print("analyze", time.perf_counter() - t, flush=True)
t = time.perf_counter(); assert o.factor() == 0; print("factor", time.perf_counter() - t, o.info()["threads"], flush=True)
t = time.perf_counter(); rc, x = o.solve(B[:, 0]); print("solve", time.perf_counter() - t, flush=True)
"""


def cpu_baseline(config, budget_s, keep_order=0):
    """The oracle (CPU restatement of the reference BLOCK route) on one request of
    the same workload: set_matrix -> analyze -> factor -> solve (the reference API
    has no pattern-only refactor), reference thread policy (src/solver.cpp:366-390),
    in a child process stopped after budget_s seconds. If the request completes,
    its throughput is measured; otherwise 1/budget_s is an upper bound on it and
    the sample says which phases completed."""
    import subprocess
    cmd = [sys.executable, "-c", CPU_SCRIPT, str(ROOT), str(config), str(keep_order)]
    t0 = time.perf_counter()
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    try:
        out, _ = proc.communicate(timeout=budget_s)
        complete = proc.returncode == 0
    except subprocess.TimeoutExpired:
        proc.kill()  # the child started here, by its own PID
        out, _ = proc.communicate()
        complete = False
    wall = time.perf_counter() - t0
    phases = {}
    threads = 1
    for line in out.splitlines():
        f = line.split()
        if len(f) >= 2 and f[0] in ("analyze", "factor", "solve"):
            phases[f[0]] = float(f[1])
            if f[0] == "factor" and len(f) > 2:
                threads = int(f[2])
    done = ", ".join(f"{k} {v:.2f}s" for k, v in phases.items()) or "none"
    if complete and len(phases) == 3:
        total = sum(phases.values())
        return {"value": 1.0 / total, "unit": "refactor+solve/s", "cores": threads, "kind": "port",
                "sample": f"one complete request of the same workload ({done}); factor threads per the "
                          f"reference policy", "measured": True}
    return {"value": 1.0 / wall, "unit": "refactor+solve/s", "cores": max(threads, 1), "kind": "port",
            "sample": f"one request of the same workload stopped at the {budget_s:.0f} s sample limit; phases "
                      f"completed: {done}; value = 1/{wall:.1f} s is an upper bound on CPU throughput",
            "measured": False}


def bench_serving(bd, args):
    """Serving stream: 16 seeded patterns (config 4), 2000 requests with Poisson
    arrivals at a fixed rate (0.8 x the measured FP64 capacity, the same rate for
    both precisions), each a +-5% value perturbation of its pattern with 1-8
    right-hand sides; 4 slots, 64 GB HBM budget."""
    import threading
    rng = np.random.default_rng(4)
    pats = []
    t = time.time()
    for p in range(16):
        n, r, c, v, _ = bd.generate(4, pattern=p)
        vals = [bd.perturb(v, seed=40_000 + 16 * p + i) for i in range(6)]
        rhs = np.asfortranarray(rng.uniform(-1, 1, (n, 8)))
        pats.append((n, r, c, v, vals, rhs))
    log(f"serving: 16 patterns n={[p[0] for p in pats]} generated in {time.time() - t:.1f}s")
    nreq = args.requests
    req_pat = rng.integers(0, 16, nreq)
    req_val = rng.integers(0, 6, nreq)
    req_k = rng.integers(1, 9, nreq)
    gaps = rng.exponential(1.0, nreq)  # scaled by 1/lambda below
    out = {}
    lam = None
    for prec in (bd.Precision.FP64, bd.Precision.FP32_IR):
        srv = bd.Server(slots=args.slots, hbm_budget_gb=64.0, precision=prec)
        t = time.time()
        pid = []
        for k, p in enumerate(pats):
            pid.append(srv.add_pattern(p[1], p[2], p[3], p[0]))
            log(f"  {prec.name}: pattern {k} (n={p[0]}) registered, {time.time() - t:.1f}s")
        t_an = time.time() - t
        # warm every slot on every pattern (plan upload, first factor)
        ids = [srv.submit(pid[p], pats[p][4][0], pats[p][5][:, :1]) for p in range(16) for _ in range(args.slots)]
        for k, i in enumerate(ids):
            srv.wait(i)
            if (k + 1) % args.slots == 0:
                log(f"  {prec.name}: warm-up, pattern {k // args.slots} on every slot, {time.time() - t:.1f}s")
        # capacity: closed loop, all requests queued at once
        ncap = min(nreq, 400)
        t = time.perf_counter()
        ids = [srv.submit(pid[req_pat[i]], pats[req_pat[i]][4][req_val[i]], pats[req_pat[i]][5][:, :req_k[i]])
               for i in range(ncap)]
        for k, i in enumerate(ids):
            srv.wait(i)
            if (k + 1) % 100 == 0:
                log(f"  {prec.name}: capacity run {k + 1}/{ncap}")
        cap = ncap / (time.perf_counter() - t)
        if lam is None:
            lam = 0.8 * cap
        log(f"  {prec.name}: analysis {t_an:.1f}s, closed-loop capacity {cap:.1f} req/s; open loop at {lam:.1f} req/s")
        # open loop
        arrivals = np.cumsum(gaps / lam)
        ids = [None] * nreq

        def feeder():
            t0 = time.perf_counter()
            for i in range(nreq):
                d = arrivals[i] - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                p = req_pat[i]
                ids[i] = srv.submit(pid[p], pats[p][4][req_val[i]], pats[p][5][:, :req_k[i]])

        th = threading.Thread(target=feeder)
        t = time.perf_counter()
        th.start()
        res = []
        for i in range(nreq):
            while ids[i] is None:
                time.sleep(0.0002)
            X, rr = srv.wait(ids[i])
            res.append(rr)
            if (i + 1) % 500 == 0:
                log(f"    {i + 1}/{nreq} done")
        th.join()
        wall = time.perf_counter() - t
        lat = np.array([r["t_done_ms"] - r["t_submit_ms"] for r in res])
        errs = [r for r in res if r["error"] != bd.SolverError.None_]
        out[prec.name] = {
            "capacity_req_per_s": cap, "offered_req_per_s": lam, "served_req_per_s": nreq / wall,
            "solves_per_s": float(req_k.sum()) / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_rel_residual": max(r["rel_residual"] for r in res),
            "max_backward_error": max(r["backward_error"] for r in res),
            "max_ir_iterations": max(r["ir_iterations"] for r in res),
            "fp64_fallback_blocks": int(sum(r["fp64_fallback_blocks"] for r in res)),
            "repivots": int(sum(r["repivots"] for r in res)), "errors": len(errs), "analysis_s": t_an}
        log(f"  {prec.name}: p50 {out[prec.name]['p50_ms']:.2f} ms p99 {out[prec.name]['p99_ms']:.2f} ms "
            f"served {out[prec.name]['served_req_per_s']:.1f} req/s errors {len(errs)}")
        srv.close()
    line = {"metric": "served refactor+solve/s, p50/p99 latency (serving stream, 16 patterns)",
            "value": out["FP32_IR"]["capacity_req_per_s"], "unit": "refactor+solve/s (closed-loop capacity)",
            "n_gpus": 1, "higher_is_better": True, "dtype": "f32 factors + f64 refinement",
            "data": "synthetic (seeded generator)",
            "config": {"workload": f"serving: 16 patterns n=1e3..1e5, {nreq} Poisson requests at "
                                   f"{lam:.1f}/s, 1-8 RHS, {args.slots} slots, 64 GB budget"},
            "fp32ir": out["FP32_IR"], "fp64": out["FP64"],
            "fp32ir_over_fp64_capacity": out["FP32_IR"]["capacity_req_per_s"] / out["FP64"]["capacity_req_per_s"],
            "build": bd.build_info()}
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1,
                    help="workload: 1 circuit-like n=200,000 (default); 4 serving stream")
    ap.add_argument("--requests", type=int, default=2000, help="serving stream length (config 4)")
    ap.add_argument("--slots", type=int, default=4, help="serving slots (config 4)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=25.0, help="seconds of CPU baseline sample")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--keep-order", action="store_true",
                    help="skip the generator's structure-hiding permutation (same blocks and values); the "
                         "batched-blocks workload needs it for a host analysis of seconds")
    ap.add_argument("--ir-stop", choices=["residual", "backward", "componentwise"],
                    help="refinement stop rule at 1e-12 (default: componentwise for the scaled stress case, "
                         "the reference's ||r||/||b|| otherwise)")
    ap.add_argument("--profile-summary", metavar="DB", help="print the kernel table of a rocprofv3 database and exit")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="compute HBM bytes per launch of the roofline kernel from two --pmc runs, write "
                         "profiles/traffic_tail_update_circuit200k.json and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device-to-device copy (the known-size reference kernel for "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs)")
    args = ap.parse_args()
    if args.profile_summary:
        print(profile_summary(args.profile_summary))
        return
    if args.traffic:
        t = traffic_from_counters(*args.traffic)
        t["kernel"] = ROOFLINE_KERNEL
        TRAFFIC_FILE.write_text(json.dumps(t, indent=1) + "\n")
        print(json.dumps(t, indent=1))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build's path has no multi-GPU mode (see the sharding note "
              "in the design document); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    bd = load_package()
    if not bd.gpu_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        x = bd.DeviceArray.from_numpy(np.ones(COPY_BYTES // 8))
        y = bd.DeviceArray(COPY_BYTES // 8)
        for _ in range(3):
            y.copy_from(x)
        bd.hip().hipDeviceSynchronize()
        x.free()
        y.free()
        log(f"copy calibration: 3 x {COPY_BYTES} B device-to-device")
    if args.config == 4:
        bench_serving(bd, args)
        return
    t = time.time()
    n, r, c, v, B = bd.generate(args.config, keep_order=int(args.keep_order))
    b = np.asfortranarray(B)  # n x k right-hand sides
    v_sets = [v] + [bd.perturb(v, seed=1000 + i) for i in range(1, 10)]
    log(f"workload n={n} nnz={len(r)} rhs={b.shape[1]} generated in {time.time() - t:.1f}s")
    workload = {
        1: f"circuit-like n={n} nnz={len(r)}: one 20000 power-law block + 9000 blocks of size 1-40, "
           f"10 refactors at +-5%, 1 RHS",
        2: f"batched blocks n={n} nnz={len(r)}: 50000 blocks log-uniform 1-64 + 8 dense-ish 512-2048 (30% fill), "
           f"10 refactors at +-5%, {b.shape[1]} RHS",
        3: f"mixed-precision stress n={n} nnz={len(r)}: circuit-like pattern, rows/cols scaled by 10^U(-4,4), "
           f"10 refactors at +-5%, 1 RHS, FP32+IR to backward error <= 1e-12",
    }.get(args.config, f"config {args.config} n={n} nnz={len(r)}")
    if args.keep_order:
        workload += (" (generated order kept: the random permutation that hides the block structure is skipped, "
                     "so the one-time host analysis is seconds instead of a long maximum-transversal search; "
                     "blocks, fills, coupling and values are the same)")

    # Plain FP64 (the reference default, no refinement), FP64 + FP64 refinement to
    # ||r||/||b|| < 1e-12 (REFINE_ITERS semantics), and FP32 factors + FP64 refinement.
    # The badly scaled stress case stops both refinements on componentwise backward
    # error (scale-invariant); ||r||/||b|| is dominated by the largest rows there.
    ir_stop = args.ir_stop or ("componentwise" if args.config == 3 else "residual")
    os.environ.pop("USERALGO_REFINE_ITERS", None)
    os.environ.pop("BLOCKDAGLU_GPU_IR_STOP", None)
    res64p = run_gpu(bd, n, r, c, v_sets, b, 0, args.steps, args.warmup, False)
    if ir_stop != "residual":
        os.environ["BLOCKDAGLU_GPU_IR_STOP"] = ir_stop
    os.environ["USERALGO_REFINE_ITERS"] = "10"
    res64 = run_gpu(bd, n, r, c, v_sets, b, 0, args.steps, args.warmup, False)
    os.environ.pop("USERALGO_REFINE_ITERS", None)
    res32 = run_gpu(bd, n, r, c, v_sets, b, 1, args.steps, args.warmup, not args.no_profile)
    os.environ.pop("BLOCKDAGLU_GPU_IR_STOP", None)

    def summarize(res):
        times, wall, gpu_ms, x, last_vals, rep, prof = res
        ts = np.array(times)
        return {"value": args.steps / wall, "ms_per_step": 1e3 * wall / args.steps,
                "gpu_ms_per_step": gpu_ms / args.steps,
                "p50_ms": float(np.percentile(ts, 50) * 1e3), "p99_ms": float(np.percentile(ts, 99) * 1e3),
                "rel_residual": rel_residual(n, r, c, last_vals, x, b),
                "backward_error": backward_error(n, r, c, last_vals, x, b),
                "componentwise_error": componentwise_error(n, r, c, last_vals, x, b),
                "ir_iterations": rep["ir_iterations"], "fp64_fallback_blocks": rep["fp64_fallback_blocks"]}

    s64p, s64, s32 = summarize(res64p), summarize(res64), summarize(res32)
    prof = res32[6]
    roof = None
    if prof and prof["tail_gemm"]["launches"] > 0:
        g = prof["tail_gemm"]
        avg_s = g["ms"] / g["launches"] / 1e3
        flops = g["flops"] / g["launches"]
        byts = g["bytes"] / g["launches"]
        tf = flops / avg_s / 1e12
        gbs = byts / avg_s / 1e9
        intensity = flops / byts
        ridge = MFMA_PEAK["f32"] * 1e12 / (HBM_PEAK_GBS * 1e9)
        if intensity < ridge:
            roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
        else:
            roof = {"bound": "mfma", "achieved": tf, "peak": MFMA_PEAK["f32"], "unit": "TFLOP/s",
                    "frac": tf / MFMA_PEAK["f32"]}
        traffic = None
        if TRAFFIC_FILE.exists():
            traffic = json.loads(TRAFFIC_FILE.read_text())["bytes_per_launch"]
        roof.update({"traffic": traffic, "kernel": ROOFLINE_KERNEL, "avg_launch_us": avg_s * 1e6,
                     "launches_per_step": g["launches"] / args.steps, "algorithmic_bytes_per_launch": byts,
                     "algorithmic_flops_per_launch": flops, "tflops": tf, "gbs": gbs})
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(args.config, args.cpu_budget, int(args.keep_order))
        log(f"cpu baseline: {cpu['sample']}")
    out = {
        "metric": METRIC,
        "value": s32["value"],
        "unit": "refactor+solve/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factors + f64 refinement",
        "data": "synthetic (seeded generator)",
        "config": {"workload": workload, "parallelism": "single GPU", "keep_order": bool(args.keep_order)},
        "p50_ms": s32["p50_ms"], "p99_ms": s32["p99_ms"], "rel_residual": s32["rel_residual"],
        "backward_error": s32["backward_error"], "componentwise_error": s32["componentwise_error"],
        "ir_stop_rule": ir_stop,
        "forward_error_vs_fp64": float(np.linalg.norm(res32[3] - res64[3]) / np.linalg.norm(res64[3])),
        "ir_iterations": s32["ir_iterations"], "fp64_fallback_blocks": s32["fp64_fallback_blocks"],
        "fp64": {"value": s64["value"], "ms_per_step": s64["ms_per_step"], "p50_ms": s64["p50_ms"],
                 "p99_ms": s64["p99_ms"], "rel_residual": s64["rel_residual"], "ir_iterations": s64["ir_iterations"],
                 "backward_error": s64["backward_error"], "componentwise_error": s64["componentwise_error"],
                 "dtype": "f64 + f64 refinement"},
        "fp64_plain": {"value": s64p["value"], "ms_per_step": s64p["ms_per_step"], "p50_ms": s64p["p50_ms"],
                       "p99_ms": s64p["p99_ms"], "rel_residual": s64p["rel_residual"],
                       "backward_error": s64p["backward_error"],
                       "componentwise_error": s64p["componentwise_error"], "dtype": "f64"},
        "fp32ir_over_fp64": s32["value"] / s64["value"],
        "fp32ir_over_fp64_plain": s32["value"] / s64p["value"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_profile": prof,
        "build": bd.build_info(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
