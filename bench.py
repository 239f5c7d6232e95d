#!/usr/bin/env python3
"""Refactor+solve throughput of the MI355X BLOCK engine.

Default workload (one GPU): the serving stream. 16 seeded sparsity patterns
built like the circuit-like system (n log-uniform in 1e3..1e5, one giant block
of 10% of n, blocks of size 1-40, coupling 0.5/row, hidden by random row and
column permutations); each request brings +-5% new values for one pattern and
1-8 right-hand sides. Requests run on 4 slots (engine contexts) under one 64 GB
HBM budget, with symbolic state and factors resident. Reported:
  value            closed-loop capacity (refactor+solve requests/s) with FP32
                   factors + FP64 refinement to ||b-Ax||/||b|| < 1e-12; a step
                   is a batch of 100 requests
  p50/p99          arrival-to-result latency of 2000 Poisson arrivals at 0.8 x
                   the smallest capacity of the three passes (the same rate for
                   every pass)
  fp64_plain       the same for the plain FP64 GPU path (no refinement, the
                   reference's default; its residual misses the 1e-12 gate)
  fp64_ir          FP64 factors with FP64 refinement (the FP64 configuration
                   that meets the gate)
  closed_loop      refactor+solve steps on the circuit-like n = 200,000 system
                   (one 20,000 power-law block + 9,000 blocks of size 1-40):
                   plain FP64, FP64 + refinement, FP32 + FP64 refinement
  roofline         the dominant kernel of that system (the fused dense-tail
                   bulk trailing update, K = 128), per-launch algorithmic
                   bytes/flops over its average HIP-event duration in the timed
                   launch structure
  cpu_baseline     the CPU restatement of the reference BLOCK route (oracle/)
                   on complete requests of the serving stream, each
                   set_matrix -> analyze -> factor -> solve
--config 0..3 runs the closed loop on that configuration instead.
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


TRAFFIC_FILE = ROOT / "profiles" / "traffic_tail_update_diag_circuit200k.json"
ROOFLINE_KERNEL = "k_tail_update_diag<float, *, 128>"


def is_roofline_kernel(name: str) -> bool:
    """Both tile variants of the FP32 bulk trailing update with K = 128 (the
    launch the timed steps run for each pair of 64-column panels: the update of
    the rest of the tail plus the next diagonal block's LU), by demangled or
    mangled name. The strip launches between paired panels (K = 64) are not
    included."""
    if "k_tail_update_diag<float," in name:
        return name.replace(" ", "").split(">")[0].endswith(",128")
    return "k_tail_update_diagIfLi" in name and "ELi128EEE" in name


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
        "rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_fetch -o run -- python3 bench.py --config 1 --steps 2 "
        "--warmup 1 --no-cpu-baseline --copy-calibration",
        "rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_write -o run -- python3 bench.py --config 1 --steps 2 "
        "--warmup 1 --no-cpu-baseline --copy-calibration",
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
    # The BLOCK route is the path measured (AUTO would send systems below
    # 50,000 unknowns to the global-KLU route).
    os.environ["USERALGO_MODE"] = "block"
    assert s.analyze(), s.last_error_message()
    assert s.selected_mode() == "BLOCK"
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
import bench
bench.cpu_child(sys.argv[2:])
"""


def cpu_info() -> dict:
    """The host CPU: model, physical cores of the machine and the CPUs this
    process may run on."""
    info = {"model": "unknown", "physical_cores": None, "allowed_cpus": len(os.sched_getaffinity(0))}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name", "unknown")
        if "Core(s) per socket" in kv and "Socket(s)" in kv:
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except Exception:
        pass
    return info


def serving_stream(bd, nreq, seed=4):
    """The serving workload: 16 seeded patterns built like the circuit-like
    system (n log-uniform in 1e3..1e5), six +-5% value sets per pattern, and a
    request stream (pattern, value set, right-hand-side count 1..8, exponential
    inter-arrival gaps for unit rate) from one seeded generator."""
    rng = np.random.default_rng(seed)
    pats = []
    for p in range(16):
        n, r, c, v, _ = bd.generate(4, pattern=p)
        vals = [bd.perturb(v, seed=40_000 + 16 * p + i) for i in range(6)]
        rhs = np.asfortranarray(rng.uniform(-1, 1, (n, 8)))
        pats.append((n, r, c, v, vals, rhs))
    req_pat = rng.integers(0, 16, nreq)
    req_val = rng.integers(0, 6, nreq)
    req_k = rng.integers(1, 9, nreq)
    gaps = rng.exponential(1.0, nreq)
    return pats, req_pat, req_val, req_k, gaps


def cpu_child(argv):
    """Child process of the CPU baseline: the oracle (CPU restatement of the
    reference BLOCK route) runs complete requests one after another, each
    set_matrix -> analyze -> factor -> solve of every right-hand side (the
    reference API has no pattern-only refactor), and prints one line per
    completed request. argv: config, keep_order, threads (0: the reference's
    thread policy), requests."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("blockdaglu_amd", ROOT / "block-daglu_amd" / "__init__.py")
    bd = importlib.util.module_from_spec(spec)
    sys.modules["blockdaglu_amd"] = bd
    spec.loader.exec_module(bd)
    from oracle import oracle as orc
    config, keep_order, threads, nreq = (int(a) for a in argv[:4])
    if config == 4:
        pats, req_pat, req_val, req_k, _ = serving_stream(bd, nreq)
        items = [(pats[req_pat[i]], pats[req_pat[i]][4][req_val[i]], pats[req_pat[i]][5][:, :req_k[i]])
                 for i in range(nreq)]
        items = ((p[0], p[1], p[2], vals, B) for p, vals, B in items)
    else:
        n, r, c, v, B = bd.generate(config, keep_order=keep_order)
        items = ((n, r, c, bd.perturb(v, seed=1000 + i) if i else v, B) for i in range(nreq))
    for (n, r, c, vals, B) in items:
        t0 = time.perf_counter()
        o = orc.OracleSolver(r, c, vals, n)
        assert o.analyze() == 0
        t1 = time.perf_counter()
        assert o.factor(threads) == 0
        t2 = time.perf_counter()
        for j in range(B.shape[1]):
            rc, _ = o.solve(np.ascontiguousarray(B[:, j]))
            assert rc == 0
        t3 = time.perf_counter()
        print(f"request {t3 - t0:.6f} analyze {t1 - t0:.6f} factor {t2 - t1:.6f} solve {t3 - t2:.6f} "
              f"threads {o.info()['threads']}", flush=True)


def cpu_baseline(config, budget_s, keep_order=0, nreq=1000):
    """The oracle on complete requests of the same workload, in a child process
    on this host, run twice: with the reference's thread policy
    (src/solver.cpp:341-390) and with one factor thread per allowed physical
    core. Each run keeps issuing requests in stream order until budget_s
    seconds have passed (then the child is stopped) and reports completed
    requests per second over the completed requests only, with the median and
    p95 request time; a run that completes none reports nothing measured."""
    import subprocess
    cpu = cpu_info()
    cores = min(x for x in (cpu["physical_cores"], cpu["allowed_cpus"]) if x)
    runs = {}
    for label, threads in (("reference_policy", 0), ("physical_cores", cores)):
        cmd = [sys.executable, "-c", CPU_SCRIPT, str(ROOT), str(config), str(keep_order), str(threads), str(nreq)]
        proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
        try:
            out, _ = proc.communicate(timeout=budget_s)
        except subprocess.TimeoutExpired:
            proc.kill()  # the child started here, by its own PID
            out, _ = proc.communicate()
        reqs = [dict(zip(f.split()[0::2], map(float, f.split()[1::2]))) for f in out.splitlines()
                if f.startswith("request ")]
        if not reqs:
            runs[label] = {"measured": False, "completed": 0, "budget_s": budget_s}
            continue
        t = np.array([q["request"] for q in reqs])
        runs[label] = {"measured": True, "completed": len(reqs), "value": len(reqs) / t.sum(),
                       "median_ms": float(np.median(t) * 1e3), "p95_ms": float(np.percentile(t, 95) * 1e3),
                       "threads": int(reqs[0]["threads"]),
                       "analyze_ms": float(np.median([q["analyze"] for q in reqs]) * 1e3),
                       "factor_ms": float(np.median([q["factor"] for q in reqs]) * 1e3),
                       "solve_ms": float(np.median([q["solve"] for q in reqs]) * 1e3)}
        log(f"cpu baseline ({label}): {runs[label]}")
    best = max((r for r in runs.values() if r["measured"]), key=lambda r: r["value"], default=None)
    if best is None:
        return {"value": None, "unit": "refactor+solve/s", "cores": cores, "kind": "port", "measured": False,
                "sample": f"no request of the workload completed within {budget_s:.0f} s", "cpu": cpu, "runs": runs}
    return {"value": best["value"], "unit": "refactor+solve/s", "cores": best["threads"], "kind": "port",
            "measured": True,
            "sample": f"complete requests of the same workload in stream order for up to {budget_s:.0f} s per "
                      f"thread setting (set_matrix -> analyze -> factor -> solve each); value = the faster "
                      f"setting, completed requests / their total time",
            "cpu": cpu, "runs": runs}


REQS_PER_STEP = 100  # serving stream: one step = one batch of 100 requests


def bench_serving(bd, args):
    """Serving stream: 16 seeded patterns (n 1e3..1e5), each request a +-5%
    value perturbation of its pattern with 1-8 right-hand sides; 4 slots, one
    64 GB HBM budget. Per precision:
      warm-up    every pattern once on every slot, then W x 100 requests;
      capacity   K x 100 requests queued at once (closed loop), timed from the
                 first submission to the last result: the refactor+solve/s value;
      open loop  the --requests stream with Poisson arrivals at a fixed rate
                 (0.8 x the smallest measured capacity, the same rate for every
                 pass): arrival-to-result p50/p99 latency.
    Three passes: plain FP64, FP64 with FP64 refinement (the FP64 configuration
    that meets the 1e-12 residual gate) and FP32 factors with FP64 refinement.
    FP32+IR solutions of the first 100 open-loop requests are compared with the
    FP64+IR ones of the same requests (forward difference)."""
    import threading
    t = time.time()
    nstream = max(args.requests, (args.steps + args.warmup) * REQS_PER_STEP)
    pats, req_pat, req_val, req_k, gaps = serving_stream(bd, nstream)
    log(f"serving: 16 patterns n={[p[0] for p in pats]} generated in {time.time() - t:.1f}s")
    out, keep = {}, {}
    lam = None
    ncap = args.steps * REQS_PER_STEP
    nwarm = args.warmup * REQS_PER_STEP
    dev = not args.host_buffers
    if dev:
        # Inputs resident in HBM before anything is timed: every value set in
        # canonical order, every pattern's right-hand sides, and one solution
        # region per request index (phases run one after another, and the
        # indices inside a phase are distinct).
        dvals, drhs, dwarm = [], [], []
        for n, r, c, v, vals, rhs in pats:
            order = bd.canonical_order(r, c)
            assert len(np.unique(np.asarray(r, np.int64) * n + c)) == len(r)  # no duplicates: canonical = sorted
            dvals.append([bd.DeviceArray.from_numpy(vs[order]) for vs in vals])
            drhs.append(bd.DeviceArray.from_numpy(rhs.ravel(order="F")))
            dwarm.append([bd.DeviceArray(n) for _ in range(args.slots)])
        sizes = np.array([pats[req_pat[i]][0] * req_k[i] for i in range(nstream)], np.int64)
        xoff = np.concatenate([[0], np.cumsum(sizes)])
        xbuf = bd.DeviceArray(int(xoff[-1]))
        log(f"serving: inputs resident in HBM ({xoff[-1] * 8 / 1e9:.2f} GB of solution space)")

    def submit(srv, pid, i):
        p = req_pat[i]
        if dev:
            return srv.submit_device(pid[p], dvals[p][req_val[i]].address, drhs[p].address,
                                     xbuf.address + 8 * int(xoff[i]), int(req_k[i]))
        return srv.submit(pid[p], pats[p][4][req_val[i]], pats[p][5][:, :req_k[i]])


    # Plain FP64 sets the arrival rate; FP64 with FP64 refinement is the FP64
    # configuration that meets the 1e-12 residual gate (the reference's
    # refinement, USERALGO_REFINE_ITERS, here up to 10 steps).
    passes = (("FP64", bd.Precision.FP64, None), ("FP64_IR", bd.Precision.FP64, "10"),
              ("FP32_IR", bd.Precision.FP32_IR, None))

    def set_refine(refine):
        if refine is None:
            os.environ.pop("USERALGO_REFINE_ITERS", None)
        else:
            os.environ["USERALGO_REFINE_ITERS"] = refine

    servers, caps = {}, {}
    # Phase 1, per pass: registration, warm-up, closed-loop capacity.
    for name, prec, refine in passes:
        set_refine(refine)
        srv = bd.Server(slots=args.slots, hbm_budget_gb=64.0, precision=prec)
        t = time.time()
        pid = []
        for k, p in enumerate(pats):
            pid.append(srv.add_pattern(p[1], p[2], p[3], p[0]))
        t_an = time.time() - t
        log(f"  {name}: 16 patterns registered (host analysis) in {t_an:.1f}s")
        if dev:
            ids = [srv.submit_device(pid[p], dvals[p][0].address, drhs[p].address, dwarm[p][j].address, 1)
                   for p in range(16) for j in range(args.slots)]
        else:
            ids = [srv.submit(pid[p], pats[p][4][0], pats[p][5][:, :1]) for p in range(16) for _ in range(args.slots)]
        for i in ids:
            srv.wait(i)
        ids = [submit(srv, pid, i) for i in range(nwarm)]
        for i in ids:
            srv.wait(i)
        log(f"  {name}: warm-up done ({16 * args.slots} + {nwarm} requests), {time.time() - t:.1f}s")
        # capacity: closed loop, all requests queued at once
        t0 = time.perf_counter()
        ids = [submit(srv, pid, i % nstream) for i in range(nwarm, nwarm + ncap)]
        step_ends = []
        for k, i in enumerate(ids):
            srv.wait(i)
            if (k + 1) % REQS_PER_STEP == 0:
                step_ends.append(time.perf_counter() - t0)
        wall_cap = time.perf_counter() - t0
        cap = ncap / wall_cap
        step_ms = np.diff(np.concatenate([[0.0], step_ends])) * 1e3
        log(f"  {name}: closed-loop capacity {cap:.1f} req/s over {ncap} requests")
        servers[name] = (srv, pid)
        caps[name] = (cap, wall_cap, step_ms, t_an)
    # Phase 2: open loop, every pass at the same Poisson rate, 0.8 x the smallest
    # capacity (so that no pass is offered more than it can serve).
    lam = 0.8 * min(c[0] for c in caps.values())
    for name, prec, refine in passes:
        set_refine(refine)
        srv, pid = servers[name]
        cap, wall_cap, step_ms, t_an = caps[name]
        log(f"  {name}: open loop at {lam:.1f} req/s")
        nreq = args.requests
        arrivals = np.cumsum(gaps[:nreq] / lam)
        oids = [None] * nreq

        def feeder():
            t_feed = time.perf_counter()
            for i in range(nreq):
                d = arrivals[i] - (time.perf_counter() - t_feed)
                if d > 0:
                    time.sleep(d)
                oids[i] = submit(srv, pid, i)

        th = threading.Thread(target=feeder)
        t0 = time.perf_counter()
        th.start()
        res, xs = [], []
        for i in range(nreq):
            while oids[i] is None:
                time.sleep(0.0002)
            X, rr = srv.wait(oids[i])
            res.append(rr)
            if i < 100 and not dev:
                xs.append(X)
            if (i + 1) % 500 == 0:
                log(f"    {i + 1}/{nreq} done")
        th.join()
        wall = time.perf_counter() - t0
        if dev:
            xall = xbuf.numpy()
            xs = [xall[xoff[i]:xoff[i + 1]].reshape(pats[req_pat[i]][0], req_k[i], order="F") for i in range(100)]
        keep[name] = xs
        lat = np.array([r["t_done_ms"] - r["t_submit_ms"] for r in res])
        qwait = np.array([r["t_start_ms"] - r["t_submit_ms"] for r in res])
        errs = [r for r in res if r["error"] != bd.SolverError.None_]
        out[name] = {
            "capacity_req_per_s": cap, "capacity_ms_per_step": float(wall_cap / args.steps * 1e3),
            "step_ms_p50": float(np.percentile(step_ms, 50)), "step_ms_p99": float(np.percentile(step_ms, 99)),
            "offered_req_per_s": lam, "served_req_per_s": nreq / wall,
            "solves_per_s": float(req_k[:nreq].sum()) / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "queue_wait_p99_ms": float(np.percentile(qwait, 99)),
            "max_rel_residual": max(r["rel_residual"] for r in res),
            "max_backward_error": max(r["backward_error"] for r in res),
            "max_ir_iterations": max(r["ir_iterations"] for r in res),
            "fp64_fallback_blocks": int(sum(r["fp64_fallback_blocks"] for r in res)),
            "fp64_refactors": int(sum(r["ir_fp64_refactor"] for r in res)),
            "repivots": int(sum(r["repivots"] for r in res)), "errors": len(errs), "analysis_s": t_an}
        log(f"  {name}: p50 {out[name]['p50_ms']:.2f} ms p99 {out[name]['p99_ms']:.2f} ms "
            f"served {out[name]['served_req_per_s']:.1f} req/s errors {len(errs)} "
            f"max rel residual {out[name]['max_rel_residual']:.2e}")
        srv.close()
    os.environ.pop("USERALGO_REFINE_ITERS", None)
    fwd = max(float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(keep["FP32_IR"], keep["FP64_IR"]))
    return out, lam, fwd


def closed_loop(bd, args, config):
    """Closed-loop refactor+solve steps on one system (configs 0-3): plain FP64,
    FP64 + FP64 refinement, and FP32 factors + FP64 refinement, with the
    dominant kernel's roofline from a profiling pass over the same steps."""
    t = time.time()
    n, r, c, v, B = bd.generate(config, keep_order=int(args.keep_order))
    b = np.asfortranarray(B)  # n x k right-hand sides
    v_sets = [v] + [bd.perturb(v, seed=1000 + i) for i in range(1, 10)]
    log(f"workload config {config}: n={n} nnz={len(r)} rhs={b.shape[1]} generated in {time.time() - t:.1f}s")
    workload = {
        0: f"n={n} nnz={len(r)}: 200 hidden blocks of geometric size (mean 25), 10 refactors at +-5%, 1 RHS",
        1: f"circuit-like n={n} nnz={len(r)}: one 20000 power-law block + 9000 blocks of size 1-40, "
           f"10 refactors at +-5%, 1 RHS",
        2: f"batched blocks n={n} nnz={len(r)}: 50000 blocks log-uniform 1-64 + 8 dense-ish 512-2048 (30% fill), "
           f"10 refactors at +-5%, {b.shape[1]} RHS",
        3: f"mixed-precision stress n={n} nnz={len(r)}: circuit-like pattern, rows/cols scaled by 10^U(-4,4), "
           f"10 refactors at +-5%, 1 RHS, FP32+IR to backward error <= 1e-12",
    }.get(config, f"config {config} n={n} nnz={len(r)}")
    if args.keep_order:
        workload += (" (generated order kept: the random permutation that hides the block structure is skipped, "
                     "so the one-time host analysis is seconds instead of a long maximum-transversal search; "
                     "blocks, fills, coupling and values are the same)")
    # The badly scaled stress case stops both refinements on componentwise
    # backward error (scale-invariant); ||r||/||b|| is dominated by the largest rows there.
    ir_stop = args.ir_stop or ("componentwise" if config == 3 else "residual")
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
                "ir_iterations": rep["ir_iterations"], "fp64_fallback_blocks": rep["fp64_fallback_blocks"],
                "ir_fp64_refactor": rep["ir_fp64_refactor"]}

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
        ridge = MFMA_PEAK["f32"] * 1e12 / (HBM_PEAK_GBS * 1e9)
        if flops / byts < ridge:
            roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
        else:
            roof = {"bound": "mfma", "achieved": tf, "peak": MFMA_PEAK["f32"], "unit": "TFLOP/s",
                    "frac": tf / MFMA_PEAK["f32"]}
        traffic = None
        if TRAFFIC_FILE.exists():
            traffic = json.loads(TRAFFIC_FILE.read_text())["bytes_per_launch"]
        roof.update({"traffic": traffic, "kernel": ROOFLINE_KERNEL, "workload": f"config {config}",
                     "avg_launch_us": avg_s * 1e6, "launches_per_step": g["launches"] / args.steps,
                     "algorithmic_bytes_per_launch": byts, "algorithmic_flops_per_launch": flops, "tflops": tf,
                     "tflops_frac_of_mfma_peak": tf / MFMA_PEAK["f32"], "gbs": gbs})
    return {
        "workload": workload, "value": s32["value"], "ms_per_step": s32["ms_per_step"], "p50_ms": s32["p50_ms"],
        "p99_ms": s32["p99_ms"], "rel_residual": s32["rel_residual"], "backward_error": s32["backward_error"],
        "componentwise_error": s32["componentwise_error"], "ir_stop_rule": ir_stop,
        "forward_error_vs_fp64": float(np.linalg.norm(res32[3] - res64[3]) / np.linalg.norm(res64[3])),
        "ir_iterations": s32["ir_iterations"], "fp64_fallback_blocks": s32["fp64_fallback_blocks"],
        "ir_fp64_refactor": s32["ir_fp64_refactor"],
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
        "roofline": roof, "kernel_profile": prof}


def plugin_evidence(bd, args):
    """The competition plugin path end to end: writes cases from the seeded
    workloads (the 5,000-unknown system and three serving patterns, b = A x*
    with x* ~ U(-1, 1)) in the plugin benchmark's file layout, runs the oracle
    (the CPU restatement of the reference BLOCK route) as the baseline column
    with the reference's policy (1 warm-up + repetitions, one fresh solver per
    run: set_matrix -> analyze -> factor -> solve), then runs
    bdlu_plugin_bench, which loads libblockdaglu.so through CreateAlgorithm /
    DestroyAlgorithm and writes runs.csv, margin.csv, timing_summary.csv and
    summary.txt into the output directory."""
    import subprocess
    from oracle import oracle as orc
    out = Path(args.plugin_evidence).resolve()
    cases = out / "cases"
    cases.mkdir(parents=True, exist_ok=True)
    picks = [("5000-unknown system", bd.config_spec(0))] + [
        (f"serving pattern {p}", bd.config_spec(4, p)) for p in (11, 8, 12)]
    reps = args.plugin_reps
    base_rows = ["alg,case,repetition,strategy,n,nnz,reorder_ms,lu_ms,solve_ms,total_ms,rel_err,rel_residual,ok"]
    for cid, (label, spec) in enumerate(picks, start=1):
        n, r, c, v, _ = bd.generate(spec)
        xs = np.random.default_rng(900 + cid).uniform(-1, 1, n)
        b = np.zeros(n)
        np.add.at(b, r, v * xs[c])
        d = cases / str(cid)
        d.mkdir(exist_ok=True)
        with open(d / f"{cid}_matrix.coo", "w") as f:
            f.write(f"{n} {len(r)}\n")
            f.writelines(f"{i} {j} {x!r}\n" for i, j, x in zip(r.tolist(), c.tolist(), v.tolist()))
        np.savetxt(d / f"{cid}_rhs.csv", b, fmt="%.17g")
        np.savetxt(d / f"{cid}_xstar.csv", xs, fmt="%.17g")
        for rep in range(reps + 1):
            t0 = time.perf_counter()
            o = orc.OracleSolver(r, c, v, n)
            ok = o.analyze() == 0
            t1 = time.perf_counter()
            ok = ok and o.factor() == 0
            t2 = time.perf_counter()
            rc, x = o.solve(b) if ok else (1, np.full(n, np.inf))
            t3 = time.perf_counter()
            ok = ok and rc == 0
            if rep == 0:
                continue  # warm-up
            rel_err = float(np.linalg.norm(x - xs) / np.linalg.norm(xs))
            rr = rel_residual(n, r, c, v, x, b)
            base_rows.append(f"CPUReference,{cid},{rep},BLOCK,{n},{len(r)},{1e3 * (t1 - t0):.3f},"
                             f"{1e3 * (t2 - t1):.3f},{1e3 * (t3 - t2):.3f},{1e3 * (t3 - t0):.3f},"
                             f"{rel_err:.6e},{rr:.6e},{int(ok)}")
        log(f"plugin evidence: case {cid} ({label}, n={n}) written, baseline timed")
    (cases / "baseline_runs.csv").write_text("\n".join(base_rows) + "\n")
    env = dict(os.environ, BLOCKDAGLU_BENCH_OUT=str(out), BLOCKDAGLU_BENCH_REPETITIONS=str(reps),
               BLOCKDAGLU_BENCH_WARMUP="1")
    exe = ROOT / "block-daglu_amd" / "lib" / "bdlu_plugin_bench"
    lib = ROOT / "block-daglu_amd" / "lib" / "libblockdaglu.so"
    rc = subprocess.run([str(exe), "--lib", str(lib), str(cases)], env=env).returncode
    print((out / "summary.txt").read_text() if (out / "summary.txt").exists() else "no summary written")
    sys.exit(rc)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps: closed-loop refactor+solve steps (configs 0-3) or batches of 100 requests "
                         "(serving stream)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=4,
                    help="workload: 4 serving stream (default, with the circuit-like closed loop beside it); "
                         "0-3 closed loop on one system")
    ap.add_argument("--requests", type=int, default=2000, help="open-loop serving stream length")
    ap.add_argument("--slots", type=int, default=4, help="serving slots")
    ap.add_argument("--host-buffers", action="store_true",
                    help="serving requests hand over host buffers (values, right-hand sides in, solutions out: "
                         "the PCIe-inclusive rate); default: inputs resident in HBM")
    ap.add_argument("--no-closed-loop", action="store_true", help="serving stream only")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline sample per thread setting")
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
                         f"{TRAFFIC_FILE.relative_to(ROOT)} and exit")
    ap.add_argument("--plugin-evidence", metavar="DIR",
                    help="run the competition plugin (CreateAlgorithm) on seeded cases against the CPU "
                         "reference restatement and write the evidence CSVs into DIR, then exit")
    ap.add_argument("--plugin-reps", type=int, default=10, help="repetitions per case for --plugin-evidence")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device-to-device copy (the known-size reference kernel for "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs)")
    args = ap.parse_args()
    if args.profile_summary:
        print(profile_summary(args.profile_summary))
        return
    if args.traffic:
        t = traffic_from_counters(*args.traffic)
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
    log(f"library build {bd.build_info()}")
    if args.plugin_evidence:
        plugin_evidence(bd, args)
    if args.copy_calibration:
        x = bd.DeviceArray.from_numpy(np.ones(COPY_BYTES // 8))
        y = bd.DeviceArray(COPY_BYTES // 8)
        for _ in range(3):
            y.copy_from(x)
        bd.hip().hipDeviceSynchronize()
        x.free()
        y.free()
        log(f"copy calibration: 3 x {COPY_BYTES} B device-to-device")
    if args.config != 4:
        cl = closed_loop(bd, args, args.config)
        cpu = None
        if not args.no_cpu_baseline:
            cpu = cpu_baseline(args.config, args.cpu_budget, int(args.keep_order), nreq=10)
        line = {"metric": METRIC, "value": cl["value"], "unit": "refactor+solve/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": cl["ms_per_step"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32 factors + f64 refinement", "data": "synthetic (seeded generator)",
                "config": {"workload": cl["workload"], "parallelism": "single GPU",
                           "keep_order": bool(args.keep_order)}}
        line.update({k: v for k, v in cl.items() if k not in ("workload", "value", "ms_per_step")})
        line.update({"cpu_baseline": cpu, "build": bd.build_info()})
        print(json.dumps(line))
        return
    srv, lam, fwd = bench_serving(bd, args)
    f32, f64, f64ir = srv["FP32_IR"], srv["FP64"], srv["FP64_IR"]
    line = {
        "metric": METRIC, "value": f32["capacity_req_per_s"], "unit": "refactor+solve/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": f32["capacity_ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factors + f64 refinement", "data": "synthetic (seeded generator)",
        "config": {"workload": f"serving stream: 16 patterns n=1e3..1e5 (circuit-like, giant block 10% of n), "
                               f"+-5% value perturbations, 1-8 RHS per request, {args.slots} slots, one 64 GB "
                               f"HBM budget; a step is a batch of {REQS_PER_STEP} requests queued at once "
                               f"(closed loop); latency from {args.requests} Poisson arrivals at {lam:.1f}/s; "
                               + ("requests hand over host buffers (PCIe-inclusive)" if args.host_buffers else
                                  "request values and right-hand sides resident in HBM, solutions written to HBM"),
                   "parallelism": "single GPU"},
        "p50_ms": f32["p50_ms"], "p99_ms": f32["p99_ms"], "latency": "arrival to result, open loop",
        "rel_residual": f32["max_rel_residual"], "backward_error": f32["max_backward_error"],
        "forward_error_vs_fp64": fwd, "ir_iterations": f32["max_ir_iterations"],
        "fp32ir": f32, "fp64_plain": f64, "fp64_ir": f64ir,
        "fp32ir_over_fp64_plain": f32["capacity_req_per_s"] / f64["capacity_req_per_s"],
        "fp32ir_over_fp64_ir": f32["capacity_req_per_s"] / f64ir["capacity_req_per_s"],
    }
    if not args.no_closed_loop:
        cl = closed_loop(bd, args, 1)
        line["roofline"] = cl.pop("roofline")
        line["closed_loop"] = cl
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(4, args.cpu_budget, 0, nreq=args.requests)
    line.update({"cpu_baseline": cpu, "build": bd.build_info()})
    print(json.dumps(line))


if __name__ == "__main__":
    main()
