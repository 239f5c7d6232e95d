"""The competition plugin on the GPU: bench.py --plugin-evidence writes seeded
cases, times the oracle as the baseline, and runs bdlu_plugin_bench, which
loads libblockdaglu.so through CreateAlgorithm and drives prepareData ->
rearrangeMatrix -> decompositionLU -> solveLinearSystem. The evidence files
must carry the reference's columns (apps/benchmark/main.cpp:309-432) and pass
its gates: every run ok, and each case's error against x* no larger than the
baseline's."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

HEADERS = {
    "runs.csv": "alg,case,repetition,strategy,n,nnz,reorder_ms,lu_ms,solve_ms,total_ms,rel_err,rel_residual,ok",
    "margin.csv": "case,n,nnz,xstar_rel_residual,ss_rel_residual,ua_rel_residual,eps_suitesparse,err_user,"
                  "margin_eps_minus_user,pass_user_le_eps",
    "timing_summary.csv": "alg,case,strategy,n,nnz,repetitions,median_total_ms,p95_total_ms,median_rel_err,"
                          "median_rel_residual,ok",
}


def test_plugin_evidence(gpu, tmp_path):
    print("case: plugin evidence (4 cases, 2 repetitions)", flush=True)
    sys.stdout.flush()
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--plugin-evidence", str(tmp_path),
                        "--plugin-reps", "2"], capture_output=True, text=True, timeout=110)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    for name, header in HEADERS.items():
        assert (tmp_path / name).read_text().splitlines()[0] == header
    margin = (tmp_path / "margin.csv").read_text().splitlines()[1:]
    assert len(margin) == 4 and all(line.endswith(",1") for line in margin)
    summary = dict(line.split(",") for line in (tmp_path / "summary.txt").read_text().splitlines())
    assert summary["BlockDAGLU_all_ok"] == "1" and summary["CPUReference_all_ok"] == "1"
    assert float(summary["speedup_suitesparse_over_blockdaglu"]) > 0
