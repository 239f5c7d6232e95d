"""The unchanged C++ API on the GPU: tests/cpp/api_check.cpp, compiled against
include/blockdaglu/solver.h alone and linked to libblockdaglu.so, runs the
reference's known answers (apps/tests/main.cpp:93-204: error codes, the 3x3
duplicate system, the n = 1100 cache matrices with x = 1, the export identity,
two solvers on two threads) as a fresh child process, under the reference's
default AUTO routing (these small systems take the global-KLU route)."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.any_route]

ROOT = Path(__file__).resolve().parents[1]


def test_cpp_api_known_answers(gpu):
    print("case: C++ API program (known answers)", flush=True)
    sys.stdout.flush()
    exe = ROOT / "block-daglu_amd" / "lib" / "bdlu_api_check"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=110)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "checks passed" in r.stdout
