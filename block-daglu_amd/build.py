"""Build recipe for the MI355X BLOCK-route engine.

Produces, in-tree:
  block-daglu_amd/lib/libbdlu_gpu.so     HIP kernels + device C-ABI (hipcc, gfx950)
  block-daglu_amd/lib/libblockdaglu.so   host C++ (Solver, analysis, planner, C-ABI)
  block-daglu_amd/lib/bdlu_api_check     tests/cpp/api_check.cpp against the public C++ API
  block-daglu_amd/lib/bdlu_plugin_bench  tests/cpp/plugin_bench.cpp, loads the plugin exports at run time
  oracle/liboracle.so                    the oracle (test infrastructure; built by
                                         oracle/Makefile, only when asked)
Usage: python block-daglu_amd/build.py [--oracle] [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("BLOCKDAGLU_OFFLOAD_ARCH", "gfx950")

HIP_SRCS = ["csrc/hip/kernels.hip", "csrc/hip/engine.hip"]
HOST_SRCS = [
    "csrc/host/canon.cpp", "csrc/host/btf.cpp", "csrc/host/decomp.cpp", "csrc/host/colamd.cpp",
    "csrc/host/symbolic.cpp", "csrc/host/denselu.cpp", "csrc/host/plan.cpp", "csrc/host/solver.cpp",
    "csrc/host/capi.cpp", "csrc/host/gen.cpp", "csrc/host/export.cpp", "csrc/host/serve.cpp",
    "csrc/host/plugin.cpp", "csrc/host/route.cpp",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    # no packed FP32 instructions: some of their forms have produced wrong results on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}",
              f"-I{PKG / 'csrc/host'}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")


def source_commit() -> str:
    """Commit of the sources being built (with "+dirty" for uncommitted changes),
    recorded inside libbdlu_gpu.so and returned by bdlu_gpu_build_info()."""
    try:
        head = subprocess.run(["git", "-C", str(ROOT), "rev-parse", "--short=12", "HEAD"], capture_output=True,
                              text=True, check=True).stdout.strip()
        dirty = subprocess.run(["git", "-C", str(ROOT), "status", "--porcelain", "--untracked-files=no"],
                               capture_output=True, text=True, check=True).stdout.strip()
        return head + ("+dirty" if dirty else "")
    except (OSError, subprocess.CalledProcessError):
        return "unknown"


def build(oracle: bool = False, jobs: int = 8) -> None:
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    steps = []
    stamp = f'-DBDLU_BUILD_INFO="{source_commit()} {ARCH}"'
    for src in HIP_SRCS:
        obj = OBJ / (Path(src).stem + ".hip.o")
        steps.append([HIPCC, *HIP_FLAGS, stamp, "-c", str(PKG / src), "-o", str(obj)])
    for src in HOST_SRCS:
        obj = OBJ / (Path(src).stem + ".o")
        steps.append([CXX, *HOST_FLAGS, "-c", str(PKG / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, steps))
    gpu = LIB / "libbdlu_gpu.so"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(gpu) + ".tmp",
          *[str(OBJ / (Path(s).stem + ".hip.o")) for s in HIP_SRCS]])
    os.replace(str(gpu) + ".tmp", gpu)
    host = LIB / "libblockdaglu.so"
    _run([CXX, "-shared", "-fPIC", "-o", str(host) + ".tmp",
          *[str(OBJ / (Path(s).stem + ".o")) for s in HOST_SRCS],
          f"-L{LIB}", "-lbdlu_gpu", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(str(host) + ".tmp", host)
    # A program written only against the public C++ header, linked like a
    # caller of the reference library (the drop-in check the tests run).
    check = LIB / "bdlu_api_check"
    _run([CXX, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "api_check.cpp"),
          "-o", str(check) + ".tmp", f"-L{LIB}", "-lblockdaglu", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(str(check) + ".tmp", check)
    # The plugin benchmark loads libblockdaglu.so at run time (no link to it).
    pbench = LIB / "bdlu_plugin_bench"
    _run([CXX, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "plugin_bench.cpp"),
          "-o", str(pbench) + ".tmp", "-ldl"])
    os.replace(str(pbench) + ".tmp", pbench)
    if oracle:
        _run(["make", "-s", "-C", str(ROOT / "oracle")])
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    build(oracle=a.oracle, jobs=a.jobs)
    print("built", LIB)
