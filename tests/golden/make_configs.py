"""Writes tests/golden/cfg_<case>.npz: the oracle's results on the seeded
workloads of every benchmark configuration, for the GPU parity tests.

    python tests/golden/make_configs.py [case ...]    (default: every case)

The oracle (oracle/, the CPU restatement of the reference BLOCK route: BTF,
merge, dense LU, COLAMD + KLU-rule left-looking LU, block solve, FP64
refinement) runs here once per case; the GPU tests regenerate the same inputs
with the seeded generator and compare. Each file holds (npz, allow_pickle=False):
  digest        sha256 of the generated inputs (rows, cols, values, right-hand
                sides), so a changed generator is caught before any comparison
  n, nnz, nrhs
  P, Q, R       the BTF row/column permutations and merged block bounds
                (full arrays when n <= 20000, else only their sha256 digests)
  P_sha, Q_sha, R_sha
  sparse        block ids of the blocks above the dense threshold; q<i>, prow<i>
                their COLAMD column order and KLU-rule pivot rows (local indices)
  x             the oracle's solution (refined to ||r||/||b|| < 1e-12), full when
                n <= 20000; else x_idx (4096 fixed row indices) and x_sample
  x_norm        2-norm of each solution column
Cases: the 5,000-unknown system (config 0); the circuit-like n = 200,000 system
(config 1) and its 10^+-4 row/column scaled twin (config 3); the batched-blocks
system (config 2) at full size in generated order, and reduced (4,000 small
blocks + 3 dense-ish blocks of 512-900) with the structure-hiding permutation;
four serving patterns (config 4) across the n range.
"""
from __future__ import annotations

import hashlib
import importlib.util
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

FULL_MAX = 20000
NSAMPLE = 4096


def _package():
    spec = importlib.util.spec_from_file_location("blockdaglu_amd", ROOT / "block-daglu_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["blockdaglu_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def case_specs(bd):
    """name -> generator spec (shared with tests/test_gpu_configs.py)."""
    out = {}
    out["cfg0"] = bd.config_spec(0)
    out["cfg1"] = bd.config_spec(1)
    out["cfg3"] = bd.config_spec(3)
    sp = bd.config_spec(2)
    sp.keep_order = 1
    out["cfg2_keep"] = sp
    sp = bd.config_spec(2)
    sp.n_small_blocks, sp.n_large_blocks, sp.large_size_lo, sp.large_size_hi = 4000, 3, 512, 900
    out["cfg2_small"] = sp
    for p in (11, 8, 12, 3):
        out[f"cfg4_p{p}"] = bd.config_spec(4, p)
    return out


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def sample_rows(n: int) -> np.ndarray:
    return np.sort(np.random.default_rng(12345).choice(n, min(n, NSAMPLE), replace=False)).astype(np.int64)


def make(bd, orc, name, spec):
    t0 = time.time()
    n, r, c, v, B = bd.generate(spec)
    o = orc.OracleSolver(r, c, v, n)
    assert o.analyze() == 0
    t1 = time.time()
    assert o.factor(os.cpu_count() or 1) == 0
    t2 = time.time()
    P, Q, R = o.permutations()
    out = {"digest": np.array(digest(r, c, v, B)), "n": np.array(n), "nnz": np.array(len(r)),
           "nrhs": np.array(B.shape[1]), "P_sha": np.array(digest(P)), "Q_sha": np.array(digest(Q)),
           "R_sha": np.array(digest(R))}
    if n <= FULL_MAX:
        out.update({"P": P, "Q": Q, "R": R})
    sizes = np.diff(R)
    sparse = np.nonzero(sizes > 64)[0]
    out["sparse"] = sparse.astype(np.int32)
    for i, b in enumerate(sparse):
        q, prow = o.sparse_order(int(b), int(sizes[b]))
        out[f"q{i}"], out[f"prow{i}"] = q, prow
    X = np.zeros_like(B)
    for j in range(B.shape[1]):
        rc, X[:, j] = o.solve(np.ascontiguousarray(B[:, j]), 3)
        assert rc == 0
    out["x_norm"] = np.linalg.norm(X, axis=0)
    if n <= FULL_MAX:
        out["x"] = X
    else:
        idx = sample_rows(n)
        out["x_idx"], out["x_sample"] = idx, X[idx]
    np.savez_compressed(HERE / f"cfg_{name}.npz", **out)
    print(f"{name}: n={n} nnz={len(r)} blocks={len(R) - 1} sparse={len(sparse)} analyze {t1 - t0:.1f}s "
          f"factor {t2 - t1:.1f}s total {time.time() - t0:.1f}s", flush=True)


def main(names):
    bd = _package()
    from oracle import oracle as orc
    specs = case_specs(bd)
    order = names or ["cfg0", "cfg4_p11", "cfg4_p8", "cfg4_p12", "cfg4_p3", "cfg2_small", "cfg2_keep", "cfg1", "cfg3"]
    for name in order:
        make(bd, orc, name, specs[name])


if __name__ == "__main__":
    main(sys.argv[1:])
