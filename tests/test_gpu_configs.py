"""GPU parity on every benchmark configuration against the oracle's fixtures
(tests/golden/cfg_*.npz, written by tests/golden/make_configs.py from the
oracle on the same seeded inputs), through the C-ABI:
  * identical BTF permutations P, Q and merged block bounds R (each system
    analysed afresh: USERALGO_CACHE=0, since resident plans are keyed by the
    sparsity pattern and keep their pivot order);
  * for every block above the dense threshold: identical COLAMD column order
    and identical KLU-rule pivot rows over the sparse leading part (the dense
    Schur tail pivots by partial pivoting, see the tail rule in symbolic.h);
  * FP64 with FP64 refinement and FP32 factors with FP64 refinement each reach
    ||b - A x||_2 / ||b||_2 < 1e-12, the reference's stop rule (the 10^+-4
    scaled system stops on componentwise backward error <= 1e-12 instead,
    where ||r|| is dominated by rounding in the largest rows);
  * forward error against the oracle's refined solution, max_i |x_i - xo_i| /
    max_i |xo_i| over the stored rows (all rows for n <= 20000, else 4096
    fixed rows), below 1e-9: both solutions meet ||r||/||b|| < 1e-12, and the
    bound leaves 10^3 for the condition numbers of these systems (the scaled
    system is compared on its componentwise-error bound 1e-12 * cond, taken as
    1e-6 in the same norm)."""
from __future__ import annotations

import hashlib
import importlib.util
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _specs(bd):
    spec = importlib.util.spec_from_file_location("make_configs", GOLD / "make_configs.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.case_specs(bd), mod


CASES = ["cfg0", "cfg4_p11", "cfg4_p8", "cfg4_p12", "cfg4_p3", "cfg2_small", "cfg2_keep", "cfg1", "cfg3"]


def _relres(n, r, c, v, X, B):
    out = []
    for j in range(B.shape[1]):
        y = np.zeros(n)
        np.add.at(y, r, v * X[c, j])
        out.append(np.linalg.norm(B[:, j] - y) / np.linalg.norm(B[:, j]))
    return max(out)


def _componentwise(n, r, c, v, X, B):
    worst = 0.0
    for j in range(B.shape[1]):
        y = np.zeros(n)
        np.add.at(y, r, v * X[c, j])
        den = np.abs(B[:, j]).copy()
        np.add.at(den, r, np.abs(v * X[c, j]))
        worst = max(worst, float(np.max(np.abs(B[:, j] - y) / den)))
    return worst


@pytest.mark.parametrize("case", CASES)
def test_config_parity_with_oracle(gpu, monkeypatch, case):
    bd = gpu
    path = GOLD / f"cfg_{case}.npz"
    if not path.exists():
        pytest.skip(f"fixture {path.name} not generated")
    g = np.load(path, allow_pickle=False)
    specs, mod = _specs(bd)
    n, r, c, v, B = bd.generate(specs[case])
    _say(f"case: {case} n={n} nnz={len(r)} rhs={B.shape[1]}")
    assert mod.digest(r, c, v, B) == str(g["digest"]), "generator output differs from the fixture's inputs"
    scaled = case == "cfg3"
    monkeypatch.setenv("USERALGO_REFINE_ITERS", "10")
    # A fresh analysis of this system's own values: resident plans are keyed by
    # the sparsity pattern, and the scaled system shares the circuit-like one's.
    monkeypatch.setenv("USERALGO_CACHE", "0")
    if scaled:
        monkeypatch.setenv("BLOCKDAGLU_GPU_IR_STOP", "componentwise")
    s = bd.Solver()
    assert s.set_matrix(r, c, v, n)
    assert s.analyze(), s.last_error_message()
    P, Q, R = s.permutations()
    assert mod.digest(P) == str(g["P_sha"]) and mod.digest(Q) == str(g["Q_sha"]) and mod.digest(R) == str(g["R_sha"])
    sizes = np.diff(R)
    assert np.array_equal(np.nonzero(sizes > 64)[0], g["sparse"])
    for i, b in enumerate(g["sparse"]):
        q, prow, k0 = s.sparse_order(int(b), int(sizes[b]))
        assert np.array_equal(q, g[f"q{i}"]), f"block {b}: column order"
        assert np.array_equal(prow[:k0], g[f"prow{i}"][:k0]), f"block {b}: pivot rows of the sparse part"
    rows = g["x_idx"] if "x_idx" in g else np.arange(n)
    xo = g["x_sample"] if "x_sample" in g else g["x"]
    for prec in (bd.Precision.FP64, bd.Precision.FP32_IR):
        s.set_precision(prec)
        assert s.factor(), s.last_error_message()
        X = np.asfortranarray(B.copy())
        assert s.solve_multi_in_place(X, B.shape[1]), s.last_error_message()
        rep = s.report()
        rr = _relres(n, r, c, v, X, B)
        cw = _componentwise(n, r, c, v, X, B)
        fwd = float(np.max(np.abs(X[rows] - xo)) / np.max(np.abs(xo)))
        _say(f"  {prec.name}: rel_residual {rr:.2e} componentwise {cw:.2e} forward error vs oracle {fwd:.2e} "
             f"iterations {rep['ir_iterations']} fp64 fallbacks {rep['fp64_fallback_blocks']}")
        if scaled:
            assert cw <= 1e-12
            assert fwd < 1e-6
        else:
            assert rr < 1e-12
            assert fwd < 1e-9
    s.close()
