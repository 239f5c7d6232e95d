"""Shared fixtures. The package directory is `block-daglu_amd/` (not an importable
identifier), so it is loaded by path and registered as `blockdaglu_amd`."""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def load_package():
    if "blockdaglu_amd" in sys.modules:
        return sys.modules["blockdaglu_amd"]
    spec = importlib.util.spec_from_file_location("blockdaglu_amd", ROOT / "block-daglu_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["blockdaglu_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "any_route: the test chooses the Solver route itself (no BLOCK default)")


@pytest.fixture(autouse=True)
def _block_route(request, monkeypatch):
    """Solver defaults to AUTO routing like the reference (small systems then
    take the global-KLU route); the tests of the BLOCK route force it with
    USERALGO_MODE=block, and tests marked any_route choose for themselves."""
    if "any_route" not in request.keywords:
        monkeypatch.setenv("USERALGO_MODE", "block")


@pytest.fixture(scope="session")
def bd():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o
    return o


@pytest.fixture(scope="session")
def gpu(bd):
    if not bd.gpu_available():
        pytest.skip("no GPU visible")
    return bd


def rel_residual(n, rows, cols, vals, x, b):
    """||b - A x||_2 / ||b||_2 over COO (reference apps/benchmark/main.cpp:55-80)."""
    y = np.zeros(n)
    np.add.at(y, rows, vals * x[cols])
    return float(np.linalg.norm(b - y) / np.linalg.norm(b))


def backward_error(n, rows, cols, vals, x, b):
    y = np.zeros(n)
    np.add.at(y, rows, vals * x[cols])
    rowsum = np.zeros(n)
    np.add.at(rowsum, rows, np.abs(vals))
    return float(np.max(np.abs(b - y)) / (rowsum.max() * np.max(np.abs(x)) + np.max(np.abs(b))))


def check_btf(n, r, c, P, Q, R, nmatch, g, k, merged=False):
    """Checks a block-triangular ordering against a btf.npz case: structural
    rank, the unique fine blocks (as row and column sets; with merged=True a
    block may be a union of whole fine blocks) and the upper block triangular
    form (no entry below the block diagonal)."""
    assert nmatch == int(g[f"rank{k}"])
    assert sorted(P.tolist()) == list(range(n)) and sorted(Q.tolist()) == list(range(n))
    invP = np.empty(n, np.int64)
    invP[P] = np.arange(n)
    invQ = np.empty(n, np.int64)
    invQ[Q] = np.arange(n)
    blk = np.searchsorted(R, np.arange(n), side="right") - 1
    if f"brow{k}" in g:
        brow, bcol = g[f"brow{k}"], g[f"bcol{k}"]
        if not merged:
            assert len(R) - 1 == len(np.unique(brow))
        for b in range(len(R) - 1):
            rows = P[R[b]:R[b + 1]]
            cols = Q[R[b]:R[b + 1]]
            rid, cid = np.unique(brow[rows]), np.unique(bcol[cols])
            assert np.array_equal(rid, cid), f"block {b}: rows and columns come from different fine blocks"
            if not merged:
                assert len(rid) == 1, f"block {b} mixes fine blocks {rid}"
            assert sum(np.count_nonzero(brow == i) for i in rid) == len(rows), f"block {b} splits a fine block"
        assert np.all(blk[invP[r]] <= blk[invQ[c]]), "entry below the block diagonal"
