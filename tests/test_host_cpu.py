"""Product host analysis (canonicalize, BTF, merge, decomposition, per-block
ordering and first pivot sequence) against the oracle, bit for bit, and against
the independent fixtures; the seeded workload generator. CPU only."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from conftest import check_btf

GOLD = Path(__file__).resolve().parent / "golden"


def test_host_matches_oracle_config0(bd, orc):
    n, r, c, v, _ = bd.generate(0)
    h = bd.HostAnalysis(r, c, v)
    assert h.error == bd.SolverError.None_
    o = orc.OracleSolver(r, c, v)
    assert o.analyze() == 0 and o.factor() == 0
    P, Q, R = h.permutations()
    Po, Qo, Ro = o.permutations()
    np.testing.assert_array_equal(P, Po)
    np.testing.assert_array_equal(Q, Qo)
    np.testing.assert_array_equal(R, Ro)
    info = o.info()
    assert (h.nblocks_raw, h.nblocks, h.nmatch) == (info["nblocks_raw"], info["nblocks"], info["nmatch"])
    assert h.signatures() == o.signatures()
    kinds, sizes = h.block_kinds()
    np.testing.assert_array_equal(sizes, np.diff(R))
    np.testing.assert_array_equal(kinds, (sizes > 64).astype(np.int32))  # dense threshold ubtf:1441-1450
    sparse = np.nonzero(kinds)[0]
    assert len(sparse) > 0
    for b in sparse:
        st = h.sparse_stats(int(b))
        q, prow = h.sparse_order(int(b))
        qo, prowo = o.sparse_order(int(b), st["s"])
        assert qo is not None
        np.testing.assert_array_equal(q, qo)                 # column order
        np.testing.assert_array_equal(prow[: st["k0"]], prowo[: st["k0"]])  # KLU-rule pivot rows of the sparse part
        assert st["k0"] + st["m"] == st["s"]


def test_host_btf_matches_fixtures(bd, orc):
    g = np.load(GOLD / "btf.npz")
    for k in range(int(g["ncases"])):
        n, r, c = int(g[f"n{k}"]), g[f"r{k}"], g[f"c{k}"]
        v = np.ones(len(r))
        h = bd.HostAnalysis(r, c, v, n)
        assert h.error == bd.SolverError.None_, (k, h.error)
        P, Q, R = h.permutations()
        check_btf(n, r, c, P, Q, R, h.nmatch, g, k, merged=True)
        o = orc.OracleSolver(r, c, v, n)
        assert o.analyze() == 0
        Po, Qo, Ro = o.permutations()
        assert np.array_equal(P, Po) and np.array_equal(Q, Qo) and np.array_equal(R, Ro)


def test_host_canonicalization_matches_exact(bd):
    g = np.load(GOLD / "canon.npz")
    for k in np.unique(g["case"]):
        sel = g["case"] == k
        h = bd.HostAnalysis(g["rows"][sel], g["cols"][sel], g["vals"][sel])
        es = g["exp_case"] == k
        assert h.nnz == int(np.count_nonzero(es))


def test_generator_is_deterministic(bd):
    a = bd.generate(0)
    b = bd.generate(0)
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(x, y)
    p1 = bd.perturb(a[3], seed=7)
    p2 = bd.perturb(a[3], seed=7)
    np.testing.assert_array_equal(p1, p2)
    ratio = p1 / a[3]
    assert np.all(ratio >= 0.95 - 1e-12) and np.all(ratio <= 1.05 + 1e-12)
    assert not np.array_equal(p1, bd.perturb(a[3], seed=8))


def test_generator_config_sizes(bd):
    n0, r0, *_ = bd.generate(0)
    assert n0 == 5000 and 25_000 < len(r0) < 40_000
    n1, r1, c1, v1, B1 = bd.generate(1)
    assert n1 == 200_000 and len(r1) == 1_288_516 and B1.shape == (n1, 1)
    n3, r3, c3, v3, _ = bd.generate(3)
    np.testing.assert_array_equal(r3, r1)  # config 3: config 1 pattern, scaled values
    np.testing.assert_array_equal(c3, c1)
    assert not np.array_equal(v3, v1)
    for p in (0, 7, 15):
        n4, r4, *_rest = bd.generate(4, pattern=p)
        assert 1_000 <= n4 <= 100_000 and 4 * n4 < len(r4) < 12 * n4
