"""Route selection and the global-KLU route's host analysis, on the CPU.

Routing (reference src/solver.cpp:432-565, :820-860): the product's decision
(blockdaglu_amd.route_for, the same function Solver.analyze calls) is checked
against known answers derived by hand from the reference's rules at their
boundaries (n = 50,000 and 65,000; nnz = 600,000 and 3,000,000; the dense-hub
threshold max(4096, n/8); the env overrides), so an off-by-one in a comparison
or a swapped branch fails a case. The oracle's Python restatement of the same
rules (oracle/oracle.py) must agree with every case as well.

Global-KLU analysis (src/solver.cpp:878-968 through klu_analyze): the product's
host plan for KLU's settings against the oracle's separate restatement
(oracle analyze_klu): identical BTF permutations composed with each small
block's COLAMD order, identical block bounds, and for every block above 64 the
same COLAMD column order and the same pivot rows over its sparse leading part.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.any_route


def _diag_plus(n, extra_per_row=0, seed=0):
    """n x n: full diagonal plus `extra_per_row` random off-diagonal entries per
    row (so nnz = n * (1 + extra_per_row) before duplicates merge)."""
    rng = np.random.default_rng(seed)
    r = [np.arange(n)]
    c = [np.arange(n)]
    for _ in range(extra_per_row):
        rr = np.arange(n)
        cc = (rr + 1 + rng.integers(0, n - 1, n)) % n
        r.append(rr)
        c.append(cc)
    r = np.concatenate(r).astype(np.int32)
    c = np.concatenate(c).astype(np.int32)
    v = rng.standard_normal(len(r))
    v[: n] = 10.0
    return r, c, v


def _exact_nnz(n, nnz, seed=0):
    """Diagonal plus distinct off-diagonal entries, exactly nnz in all."""
    rng = np.random.default_rng(seed)
    extra = nnz - n
    keys = set()
    rows, cols = [], []
    while len(rows) < extra:
        k = min(2 * (extra - len(rows)) + 16, 4 * extra + 16)
        rr = rng.integers(0, n, k)
        cc = rng.integers(0, n, k)
        for a, b in zip(rr.tolist(), cc.tolist()):
            if a == b or (a, b) in keys:
                continue
            keys.add((a, b))
            rows.append(a)
            cols.append(b)
            if len(rows) == extra:
                break
    r = np.concatenate([np.arange(n), np.array(rows, np.int64)]).astype(np.int32)
    c = np.concatenate([np.arange(n), np.array(cols, np.int64)]).astype(np.int32)
    v = np.ones(len(r))
    return r, c, v


def _oracle_route(orc, n, r, c):
    nnz = len(r)
    order = np.lexsort((c, r))
    klu = not orc.auto_choose_block(n, nnz, lambda: orc.profile_lite(n, r[order], c[order]))
    btf = orc.klu_btf_for(n, nnz)
    return {"klu": klu, "btf": btf, "scale": orc.klu_scale_for(n, nnz, btf), "tol": orc.klu_tol_for(n, nnz, btf)}


@pytest.fixture
def clean_env(monkeypatch):
    for k in ("USERALGO_MODE", "BLOCKDAGLU__MODE", "USERALGO_AUTO_BLOCK_N", "USERALGO_AUTO_BLOCK_NNZ",
              "USERALGO_KLU_BTF", "BLOCKDAGLU__KLU_BTF", "USERALGO_KLU_SCALE", "USERALGO_KLU_TOL",
              "USERALGO_KLU_ORDERING", "USERALGO_CACHE"):
        monkeypatch.delenv(k, raising=False)
    return monkeypatch


# (n, nnz, hub_row, expected route, btf, scale, tol)
CASES = [
    (5000, 31000, 0, "klu", 1, 0, 0.001),       # n < 50000, nnz < 2M: KLU; n <= 10000: btf
    (10000, 20000, 0, "klu", 1, 0, 0.001),      # btf boundary n <= 10000
    (10001, 20002, 0, "klu", 0, 0, 0.001),
    (49999, 99998, 0, "klu", 0, 0, 0.001),      # just below the AUTO_BLOCK_N threshold
    (50000, 100000, 0, "block", 0, 0, 0.001),   # at the threshold; profile clean
    (20000, 600000, 0, "block", 1, 0, 0.001),   # n <= 25000 and nnz >= 600000 (btf also 1 there)
    (20000, 599999, 0, "klu", 0, 0, 0.001),
    (40000, 600000, 0, "klu", 1, 2, 0.001),     # 35000..50000 with 500k..800k: btf 1, scale 2
    (65000, 3000000, 0, "klu", 0, 2, 0.001),    # n >= 65000 and nnz >= 3M: global KLU, max scaling
    (64999, 3000000, 0, "block", 0, 0, 0.001),  # nnz >= 2M: block; profile clean
    (50000, 100000, 6250, "klu", 0, 0, 0.001),  # a dense hub row of n/8 entries -> global KLU
    (50000, 100000, 6249, "block", 0, 0, 0.001),
]


@pytest.mark.parametrize("n,nnz,hub,route,btf,scale,tol", CASES)
def test_route_known_answers(bd, orc, clean_env, n, nnz, hub, route, btf, scale, tol):
    r, c, v = _exact_nnz(n, nnz, seed=n)
    if hub:
        # Fill one row up to exactly `hub` entries (new distinct columns) and
        # drop as many off-diagonal entries of other rows, so nnz is unchanged.
        row = n // 3
        used = set(c[r == row].tolist())
        add = hub - len(used)
        cols = [j for j in range(n) if j not in used][:add]
        keep = np.ones(len(r), bool)
        keep[np.nonzero((r != row) & (r != c))[0][:add]] = False
        r = np.concatenate([r[keep], np.full(add, row, np.int32)])
        c = np.concatenate([c[keep], np.array(cols, np.int32)])
        v = np.concatenate([v[keep], np.ones(add)])
        assert len(r) == nnz and int(np.count_nonzero(r == row)) == hub
    got = bd.route_for(r, c, v, n)
    assert got["klu"] == (route == "klu"), got
    assert (got["btf"], got["scale"], got["ordering"]) == (btf, scale, 1)
    assert got["tol"] == tol
    ref = _oracle_route(orc, n, r, c)
    assert ref["klu"] == got["klu"] and ref["btf"] == btf and ref["scale"] == scale and ref["tol"] == tol


def test_route_env_overrides(bd, clean_env):
    r, c, v = _diag_plus(5000, 2)
    assert bd.route_for(r, c, v)["klu"]
    clean_env.setenv("USERALGO_MODE", "BLOCK")  # case-insensitive (src/solver.cpp:825)
    assert not bd.route_for(r, c, v)["klu"]
    clean_env.setenv("USERALGO_MODE", "klu")
    assert bd.route_for(r, c, v)["klu"]
    clean_env.setenv("USERALGO_MODE", "nonsense")  # unknown -> AUTO
    assert bd.route_for(r, c, v)["klu"]
    clean_env.delenv("USERALGO_MODE")
    # The reference's alias rule reads BLOCKDAGLU__X (two underscores) before USERALGO_X.
    clean_env.setenv("BLOCKDAGLU__MODE", "block")
    clean_env.setenv("USERALGO_MODE", "klu")
    assert not bd.route_for(r, c, v)["klu"]
    clean_env.delenv("BLOCKDAGLU__MODE")
    clean_env.delenv("USERALGO_MODE")
    clean_env.setenv("USERALGO_AUTO_BLOCK_N", "4000")
    assert not bd.route_for(r, c, v)["klu"]
    clean_env.setenv("USERALGO_KLU_BTF", "0")
    clean_env.setenv("USERALGO_KLU_SCALE", "1")
    clean_env.setenv("USERALGO_KLU_TOL", "0.1")
    got = bd.route_for(r, c, v)
    assert (got["btf"], got["scale"], got["tol"]) == (0, 1, 0.1)


def _compare_klu_analysis(bd, orc, n, r, c, v, btf, scale, tol=0.001):
    h = bd.HostAnalysis(r, c, v, n, klu=(btf, scale, tol))
    assert h.error == 0
    o = orc.OracleSolver(r, c, v, n)
    assert o.analyze_klu(btf, scale, tol) == 0
    assert o.factor() == 0
    P, Q, R = h.permutations()
    Po, Qo, Ro = o.klu_permutations()
    assert np.array_equal(R, Ro)
    assert np.array_equal(P, Po) and np.array_equal(Q, Qo)
    kinds, sizes = h.block_kinds()
    assert np.array_equal(sizes, np.diff(Ro))
    assert np.array_equal(kinds, (sizes > 64).astype(np.int32))
    nsparse = 0
    for b in np.nonzero(sizes > 64)[0]:
        st = h.sparse_stats(int(b))
        q, prow = h.sparse_order(int(b))
        qo, prowo = o.sparse_order(int(b), int(sizes[b]))
        assert np.array_equal(q, qo), f"block {b}: column order"
        k0 = st["k0"]
        assert np.array_equal(prow[:k0], prowo[:k0]), f"block {b}: pivot rows of the sparse part"
        nsparse += 1
    return nsparse


def test_klu_analysis_matches_oracle_btf(bd, orc):
    n, r, c, v, _ = bd.generate(0)
    assert _compare_klu_analysis(bd, orc, n, r, c, v, btf=1, scale=0) > 0


@pytest.mark.parametrize("scale", [1, 2])
def test_klu_analysis_matches_oracle_scaled(bd, orc, scale):
    n, r, c, v, _ = bd.generate(0)
    rng = np.random.default_rng(7)
    v = v * (10.0 ** rng.uniform(-4, 4, n))[r] * (10.0 ** rng.uniform(-4, 4, n))[c]
    _compare_klu_analysis(bd, orc, n, r, c, v, btf=1, scale=scale)


def test_klu_analysis_matches_oracle_one_block(bd, orc):
    """btf = 0: the whole matrix is one block under COLAMD."""
    spec = bd.config_spec(4, 11)
    n, r, c, v, _ = bd.generate(spec)
    assert n < 5000
    assert _compare_klu_analysis(bd, orc, n, r, c, v, btf=0, scale=0) == 1
