"""The COLAMD column order (klu_analyze with ordering = 1 inside the
reference's KLU branch, ubtf_block_solver.cpp:493-499, :1804) of the product
host and of the oracle, two separately written restatements of the published
method, against the COLAMD that SuperLU bundles (scipy.sparse.linalg.splu with
permc_spec="COLAMD"; SymmetricMode skips SuperLU's own elimination-tree
postorder, so its perm_c is COLAMD's output: order = inverse(perm_c)). The
order depends on the pattern only, so every case uses random values. The
patterns cover the cases that decide COLAMD's ties: sparse random, dense rows
and columns (killed and ordered last), 30% fill (mostly dense), banded,
power-law column degrees, and a jumbled input (unsorted, duplicated rows).
Structurally singular patterns are left out: SuperLU returns no order for them,
and the reference only orders strongly connected blocks. CPU only."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def _superlu_colamd(A):
    A = A.tocsc().astype(float)
    A.data = np.random.default_rng(A.nnz).normal(size=A.nnz)
    pc = sla.splu(A, permc_spec="COLAMD", options=dict(SymmetricMode=True)).perm_c
    order = np.empty(len(pc), np.int64)
    order[pc] = np.arange(len(pc))
    return order


def _cases():
    rng = np.random.default_rng(20)
    out = []
    for t in range(25):
        n = int(rng.integers(5, 300))
        out.append(("random", sp.random(n, n, density=min(float(rng.uniform(1, 8)) / n, 0.9), random_state=t,
                                        format="csc") + sp.eye(n)))
    for t in range(10):
        n = int(rng.integers(60, 400))
        A = sp.random(n, n, density=3.0 / n, random_state=50 + t, format="lil")
        for _ in range(2):
            A[int(rng.integers(n)), :] = 1.0
            A[:, int(rng.integers(n))] = 1.0
        out.append(("dense rows and columns", A.tocsc() + sp.eye(n)))
    for t in range(5):
        n = int(rng.integers(64, 300))
        out.append(("30% fill", sp.random(n, n, density=0.3, random_state=80 + t, format="csc") + sp.eye(n)))
    for t in range(4):
        n = int(rng.integers(100, 800))
        out.append(("banded", sp.diags([np.ones(n - k) for k in (0, 1, 2, 5)], [0, 1, 2, 5], format="csc")
                    + sp.diags(np.ones(n - 3), -3)))
    for t in range(6):
        n = int(rng.integers(300, 3000))
        r = rng.integers(0, n, 6 * n)
        c = np.minimum((rng.pareto(1.0, 6 * n) * 3).astype(int), n - 1)
        out.append(("power-law", sp.csc_matrix((np.ones(6 * n), (r, c)), shape=(n, n)) + sp.eye(n)))
    for name, A in out:
        A = A.tocsc()
        A.sum_duplicates()
        A.sort_indices()
        yield name, A


CASES = list(_cases())


@pytest.mark.parametrize("which", ["product", "oracle"])
def test_colamd_equals_superlu_colamd(bd, orc, which):
    order_of = bd.colamd_order if which == "product" else orc.colamd
    bad = []
    for k, (name, A) in enumerate(CASES):
        n = A.shape[0]
        if not np.array_equal(order_of(n, A.indptr, A.indices), _superlu_colamd(A)):
            bad.append((k, name, n))
    assert not bad, f"{which} COLAMD differs from SuperLU's on {bad}"


def test_colamd_jumbled_input_orders_as_clean(bd, orc):
    A = (sp.random(200, 200, density=0.03, random_state=7, format="csc") + sp.eye(200)).tocsc()
    A.sort_indices()
    ip, ix = A.indptr, A.indices
    cols = [ix[ip[j]:ip[j + 1]] for j in range(200)]
    jix = np.concatenate([np.concatenate([c[::-1], c[:1]]) for c in cols])
    jip = np.concatenate([[0], np.cumsum([len(c) + 1 for c in cols])])
    ref = _superlu_colamd(A)
    assert np.array_equal(bd.colamd_order(200, jip, jix), ref)
    assert np.array_equal(orc.colamd(200, jip, jix), ref)
