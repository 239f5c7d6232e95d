"""Writes the golden fixtures under tests/golden/ from independent implementations
(numpy/scipy and exact Python arithmetic), never from the oracle or the product.

    python tests/golden/make_golden.py

Fixtures (npz, allow_pickle=False):
  canon.npz  - COO inputs with duplicates and explicit zeros; expected canonical
               (row, col, value) triplets: sorted by (row, col), duplicates summed
               exactly (math.fsum of small integers, so every summation order
               agrees), exact zeros dropped.
  btf.npz    - square sparse matrices with a hidden block-triangular structure
               (random row/column permutations); expected structural rank
               (scipy structural_rank) and the fine block partition: for a
               structurally nonsingular matrix the strongly connected
               components of the matched column digraph are unique as
               (row set, column set) pairs whatever the matching.
  dense.npz  - dense matrices 1..64; expected LAPACK getrf pivots (first largest
               |a_ik| wins ties, the same rule as the reference's dense LU) and
               the LU values.
  solve.npz  - nonsingular sparse systems with hidden block structure; expected
               x from scipy's sparse LU (SuperLU), in float64.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components, maximum_bipartite_matching, structural_rank
from scipy.sparse.linalg import spsolve

OUT = Path(__file__).resolve().parent


def canon_cases(rng):
    rows, cols, vals, case = [], [], [], []
    exp_r, exp_c, exp_v, exp_case = [], [], [], []
    for k in range(6):
        n = int(rng.integers(3, 40))
        m = int(rng.integers(n, 6 * n))
        r = rng.integers(0, n, m)
        c = rng.integers(0, n, m)
        v = rng.integers(-4, 5, m).astype(np.float64)  # small integers: exact sums
        # force some cancelling duplicates and explicit zeros
        r = np.concatenate([r, r[:3], [0]])
        c = np.concatenate([c, c[:3], [n - 1]])
        v = np.concatenate([v, -v[:3], [0.0]])
        acc: dict[tuple[int, int], list[float]] = {}
        for a, b, x in zip(r.tolist(), c.tolist(), v.tolist()):
            acc.setdefault((a, b), []).append(x)
        for (a, b) in sorted(acc):
            s = math.fsum(acc[(a, b)])
            if s != 0.0:
                exp_r.append(a)
                exp_c.append(b)
                exp_v.append(s)
                exp_case.append(k)
        rows.append(r)
        cols.append(c)
        vals.append(v)
        case.append(np.full(len(r), k))
    np.savez(OUT / "canon.npz", rows=np.concatenate(rows).astype(np.int32), cols=np.concatenate(cols).astype(np.int32),
             vals=np.concatenate(vals), case=np.concatenate(case).astype(np.int32),
             exp_rows=np.array(exp_r, np.int32), exp_cols=np.array(exp_c, np.int32), exp_vals=np.array(exp_v),
             exp_case=np.array(exp_case, np.int32))


def hidden_btf(rng, n, max_block, coupling, singular_rows=0):
    """Block upper triangular pattern (each block strongly connected via a cycle
    plus random entries), then random row/column permutations."""
    sizes = []
    left = n
    while left > 0:
        s = int(min(left, rng.integers(1, max_block + 1)))
        sizes.append(s)
        left -= s
    starts = np.cumsum([0] + sizes)
    r, c = [], []
    for b, s in enumerate(sizes):
        o = starts[b]
        for i in range(s):
            r.append(o + i)
            c.append(o + i)
            if s > 1:
                r.append(o + i)
                c.append(o + (i + 1) % s)
            for _ in range(int(rng.poisson(1.5))):
                r.append(o + i)
                c.append(o + int(rng.integers(0, s)))
        for i in range(s):  # coupling into later blocks (upper triangular)
            if starts[b + 1] < n and rng.random() < coupling:
                r.append(o + i)
                c.append(int(rng.integers(starts[b + 1], n)))
    r = np.array(r)
    c = np.array(c)
    key = np.unique(r * n + c)
    r, c = key // n, key % n
    if singular_rows:
        drop = rng.choice(n, singular_rows, replace=False)
        keep = ~np.isin(r, drop)
        r, c = r[keep], c[keep]
    pr = rng.permutation(n)
    pc = rng.permutation(n)
    return pr[r], pc[c]


def fine_blocks(n, r, c):
    """(row set, column set) of each fine block, from scipy's matching and SCCs."""
    A = sp.csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    match = maximum_bipartite_matching(A, perm_type="column")  # match[i] = column of row i
    row_of_col = np.full(n, -1)
    row_of_col[match] = np.arange(n)
    # column digraph: j -> k when the row matched to column j has an entry in column k
    jr = match[r]
    G = sp.csr_matrix((np.ones(len(r)), (jr, c)), shape=(n, n))
    ncomp, lab = connected_components(G, directed=True, connection="strong")
    blocks = []
    for k in range(ncomp):
        cs = np.sort(np.nonzero(lab == k)[0])
        rs = np.sort(row_of_col[cs])
        blocks.append((rs, cs))
    return blocks


def btf_cases(rng):
    out = {}
    specs = [(30, 5, 0.3, 0), (120, 12, 0.2, 0), (300, 40, 0.1, 0), (200, 1, 0.5, 0), (80, 80, 0.0, 0),
             (150, 20, 0.3, 3)]
    for k, (n, mb, cp, sing) in enumerate(specs):
        r, c = hidden_btf(rng, n, mb, cp, sing)
        A = sp.csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
        out[f"r{k}"] = r.astype(np.int32)
        out[f"c{k}"] = c.astype(np.int32)
        out[f"n{k}"] = np.int32(n)
        out[f"rank{k}"] = np.int32(structural_rank(A))
        if structural_rank(A) == n:
            bl = fine_blocks(n, r, c)
            # block id per row and per column (ids in an arbitrary order)
            brow = np.zeros(n, np.int32)
            bcol = np.zeros(n, np.int32)
            for i, (rs, cs) in enumerate(bl):
                brow[rs] = i
                bcol[cs] = i
            out[f"brow{k}"] = brow
            out[f"bcol{k}"] = bcol
    out["ncases"] = np.int32(len(specs))
    np.savez(OUT / "btf.npz", **out)


def dense_cases(rng):
    out = {}
    sizes = [1, 2, 3, 5, 8, 17, 32, 40, 63, 64]
    for k, s in enumerate(sizes):
        a = rng.standard_normal((s, s))
        if k % 3 == 0:
            a += np.diag(rng.standard_normal(s) * 0.01)  # weak diagonal: forces row swaps
        lu, piv = sla.lu_factor(a)
        out[f"a{k}"] = a
        out[f"lu{k}"] = lu
        out[f"piv{k}"] = piv.astype(np.int32)  # 0-based LAPACK swap sequence
    out["ncases"] = np.int32(len(sizes))
    np.savez(OUT / "dense.npz", **out)


def solve_cases(rng):
    out = {}
    specs = [(200, 30, 0.3), (600, 90, 0.2), (400, 1, 0.5)]
    for k, (n, mb, cp) in enumerate(specs):
        r, c = hidden_btf(rng, n, mb, cp)
        v = rng.standard_normal(len(r))
        A = sp.csr_matrix((v, (r, c)), shape=(n, n)).tolil()
        # diagonal dominance on the matched entries keeps the systems well posed
        d = np.abs(A).sum(axis=1).A1
        A = A.tocsr()
        match = maximum_bipartite_matching(sp.csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n)),
                                           perm_type="column")
        for i in range(n):
            A[i, match[i]] = (0.6 * d[i] + 0.5) * (1 if rng.random() < 0.5 else -1)
        A = A.tocoo()
        b = rng.uniform(-1, 1, n)
        x = spsolve(A.tocsc(), b)
        out[f"r{k}"] = A.row.astype(np.int32)
        out[f"c{k}"] = A.col.astype(np.int32)
        out[f"v{k}"] = A.data
        out[f"b{k}"] = b
        out[f"x{k}"] = x
    out["ncases"] = np.int32(len(specs))
    np.savez(OUT / "solve.npz", **out)


def main():
    rng = np.random.default_rng(20261203)
    canon_cases(rng)
    btf_cases(rng)
    dense_cases(rng)
    solve_cases(rng)
    print("wrote", sorted(p.name for p in OUT.glob("*.npz")))


if __name__ == "__main__":
    main()
