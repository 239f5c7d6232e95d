"""ORACLE wrapper — test infrastructure, not product code.

ctypes view of oracle/liboracle.so, the CPU restatement of the reference BLOCK
route (see oracle/oracle.cpp for what each part restates). Imported only by
tests/, __graft_entry__.smoke() (as the checker) and bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        sig = {
            "oracle_create": (_P, [C.c_int64, _I32P, _I32P, _F64P, C.c_int32]),
            "oracle_free": (None, [_P]),
            "oracle_analyze": (C.c_int, [_P]),
            "oracle_analyze_klu": (C.c_int, [_P, C.c_int, C.c_int, C.c_double]),
            "oracle_factor": (C.c_int, [_P, C.c_int]),
            "oracle_solve": (C.c_int, [_P, _F64P, C.c_int]),
            "oracle_set_values": (None, [_P, _F64P]),
            "oracle_n": (C.c_int32, [_P]),
            "oracle_nnz": (C.c_int64, [_P]),
            "oracle_canonical": (None, [_P, _I32P, _I32P, _F64P]),
            "oracle_info": (None, [_P, _I32P, _I32P, _I32P, _I32P]),
            "oracle_perm": (None, [_P, _I32P, _I32P, _I32P]),
            "oracle_dense_factor": (C.c_int32, [_P, C.c_int32, _F64P, _I32P]),
            "oracle_sparse_order": (C.c_int32, [_P, C.c_int32, _I32P, _I32P]),
            "oracle_timings": (None, [_P, _F64P]),
            "oracle_signatures": (None, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64)]),
            "oracle_btf_order": (C.c_int32, [C.c_int32, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P]),
            "oracle_merge": (C.c_int32, [_I32P, C.c_int32, C.c_int32, C.c_int32, _I32P]),
            "oracle_colamd": (None, [C.c_int32, _I32P, _I32P, _I32P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class OracleSolver:
    """Reference BLOCK route on the CPU: set_matrix/analyze/factor/solve."""

    def __init__(self, rows, cols, vals, dimension=-1):
        self.L = lib()
        r = np.ascontiguousarray(rows, np.int32)
        c = np.ascontiguousarray(cols, np.int32)
        v = np.ascontiguousarray(vals, np.float64)
        self.h = self.L.oracle_create(len(r), _p(r, _I32P), _p(c, _I32P), _p(v, _F64P), dimension)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_free(self.h)
            self.h = None

    def analyze(self) -> int:
        return self.L.oracle_analyze(self.h)

    def analyze_klu(self, btf: int, scale: int, tol: float = 0.001) -> int:
        """The global-KLU route's analysis with KLU's settings (every block by the
        KLU rule, whole-row scaling for scale 1 / 2)."""
        return self.L.oracle_analyze_klu(self.h, btf, scale, tol)

    def klu_permutations(self, small_max: int = 64):
        """After analyze_klu + factor: the row/column permutations with each
        block's COLAMD order applied symmetrically to the blocks of
        2..small_max (KLU's P = Pbtf[Pblk], Q = Qbtf[Pblk]), and R."""
        P, Q, R = self.permutations()
        P2, Q2 = P.copy(), Q.copy()
        for b in range(len(R) - 1):
            lo, s = int(R[b]), int(R[b + 1] - R[b])
            if 2 <= s <= small_max:
                q, _ = self.sparse_order(b, s)
                P2[lo:lo + s] = P[lo + q]
                Q2[lo:lo + s] = Q[lo + q]
        return P2, Q2, R

    def factor(self, threads: int = 0) -> int:
        return self.L.oracle_factor(self.h, threads)

    def solve(self, b, refine_iters: int = 0):
        x = np.ascontiguousarray(b, np.float64).copy()
        rc = self.L.oracle_solve(self.h, _p(x, _F64P), refine_iters)
        return rc, x

    def set_values(self, canonical_vals):
        v = np.ascontiguousarray(canonical_vals, np.float64)
        self.L.oracle_set_values(self.h, _p(v, _F64P))

    @property
    def n(self):
        return self.L.oracle_n(self.h)

    def canonical(self):
        nnz = self.L.oracle_nnz(self.h)
        r = np.zeros(nnz, np.int32)
        c = np.zeros(nnz, np.int32)
        v = np.zeros(nnz, np.float64)
        self.L.oracle_canonical(self.h, _p(r, _I32P), _p(c, _I32P), _p(v, _F64P))
        return r, c, v

    def info(self):
        a, b, c, d = (C.c_int32() for _ in range(4))
        self.L.oracle_info(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return {"nblocks_raw": a.value, "nblocks": b.value, "nmatch": c.value, "threads": d.value}

    def permutations(self):
        info = self.info()
        n = self.n
        P = np.zeros(n, np.int32)
        Q = np.zeros(n, np.int32)
        R = np.zeros(info["nblocks"] + 1, np.int32)
        self.L.oracle_perm(self.h, _p(P, _I32P), _p(Q, _I32P), _p(R, _I32P))
        return P, Q, R

    def dense_factor(self, b, size):
        lu = np.zeros(size * size, np.float64)
        piv = np.zeros(size, np.int32)
        s = self.L.oracle_dense_factor(self.h, b, _p(lu, _F64P), _p(piv, _I32P))
        return (None, None) if s < 0 else (lu.reshape(size, size), piv)

    def sparse_order(self, b, size):
        q = np.zeros(size, np.int32)
        prow = np.zeros(size, np.int32)
        s = self.L.oracle_sparse_order(self.h, b, _p(q, _I32P), _p(prow, _I32P))
        return (None, None) if s < 0 else (q, prow)

    def timings(self):
        t = np.zeros(3, np.float64)
        self.L.oracle_timings(self.h, _p(t, _F64P))
        return {"analyze": t[0], "factor": t[1], "solve": t[2]}

    def signatures(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.L.oracle_signatures(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value


def btf_order(n, Ap, Ai):
    L = lib()
    Ap = np.ascontiguousarray(Ap, np.int32)
    Ai = np.ascontiguousarray(Ai, np.int32)
    P = np.zeros(n, np.int32)
    Q = np.zeros(n, np.int32)
    R = np.zeros(n + 1, np.int32)
    nm = C.c_int32()
    nb = L.oracle_btf_order(n, _p(Ap, _I32P), _p(Ai, _I32P), _p(P, _I32P), _p(Q, _I32P), _p(R, _I32P), C.byref(nm))
    return P, Q, R[: nb + 1], nm.value


def merge(R, n, merge_max=64):
    L = lib()
    R = np.ascontiguousarray(R, np.int32)
    out = np.zeros(len(R), np.int32)
    nb = L.oracle_merge(_p(R, _I32P), len(R) - 1, n, merge_max, _p(out, _I32P))
    return out[: nb + 1]


def colamd(n, indptr, indices):
    """The oracle's COLAMD order of an n x n CSC pattern (order[k] = column at position k)."""
    ip = np.ascontiguousarray(indptr, np.int32)
    ix = np.ascontiguousarray(indices, np.int32)
    out = np.zeros(n, np.int32)
    lib().oracle_colamd(n, _p(ip, _I32P), _p(ix, _I32P), _p(out, _I32P))
    return out


# ---- route selection, restated from the reference's src/solver.cpp -------------------
# Decision rules of Solver::analyze in AUTO mode and the KLU settings of the
# global route. Pure functions of (n, nnz, sampled degree profile, env values);
# restated here so that the product's choice can be checked against them.

def profile_lite(n, rows, cols):
    """compute_profile_lite (src/solver.cpp:182-230) over canonical-order entries."""
    nnz = len(rows)
    step = max(1, nnz // 200000) if nnz > 200000 else 1
    r = np.asarray(rows)[::step]
    c = np.asarray(cols)[::step]
    rd = np.bincount(r, minlength=n)
    cd = np.bincount(c, minlength=n)
    diag = np.zeros(n, bool)
    diag[r[r == c]] = True
    k = int(0.95 * (n - 1))
    return {"row_max": int(rd.max()) * step, "col_max": int(cd.max()) * step,
            "row_p95": int(np.partition(rd, k)[k]) * step, "col_p95": int(np.partition(cd, k)[k]) * step,
            "diag_cov": float(diag.sum()) / n}


def auto_choose_block(n, nnz, profile=None, auto_n=50000, auto_nnz=2000000):
    """auto_choose_block (src/solver.cpp:450-515), without the learned override;
    profile: a callable returning profile_lite's dict (evaluated only when the
    rules need it)."""
    if n >= 65000 and nnz >= 3000000:
        coarse = False
    elif n <= 25000 and nnz >= 600000:
        coarse = True
    elif n >= auto_n or nnz >= auto_nnz:
        coarse = True
    else:
        coarse = False
    if not coarse:
        return False
    if n < 40000 and nnz < 1500000:
        return True
    p = profile()
    dense_thr = max(4096, n // 8)
    hub = p["row_max"] >= dense_thr or p["col_max"] >= dense_thr
    broad_thr = max(512, n // 64)
    broad = p["row_p95"] >= broad_thr and p["col_p95"] >= broad_thr and p["diag_cov"] >= 0.9
    return not (hub or broad)


def klu_btf_for(n, nnz):
    """pick_klu_btf_for (src/solver.cpp:517-532), no env override."""
    if n <= 10000:
        return 1
    if n <= 25000 and nnz >= 600000:
        return 1
    if 35000 <= n <= 50000 and 500000 <= nnz <= 800000:
        return 1
    if 30000 <= n <= 38000 and nnz >= 1500000:
        return 1
    return 0


def klu_scale_for(n, nnz, btf):
    """pick_klu_scale_for (src/solver.cpp:534-554), no env override."""
    if btf == 1 and 35000 <= n <= 50000 and 500000 <= nnz <= 800000:
        return 2
    if btf == 1 and 30000 <= n <= 38000 and nnz >= 1500000:
        return 2
    if n >= 65000 and nnz >= 3000000:
        return 2
    return 0


def klu_tol_for(n, nnz, btf):
    """pick_klu_tol_for (src/solver.cpp:556-565), no env override; 0.001 is KLU's default."""
    if btf == 1 and 30000 <= n <= 38000 and nnz >= 1500000:
        return 1e-6
    return 0.001
