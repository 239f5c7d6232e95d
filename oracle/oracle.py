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
