"""blockdaglu_amd — MI355X-native engine for the BLOCK route of BlockDAGLU.

Python mirror of the reference C++ API (blockdaglu/include/blockdaglu/solver.h:43-65)
over the host library's C-ABI (include/bdlu_host.h):

    s = Solver()
    s.set_matrix(rows, cols, vals, dimension=-1)   # storage only
    s.analyze()                                    # host: validate, BTF, ordering, plan upload
    s.factor()                                     # device: refactorization (HIP, gfx950)
    s.solve_in_place(b)                            # device: block-DAG solve (+ refinement)

Every call returns a bool and sets last_error()/last_error_message() on failure,
as the reference does. The numeric work runs only in the HIP library; when it
cannot load or no GPU is present, factor()/analyze() fail with
BackendUnavailable — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_DIR = PKG_DIR / "lib"
HOST_LIB = LIB_DIR / "libblockdaglu.so"
GPU_LIB = LIB_DIR / "libbdlu_gpu.so"

_lib = None


class SolverError(enum.IntEnum):
    """Same enumerators, same order as blockdaglu::SolverError (solver.h:19-31)."""
    None_ = 0
    NotPrepared = 1
    InvalidDimension = 2
    InvalidIndex = 3
    NonFiniteValue = 4
    InvalidRightHandSide = 5
    BackendUnavailable = 6
    AnalysisFailed = 7
    FactorizationFailed = 8
    SolveFailed = 9
    ExportFailed = 10


class Precision(enum.IntEnum):
    FP64 = 0
    FP32_IR = 1


class _Report(C.Structure):
    _fields_ = [("nblocks", C.c_int32), ("dense_blocks", C.c_int32), ("sparse_blocks", C.c_int32),
                ("max_block", C.c_int32), ("fp64_fallback_blocks", C.c_int32), ("repivot_blocks", C.c_int32),
                ("ir_iterations", C.c_int32), ("plan_reused", C.c_int32), ("rel_residual", C.c_double),
                ("backward_error", C.c_double), ("componentwise_error", C.c_double),
                ("ir_fp64_refactor", C.c_int32)]


class GenSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_small_blocks", C.c_int32), ("size_dist", C.c_int32),
                ("size_param", C.c_double), ("size_lo", C.c_int32), ("size_hi", C.c_int32),
                ("n_small_total", C.c_int64), ("n_large_blocks", C.c_int32), ("large_size_lo", C.c_int32),
                ("large_size_hi", C.c_int32), ("large_fill", C.c_int32), ("large_fill_param", C.c_double),
                ("extra_per_row", C.c_double), ("coupling_kind", C.c_int32), ("coupling_rate", C.c_double),
                ("diag_factor", C.c_double), ("nrhs", C.c_int32), ("scale_exp", C.c_double),
                ("scale_seed", C.c_uint64), ("keep_order", C.c_int32)]


class _ServeResult(C.Structure):
    _fields_ = [("error", C.c_int32), ("slot", C.c_int32), ("ir_iterations", C.c_int32),
                ("fp64_fallback_blocks", C.c_int32), ("repivots", C.c_int32), ("rel_residual", C.c_double),
                ("backward_error", C.c_double), ("componentwise_error", C.c_double), ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_done_ms", C.c_double), ("ir_fp64_refactor", C.c_int32)]


_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)

# (name, restype, argtypes) of every symbol include/bdlu_host.h declares.
HOST_SYMBOLS = [
    ("bdlu_solver_create", _P, []),
    ("bdlu_solver_destroy", None, [_P]),
    ("bdlu_solver_set_matrix", C.c_int, [_P, C.c_int64, _I32P, _I32P, _F64P, C.c_int32]),
    ("bdlu_solver_analyze", C.c_int, [_P]),
    ("bdlu_solver_factor", C.c_int, [_P]),
    ("bdlu_solver_solve", C.c_int, [_P, _F64P, C.c_int64]),
    ("bdlu_solver_solve_multi", C.c_int, [_P, _F64P, C.c_int64, C.c_int32]),
    ("bdlu_solver_last_error", C.c_int, [_P]),
    ("bdlu_solver_last_error_message", C.c_char_p, [_P]),
    ("bdlu_solver_n", C.c_int32, [_P]),
    ("bdlu_solver_selected_mode", C.c_char_p, [_P]),
    ("bdlu_solver_set_precision", None, [_P, C.c_int]),
    ("bdlu_solver_report", None, [_P, C.POINTER(_Report)]),
    ("bdlu_solver_factor_device", C.c_int, [_P, C.c_void_p]),
    ("bdlu_solver_solve_device", C.c_int, [_P, C.c_void_p, C.c_int32, C.c_int64]),
    ("bdlu_solver_device_context", _P, [_P]),
    ("bdlu_solver_device_stream", _P, [_P]),
    ("bdlu_solver_device_plan", C.c_uint64, [_P]),
    ("bdlu_solver_nblocks", C.c_int32, [_P]),
    ("bdlu_solver_copy_perm", None, [_P, _I32P, _I32P, _I32P]),
    ("bdlu_solver_sparse_order", C.c_int32, [_P, C.c_int32, _I32P, _I32P]),
    ("bdlu_solver_export", C.c_int64, [_P, C.c_int, _I32P, _I32P, _F64P]),
    ("bdlu_serve_create", _P, [C.c_int, C.c_int, C.c_uint64, C.c_int]),
    ("bdlu_serve_destroy", None, [_P]),
    ("bdlu_serve_add_pattern", C.c_int32, [_P, C.c_int64, _I32P, _I32P, _F64P, C.c_int32]),
    ("bdlu_serve_pattern_n", C.c_int32, [_P, C.c_int32]),
    ("bdlu_serve_submit", C.c_int64, [_P, C.c_int32, _F64P, C.c_int64, _F64P, C.c_int32]),
    ("bdlu_serve_submit_device", C.c_int64, [_P, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    ("bdlu_serve_wait", C.c_int, [_P, C.c_int64, _F64P, C.POINTER(_ServeResult)]),
    ("bdlu_host_analyze", _P, [C.c_int64, _I32P, _I32P, _F64P, C.c_int32, C.POINTER(C.c_int)]),
    ("bdlu_host_analyze_klu", _P, [C.c_int64, _I32P, _I32P, _F64P, C.c_int32, C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_int)]),
    ("bdlu_route_for", C.c_int, [C.c_int64, _I32P, _I32P, _F64P, C.c_int32, _I32P, _F64P]),
    ("bdlu_host_free", None, [_P]),
    ("bdlu_host_info", None, [_P, _I32P, _I32P, _I32P, _I32P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _I32P]),
    ("bdlu_host_copy_perm", None, [_P, _I32P, _I32P, _I32P]),
    ("bdlu_host_block_kinds", None, [_P, _I32P, _I32P]),
    ("bdlu_host_sparse_stats", C.c_int, [_P, C.c_int32, _I32P, _F64P]),
    ("bdlu_host_sparse_order", C.c_int, [_P, C.c_int32, _I32P, _I32P]),
    ("bdlu_host_colamd", None, [C.c_int32, _I32P, _I32P, _I32P]),
    ("bdlu_host_signatures", None, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("bdlu_gen_config_spec", None, [C.c_int, C.c_int, C.POINTER(GenSpec)]),
    ("bdlu_gen_create", _P, [C.POINTER(GenSpec)]),
    ("bdlu_gen_info", None, [_P, _I32P, C.POINTER(C.c_int64), _I32P]),
    ("bdlu_gen_copy", None, [_P, _I32P, _I32P, _F64P, _F64P]),
    ("bdlu_gen_free", None, [_P]),
    ("bdlu_gen_perturb", None, [_F64P, C.c_int64, C.c_uint64, C.c_double]),
]

# Symbols of include/bdlu_gpu.h (device engine).
GPU_SYMBOLS = [
    ("bdlu_gpu_device_count", C.c_int, [_I32P]),
    ("bdlu_gpu_build_info", C.c_char_p, []),
    ("bdlu_gpu_create", C.c_int, [C.c_int, C.c_uint64, C.POINTER(_P)]),
    ("bdlu_gpu_destroy", None, [_P]),
    ("bdlu_gpu_last_error", C.c_char_p, [_P]),
    ("bdlu_gpu_plan_upload", C.c_int, [_P, _P, C.POINTER(C.c_uint64)]),
    ("bdlu_gpu_plan_release", C.c_int, [_P, C.c_uint64]),
    ("bdlu_gpu_plan_bytes", C.c_int, [_P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("bdlu_gpu_plan_schedule", C.c_int, [_P, C.c_uint64, _I32P]),
    ("bdlu_gpu_refactor", C.c_int, [_P, C.c_uint64, C.c_void_p, C.c_int, C.c_int, _P]),
    ("bdlu_gpu_solve", C.c_int, [_P, C.c_uint64, C.c_void_p, C.c_int32, C.c_int64, C.c_int, _P, _P]),
    ("bdlu_gpu_solve_noalloc", C.c_int, [_P, C.c_uint64, C.c_void_p, C.c_int32, C.c_int64, _P, _P]),
    ("bdlu_gpu_plan_values", C.c_int, [_P, C.c_uint64, _F64P]),
    ("bdlu_gpu_stream", _P, [_P]),
    ("bdlu_gpu_plan_dense_pivots", C.c_int, [_P, C.c_uint64, _I32P]),
    ("bdlu_gpu_set_side_streams", C.c_int, [_P, C.c_int]),
    ("bdlu_gpu_set_solve_grid_max", C.c_int, [_P, C.c_int32]),
    ("bdlu_gpu_set_device_refinement", C.c_int, [_P, C.c_int]),
    ("bdlu_gpu_copy", C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("bdlu_gpu_profile", C.c_int, [_P, C.c_int]),
    ("bdlu_gpu_profile_read", C.c_int, [_P, C.c_int, _F64P, C.POINTER(C.c_int64), _F64P, _F64P]),
    ("bdlu_gpu_debug_dense_lu", C.c_int, [_P, C.c_int32, _F64P, C.c_int, C.c_int, C.c_int32]),
]

PROF_TAIL_GEMM, PROF_TAIL_PANEL, PROF_DENSE_LU, PROF_SPARSE, PROF_SOLVE = range(5)


class IrOpts(C.Structure):
    """bdlu_ir_opts (include/bdlu_gpu.h)."""
    _fields_ = [("max_iters", C.c_int32), ("stop_rule", C.c_int32), ("tol", C.c_double),
                ("fp64_fallback", C.c_int32)]


class SolveStats(C.Structure):
    """bdlu_solve_stats (include/bdlu_gpu.h)."""
    _fields_ = [("iterations", C.c_int32), ("rel_residual", C.c_double), ("backward_error", C.c_double),
                ("componentwise_error", C.c_double), ("fp64_refactor", C.c_int32)]


_hip = None


def hip() -> C.CDLL:
    """The HIP runtime the engine library is linked against (already loaded by
    lib()), for device buffers and events in callers such as the benchmark."""
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]), ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipMemcpyAsync", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]),
                           ("hipDeviceSynchronize", []), ("hipStreamSynchronize", [C.c_void_p]),
                           ("hipStreamCreateWithFlags", [C.POINTER(C.c_void_p), C.c_uint]),
                           ("hipEventCreate", [C.POINTER(C.c_void_p)]), ("hipEventDestroy", [C.c_void_p]),
                           ("hipEventRecord", [C.c_void_p, C.c_void_p]), ("hipEventSynchronize", [C.c_void_p]),
                           ("hipEventElapsedTime", [C.POINTER(C.c_float), C.c_void_p, C.c_void_p])):
            f = getattr(H, name)
            f.restype = C.c_int
            f.argtypes = args
        _hip = H
    return _hip


def _hipcheck(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


_copy_stream = None


def _copies() -> C.c_void_p:
    """A non-blocking stream of this module for host<->device copies. Copies on
    the legacy default stream (plain hipMemcpy) fail with
    hipErrorStreamCaptureImplicit while any thread of the process captures a
    graph, and the engine captures its launch sequences on first use."""
    global _copy_stream
    if _copy_stream is None:
        st = C.c_void_p()
        _hipcheck(hip().hipStreamCreateWithFlags(C.byref(st), 1), "hipStreamCreateWithFlags")
        _copy_stream = st
    return _copy_stream


class DeviceArray:
    """A float64 buffer in device memory (HBM) with host copies in and out
    (on a non-blocking stream of their own, complete when the call returns)."""
    H2D, D2H, D2D = 1, 2, 3

    def __init__(self, count: int):
        self.count = int(count)
        self.ptr = C.c_void_p()
        _hipcheck(hip().hipMalloc(C.byref(self.ptr), max(1, self.count) * 8), "hipMalloc")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size <= self.count
        st = _copies()
        _hipcheck(hip().hipMemcpyAsync(self.ptr, a.ctypes.data, a.size * 8, self.H2D, st), "hipMemcpyAsync")
        _hipcheck(hip().hipStreamSynchronize(st), "hipStreamSynchronize")

    def copy_from(self, other: "DeviceArray", stream: int = 0) -> None:
        assert other.count <= self.count
        _hipcheck(hip().hipMemcpyAsync(self.ptr, other.ptr, other.count * 8, self.D2D, C.c_void_p(stream)),
                  "hipMemcpyAsync")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.count, np.float64)
        st = _copies()
        _hipcheck(hip().hipMemcpyAsync(out.ctypes.data, self.ptr, self.count * 8, self.D2H, st), "hipMemcpyAsync")
        _hipcheck(hip().hipStreamSynchronize(st), "hipStreamSynchronize")
        return out

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def free(self) -> None:
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    """A HIP event (timing) on a given stream."""

    def __init__(self):
        self.e = C.c_void_p()
        _hipcheck(hip().hipEventCreate(C.byref(self.e)), "hipEventCreate")

    def record(self, stream: int) -> None:
        _hipcheck(hip().hipEventRecord(self.e, C.c_void_p(stream)), "hipEventRecord")

    def synchronize(self) -> None:
        _hipcheck(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        _hipcheck(hip().hipEventElapsedTime(C.byref(ms), self.e, later.e), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            hip().hipEventDestroy(self.e)
        except Exception:
            pass


def lib() -> C.CDLL:
    """Loads the host library (and, through it, the HIP engine). Raises if absent."""
    global _lib
    if _lib is None:
        if not HOST_LIB.exists() or not GPU_LIB.exists():
            raise RuntimeError(f"blockdaglu_amd native libraries are not built ({LIB_DIR}); run build()")
        C.CDLL(str(GPU_LIB), mode=C.RTLD_GLOBAL)
        L = C.CDLL(str(HOST_LIB))
        for name, res, args in HOST_SYMBOLS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        G = C.CDLL(str(GPU_LIB))
        for name, res, args in GPU_SYMBOLS:
            f = getattr(G, name)
            f.restype = res
            f.argtypes = args
        L.gpu = G
        _lib = L
    return _lib


def build(oracle: bool = False) -> None:
    import importlib.util
    spec = importlib.util.spec_from_file_location("_bdlu_build", PKG_DIR / "build.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.build(oracle=oracle)


def build_info() -> str:
    """Source commit and target recorded in libbdlu_gpu.so when it was built."""
    return lib().gpu.bdlu_gpu_build_info().decode()


def gpu_available() -> bool:
    n = C.c_int32(0)
    return lib().gpu.bdlu_gpu_device_count(C.byref(n)) == 0 and n.value > 0


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Solver:
    """blockdaglu::Solver over the C-ABI (reference solver.h:43-65)."""

    def __init__(self):
        self._L = lib()
        self._h = self._L.bdlu_solver_create()
        if not self._h:
            raise MemoryError("bdlu_solver_create failed")

    def close(self):
        if getattr(self, "_h", None):
            self._L.bdlu_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_matrix(self, rows, cols=None, vals=None, dimension: int = -1) -> bool:
        """set_matrix(coo_triples) or set_matrix(rows, cols, vals); storage only."""
        if cols is None:
            t = list(rows)
            rows = [x[0] for x in t]
            cols = [x[1] for x in t]
            vals = [x[2] for x in t]
        r, c, v = _i32(rows), _i32(cols), _f64(vals)
        if not (len(r) == len(c) == len(v)):
            raise ValueError("rows/cols/vals lengths differ")
        return bool(self._L.bdlu_solver_set_matrix(self._h, len(r), _ptr(r, _I32P), _ptr(c, _I32P),
                                                   _ptr(v, _F64P), int(dimension)))

    def analyze(self) -> bool:
        return bool(self._L.bdlu_solver_analyze(self._h))

    def factor(self) -> bool:
        return bool(self._L.bdlu_solver_factor(self._h))

    def solve_in_place(self, b: np.ndarray) -> bool:
        if not (isinstance(b, np.ndarray) and b.dtype == np.float64 and b.flags.c_contiguous):
            raise TypeError("b must be a C-contiguous float64 numpy array (overwritten with x)")
        return bool(self._L.bdlu_solver_solve(self._h, _ptr(b, _F64P), b.size))

    def solve_multi_in_place(self, B: np.ndarray, nrhs: int) -> bool:
        """B: n*nrhs float64, column-major (Fortran order or a flat column stack)."""
        if not (isinstance(B, np.ndarray) and B.dtype == np.float64 and
                (B.flags.f_contiguous or B.ndim == 1)):
            raise TypeError("B must be a float64 Fortran-ordered array")
        return bool(self._L.bdlu_solver_solve_multi(self._h, _ptr(B, _F64P), B.size // nrhs, nrhs))

    def factor_device(self, d_values_ptr: int) -> bool:
        return bool(self._L.bdlu_solver_factor_device(self._h, C.c_void_p(d_values_ptr)))

    def solve_device(self, d_B_ptr: int, nrhs: int, ldb: int) -> bool:
        return bool(self._L.bdlu_solver_solve_device(self._h, C.c_void_p(d_B_ptr), nrhs, ldb))

    def device_stream(self) -> int:
        return self._L.bdlu_solver_device_stream(self._h) or 0

    def device_context(self) -> int:
        return self._L.bdlu_solver_device_context(self._h) or 0

    def device_plan(self) -> int:
        return int(self._L.bdlu_solver_device_plan(self._h))

    def last_error(self) -> SolverError:
        return SolverError(self._L.bdlu_solver_last_error(self._h))

    def last_error_message(self) -> str:
        return self._L.bdlu_solver_last_error_message(self._h).decode()

    def n(self) -> int:
        return self._L.bdlu_solver_n(self._h)

    def selected_mode(self) -> str:
        return self._L.bdlu_solver_selected_mode(self._h).decode()

    def set_precision(self, p: Precision) -> None:
        self._L.bdlu_solver_set_precision(self._h, int(p))

    def report(self) -> dict:
        r = _Report()
        self._L.bdlu_solver_report(self._h, C.byref(r))
        return {k: getattr(r, k) for k, _ in _Report._fields_}

    def permutations(self):
        nb = self._L.bdlu_solver_nblocks(self._h)
        n = self.n()
        P = np.zeros(max(n, 0), np.int32)
        Q = np.zeros(max(n, 0), np.int32)
        R = np.zeros(nb + 1, np.int32)
        self._L.bdlu_solver_copy_perm(self._h, _ptr(P, _I32P), _ptr(Q, _I32P), _ptr(R, _I32P))
        return P, Q, R

    def dense_pivots(self):
        """Row pivots of the dense blocks from the last factor(): for each permuted
        row i inside a dense block, the block-local row the pivoting brought there."""
        piv = np.zeros(self.n(), np.int32)
        rc = lib().gpu.bdlu_gpu_plan_dense_pivots(self.device_context(), self.device_plan(), _ptr(piv, _I32P))
        if rc != 0:
            raise RuntimeError("bdlu_gpu_plan_dense_pivots failed")
        return piv

    def sparse_order(self, b: int, size: int):
        """(q, prow, k0) of block b if it is above the dense threshold, else None."""
        q = np.zeros(size, np.int32)
        prow = np.zeros(size, np.int32)
        k0 = self._L.bdlu_solver_sparse_order(self._h, b, _ptr(q, _I32P), _ptr(prow, _I32P))
        return None if k0 < 0 else (q, prow, k0)

    def _export(self, which):
        cnt = self._L.bdlu_solver_export(self._h, which, None, None, None)
        if cnt < 0:
            return []
        r = np.zeros(cnt, np.int32)
        c = np.zeros(cnt, np.int32)
        v = np.zeros(cnt, np.float64)
        self._L.bdlu_solver_export(self._h, which, _ptr(r, _I32P), _ptr(c, _I32P), _ptr(v, _F64P))
        return list(zip(r.tolist(), c.tolist(), v.tolist()))

    def get_rearranged_matrix(self):
        return self._export(0)

    def getP(self):
        return self._export(1)

    def getQ(self):
        return self._export(2)

    def getL(self):
        return self._export(3)

    def getU(self):
        return self._export(4)


class Server:
    """Serving loop (include/blockdaglu/serve.h): registered patterns, plans
    resident per slot, requests admitted in arrival order as slots free up, each
    slot on its own stream so in-flight requests overlap on the GPU."""

    def __init__(self, slots: int = 4, hbm_budget_gb: float = 64.0, precision: Precision = Precision.FP64,
                 device: int = 0):
        self._L = lib()
        self._h = self._L.bdlu_serve_create(device, slots, int(hbm_budget_gb * (1 << 30)), int(precision))
        if not self._h:
            raise RuntimeError("bdlu_serve_create failed")
        self._n = {}
        self._nnz = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.bdlu_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_pattern(self, rows, cols, vals, dimension: int = -1) -> int:
        r, c, v = _i32(rows), _i32(cols), _f64(vals)
        pid = self._L.bdlu_serve_add_pattern(self._h, len(r), _ptr(r, _I32P), _ptr(c, _I32P), _ptr(v, _F64P),
                                            dimension)
        if pid < 0:
            raise RuntimeError(f"add_pattern failed: {SolverError(-pid).name}")
        self._n[pid] = self._L.bdlu_serve_pattern_n(self._h, pid)
        self._nnz[pid] = len(r)
        return pid

    def submit(self, pattern: int, vals, B) -> int:
        v = _f64(vals)
        b = np.asfortranarray(B, dtype=np.float64)
        k = 1 if b.ndim == 1 else b.shape[1]
        rid = self._L.bdlu_serve_submit(self._h, pattern, _ptr(v, _F64P), len(v), _ptr(b, _F64P), k)
        if rid < 0:
            raise RuntimeError(f"submit failed: {SolverError(-rid).name}")
        self._pending = getattr(self, "_pending", {})
        self._pending[rid] = (self._n[pattern], k, b.ndim)
        return rid

    def submit_device(self, pattern: int, d_vals: int, d_B: int, d_X: int, nrhs: int) -> int:
        """A request whose inputs are already in device memory (addresses):
        d_vals the pattern's canonical values, d_B / d_X n x nrhs column-major
        right-hand sides / solutions. wait() then returns X = None."""
        rid = self._L.bdlu_serve_submit_device(self._h, pattern, C.c_void_p(d_vals), C.c_void_p(d_B),
                                               C.c_void_p(d_X), nrhs)
        if rid < 0:
            raise RuntimeError(f"submit_device failed: {SolverError(-rid).name}")
        self._pending = getattr(self, "_pending", {})
        self._pending[rid] = None
        return rid

    def wait(self, request: int):
        """-> (X, result dict). X has the shape of the submitted B (None for a
        device-resident request: the solution is in its d_X)."""
        info = self._pending.pop(request)
        res = _ServeResult()
        if info is None:
            X = None
            self._L.bdlu_serve_wait(self._h, request, None, C.byref(res))
        else:
            n, k, nd = info
            X = np.zeros((n, k), np.float64, order="F")
            self._L.bdlu_serve_wait(self._h, request, _ptr(X, _F64P), C.byref(res))
            X = X[:, 0] if nd == 1 else X
        out = {f: getattr(res, f) for f, _ in _ServeResult._fields_}
        out["error"] = SolverError(out["error"])
        return X, out


class HostAnalysis:
    """Read-only view of the host analysis (canonicalize, BTF, merge, per-block symbolic). CPU only."""

    def __init__(self, rows, cols, vals, dimension: int = -1, klu=None):
        """klu: None for the BLOCK route's analysis, or (btf, scale, tol) for
        the global-KLU route's."""
        L = lib()
        r, c, v = _i32(rows), _i32(cols), _f64(vals)
        err = C.c_int(0)
        self._L = L
        if klu is None:
            self._h = L.bdlu_host_analyze(len(r), _ptr(r, _I32P), _ptr(c, _I32P), _ptr(v, _F64P), dimension,
                                          C.byref(err))
        else:
            btf, scale, tol = klu
            self._h = L.bdlu_host_analyze_klu(len(r), _ptr(r, _I32P), _ptr(c, _I32P), _ptr(v, _F64P), dimension,
                                              int(btf), int(scale), float(tol), C.byref(err))
        self.error = SolverError(err.value)
        if not self._h:
            return
        n, nbr, nb, nm, nl = (C.c_int32() for _ in range(5))
        nnz, noff = C.c_int64(), C.c_int64()
        L.bdlu_host_info(self._h, C.byref(n), C.byref(nbr), C.byref(nb), C.byref(nm), C.byref(nnz), C.byref(noff),
                         C.byref(nl))
        self.n, self.nblocks_raw, self.nblocks, self.nmatch = n.value, nbr.value, nb.value, nm.value
        self.nnz, self.nnz_off, self.nlevels = nnz.value, noff.value, nl.value

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.bdlu_host_free(self._h)
            self._h = None

    def permutations(self):
        P = np.zeros(self.n, np.int32)
        Q = np.zeros(self.n, np.int32)
        R = np.zeros(self.nblocks + 1, np.int32)
        self._L.bdlu_host_copy_perm(self._h, _ptr(P, _I32P), _ptr(Q, _I32P), _ptr(R, _I32P))
        return P, Q, R

    def block_kinds(self):
        k = np.zeros(self.nblocks, np.int32)
        s = np.zeros(self.nblocks, np.int32)
        self._L.bdlu_host_block_kinds(self._h, _ptr(k, _I32P), _ptr(s, _I32P))
        return k, s

    def sparse_stats(self, b):
        st = np.zeros(6, np.int32)
        fs = np.zeros(4, np.float64)
        if not self._L.bdlu_host_sparse_stats(self._h, b, _ptr(st, _I32P), _ptr(fs, _F64P)):
            return None
        keys = ["s", "k0", "m", "levels", "offdiag_pivots", "scale"]
        out = dict(zip(keys, st.tolist()))
        out.update(dict(zip(["nnz_l", "nnz_u", "updates", "flops"], fs.tolist())))
        return out

    def sparse_order(self, b):
        st = self.sparse_stats(b)
        q = np.zeros(st["s"], np.int32)
        prow = np.zeros(st["s"], np.int32)
        self._L.bdlu_host_sparse_order(self._h, b, _ptr(q, _I32P), _ptr(prow, _I32P))
        return q, prow

    def signatures(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._L.bdlu_host_signatures(self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value



def route_for(rows, cols, vals, dimension: int = -1) -> dict:
    """The route Solver.analyze would take for this matrix under the current
    environment: {"klu": bool, "btf", "scale", "ordering", "tol"} (CPU only)."""
    r, c, v = _i32(rows), _i32(cols), _f64(vals)
    out = np.zeros(4, np.int32)
    tol = C.c_double()
    rc = lib().bdlu_route_for(len(r), _ptr(r, _I32P), _ptr(c, _I32P), _ptr(v, _F64P), dimension, _ptr(out, _I32P),
                              C.byref(tol))
    if rc != 0:
        raise ValueError("matrix fails validation")
    return {"klu": bool(out[0]), "btf": int(out[1]), "scale": int(out[2]), "ordering": int(out[3]),
            "tol": tol.value}


def colamd_order(n: int, indptr, indices) -> np.ndarray:
    """Column order the host analysis gives an n x n block with this CSC
    pattern (order[k] = column at position k)."""
    ip, ix = _i32(indptr), _i32(indices)
    out = np.zeros(n, np.int32)
    lib().bdlu_host_colamd(n, _ptr(ip, _I32P), _ptr(ix, _I32P), _ptr(out, _I32P))
    return out


def config_spec(cfg: int, pattern: int = 0) -> GenSpec:
    sp = GenSpec()
    lib().bdlu_gen_config_spec(cfg, pattern, C.byref(sp))
    return sp


def generate(spec_or_cfg, pattern: int = 0, **overrides):
    """Seeded workload -> (n, rows, cols, vals, rhs[n, nrhs] Fortran order)."""
    L = lib()
    sp = config_spec(spec_or_cfg, pattern) if isinstance(spec_or_cfg, int) else spec_or_cfg
    for k, v in overrides.items():
        setattr(sp, k, v)
    h = L.bdlu_gen_create(C.byref(sp))
    if not h:
        raise RuntimeError("generator failed")
    try:
        n, nnz, nrhs = C.c_int32(), C.c_int64(), C.c_int32()
        L.bdlu_gen_info(h, C.byref(n), C.byref(nnz), C.byref(nrhs))
        r = np.zeros(nnz.value, np.int32)
        c = np.zeros(nnz.value, np.int32)
        v = np.zeros(nnz.value, np.float64)
        b = np.zeros((n.value, nrhs.value), np.float64, order="F")
        L.bdlu_gen_copy(h, _ptr(r, _I32P), _ptr(c, _I32P), _ptr(v, _F64P), _ptr(b, _F64P))
    finally:
        L.bdlu_gen_free(h)
    return n.value, r, c, v, b


def perturb(vals: np.ndarray, seed: int, rel: float = 0.05) -> np.ndarray:
    out = _f64(vals).copy()
    lib().bdlu_gen_perturb(_ptr(out, _F64P), out.size, seed, rel)
    return out


def canonical_order(rows, cols) -> np.ndarray:
    """Permutation that sorts COO entries by (row, col) — the canonical order of
    a duplicate-free input (reference src/solver.cpp:692)."""
    return np.lexsort((np.asarray(cols), np.asarray(rows)))
