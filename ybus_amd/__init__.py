"""ybus_amd: ybus's GPU factor-solve path on MI355X, from Python.

A thin ctypes face over ``libybus_hip.so`` (the C ABI in ``include/ybus_hip.h``) that mirrors
the reference crate ybus_cuda's API, name for name (ybus_cuda/src/lib.rs, factor.rs, pairs.rs):
``Gpu``, ``GpuFactor``, ``GpuBlock``, ``GpuOperator``, ``Pair``, ``PairUpdate``, ``ColumnMax``,
``Timings`` and the ``GpuError`` family for ``CudaError``. All compute runs in the HIP library;
there is no CPU path, and importing this package on a machine without the built library fails
loudly.
"""
from __future__ import annotations

import ctypes as C
import mmap
import os
import sys
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libybus_hip.so")

OK, NO_DEVICE, INVALID_ARGUMENT, SHAPE_MISMATCH, TOO_LARGE, DEVICE_ERROR, OUT_OF_MEMORY, INTERNAL, PATTERN_MISMATCH = range(9)
FP64, FP32_REFINED, TRANSPOSE = 0, 1, 2
# refined solves' stop rules (ybus_hip_factor_set_refine_stop)
STOP_NORMWISE, STOP_AT_FLOOR = 0, 1
SCRATCH_BYTES = 64 << 20  # YBUS_HIP_SCRATCH_BYTES: the no-allocation workspace per device

# ---- errors: CudaError (ybus_cuda/src/lib.rs:88-124) ----


class GpuError(Exception):
    """An error from the GPU path, with the C ABI's status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class ShapeMismatch(GpuError):
    """A right-hand side doesn't have one entry per row of the factorization."""

    def __init__(self, expected: int, found: int):
        super().__init__(SHAPE_MISMATCH, f"a right-hand side has {found} entries, but the factorization has "
                                         f"{expected} rows")
        self.expected = expected
        self.found = found


class TooLarge(GpuError):
    pass


class InvalidArgument(GpuError, ValueError):
    """A caller bug the reference asserts on (a panic there)."""


class DeviceError(GpuError):
    pass


class PatternMismatch(GpuError):
    """SparseError::PatternMismatch (lu.rs:662-667): the matrix's pattern is not the factor's."""


_ERRORS = {TOO_LARGE: TooLarge, INVALID_ARGUMENT: InvalidArgument, DEVICE_ERROR: DeviceError,
           PATTERN_MISMATCH: PatternMismatch}

_lib = None


def library() -> C.CDLL:
    """The loaded HIP library; raises if it hasn't been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ybus_amd.build` "
                               "(the GPU path has no CPU fallback)")
        # One HIP runtime per process. PyTorch ships its own libamdhip64 (soname libamdhip64.so.7,
        # needed by torch as libamdhip64.so): loaded first, it is the one the library binds to;
        # loaded after the library's, torch would map a second runtime beside it, which then finds
        # no GPU. So torch, where installed, is imported before the library is loaded.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class _Timings(C.Structure):
    _fields_ = [("upload", C.c_double), ("solve", C.c_double), ("download", C.c_double), ("blocks", C.c_uint64)]


class _ColumnMax(C.Structure):
    _fields_ = [("row", C.c_int64), ("value", C.c_double), ("not_finite", C.c_int32)]


# ybus_hip_column_max as a numpy record, for maxima read as arrays
_COLUMN_MAX = np.dtype([("row", np.int64), ("value", np.float64), ("not_finite", np.int32)], align=True)
assert _COLUMN_MAX.itemsize == C.sizeof(_ColumnMax)


class _Tuning(C.Structure):
    _fields_ = [("path", C.c_int32), ("threads", C.c_int32), ("tile", C.c_int32), ("milliseconds", C.c_double),
                ("parts", C.c_int32), ("top_tile", C.c_int32), ("top_threads", C.c_int32),
                ("parts_lds_tile", C.c_int32), ("parts_lds_threads", C.c_int32), ("window", C.c_int32)]


class _KernelStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("milliseconds", C.c_double), ("bytes", C.c_uint64)]


class _RefactorStats(C.Structure):
    _fields_ = [("calls", C.c_uint64), ("reused", C.c_uint64), ("updates", C.c_uint64),
                ("refactor_ms", C.c_double), ("refresh_ms", C.c_double), ("multiply_subtracts", C.c_uint64)]


P = C.c_void_p
U64 = C.c_uint64


def _declare(L):
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
    i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
    u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
    i32 = C.c_int32
    sig = {
        "ybus_hip_last_error_message": (C.c_size_t, [C.c_char_p, C.c_size_t]),
        "ybus_hip_ctx_create": (i32, [C.POINTER(P)]),
        "ybus_hip_ctx_free": (None, [P]),
        "ybus_hip_ctx_name": (i32, [P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "ybus_hip_ctx_memory": (i32, [P, C.POINTER(U64), C.POINTER(U64)]),
        "ybus_hip_factor_upload_cholesky": (i32, [P, U64, u64p, u64p, f64p, u64p, C.POINTER(P)]),
        "ybus_hip_factor_upload_lu": (i32, [P, U64, u64p, u64p, f64p, u64p, u64p, f64p, u64p, u64p, C.POINTER(P)]),
        "ybus_hip_factor_upload_cholesky_complex": (i32, [P, U64, u64p, u64p, f64p, u64p, C.POINTER(P)]),
        "ybus_hip_factor_upload_lu_complex": (i32, [P, U64, u64p, u64p, f64p, u64p, u64p, f64p, u64p, u64p,
                                                    C.POINTER(P)]),
        "ybus_hip_factor_is_complex": (i32, [P]),
        "ybus_hip_factor_solve_many_complex": (i32, [P, f64p, U64, f64p]),
        "ybus_hip_factor_solve_transpose_many_complex": (i32, [P, f64p, U64, f64p]),
        "ybus_hip_factor_set_matrix_complex": (i32, [P, i32, u64p, u64p, f64p]),
        "ybus_hip_factor_set_refine_stop": (i32, [P, i32]),
        "ybus_hip_batch_set_refine_stop": (i32, [P, i32]),
        "ybus_hip_factor_solve_many_refined_complex": (i32, [P, f64p, U64, C.c_double, U64, f64p, u64p, f64p]),
        "ybus_hip_factor_solve_device_complex": (i32, [P, i32, C.c_double, U64, C.c_void_p, U64, U64, C.c_void_p,
                                                       U64]),
        "ybus_hip_factor_free": (None, [P]),
        "ybus_hip_factor_dim": (U64, [P]),
        "ybus_hip_factor_block_width": (U64, [P]),
        "ybus_hip_factor_set_block_width": (None, [P, U64]),
        "ybus_hip_factor_timings": (None, [P, C.POINTER(_Timings)]),
        "ybus_hip_factor_reset_timings": (None, [P]),
        "ybus_hip_factor_set_matrix": (i32, [P, i32, u64p, u64p, f64p]),
        "ybus_hip_factor_solve_many": (i32, [P, f64p, U64, f64p]),
        "ybus_hip_factor_solve_transpose_many": (i32, [P, f64p, U64, f64p]),
        "ybus_hip_factor_solve_many_refined": (i32, [P, f64p, U64, C.c_double, U64, f64p, u64p, f64p]),
        "ybus_hip_factor_adjoint": (i32, [P, i32, C.c_double, U64, f64p, f64p, U64, f64p, f64p]),
        "ybus_hip_factor_solve_device": (i32, [P, i32, C.c_double, U64, C.c_void_p, U64, U64, C.c_void_p, U64]),
        "ybus_hip_factor_solve_device_noalloc": (i32, [P, i32, C.c_void_p, U64, U64, C.c_void_p, U64]),
        "ybus_hip_factor_adjoint_device": (i32, [P, i32, C.c_double, U64, C.c_void_p, C.c_void_p, U64, U64,
                                                 C.c_void_p, C.c_void_p]),
        "ybus_hip_factor_autotune": (i32, [P, U64, i32, C.c_void_p, C.c_void_p, C.POINTER(_Tuning)]),
        "ybus_hip_factor_stage": (i32, [P, f64p, U64]),
        "ybus_hip_factor_stage_pairs": (i32, [P, i64p, i64p, U64]),
        "ybus_hip_factor_solve_staged": (i32, [P, i32, C.c_double, U64, C.POINTER(P)]),
        "ybus_hip_factor_pair_solves": (i32, [P, i64p, i64p, U64, C.POINTER(P)]),
        "ybus_hip_factor_pair_updates": (i32, [P, f64p, i64p, i64p, f64p, f64p, U64, C.POINTER(P)]),
        "ybus_hip_factor_outage_updates": (i32, [P, f64p, u64p, U64, i64p, i64p, f64p, f64p, C.POINTER(P)]),
        "ybus_hip_block_free": (None, [P]),
        "ybus_hip_block_columns": (U64, [P]),
        "ybus_hip_block_singular": (i32, [P, u8p]),
        "ybus_hip_block_refinement": (i32, [P, u64p, f64p]),
        "ybus_hip_block_solutions": (i32, [P, f64p]),
        "ybus_hip_block_maxima": (i32, [P, C.POINTER(_ColumnMax)]),
        "ybus_hip_block_product": (i32, [P, P, i64p, f64p]),
        "ybus_hip_block_product_maxima": (i32, [P, P, i64p, C.POINTER(_ColumnMax)]),
        "ybus_hip_factor_kernel_stats": (None, [P, C.POINTER(_KernelStats)]),
        "ybus_hip_factor_path": (i32, [P, U64, i32]),
        "ybus_hip_factor_reset_kernel_stats": (None, [P]),
        "ybus_hip_factor_operator": (i32, [P, U64, u64p, u64p, f64p, C.POINTER(P)]),
        "ybus_hip_operator_free": (None, [P]),
        "ybus_hip_operator_rows": (U64, [P]),
        "ybus_hip_operator_set_offset": (i32, [P, f64p]),
        "ybus_hip_operator_set_weights": (i32, [P, f64p]),
        "ybus_hip_batch_create": (i32, [P, C.POINTER(P)]),
        "ybus_hip_batch_free": (None, [P]),
        "ybus_hip_batch_add": (i32, [P, P]),
        "ybus_hip_batch_factors": (U64, [P]),
        "ybus_hip_batch_stage": (i32, [P, C.c_void_p, u64p]),
        "ybus_hip_batch_solve_staged": (i32, [P, i32, C.c_double, U64]),
        "ybus_hip_batch_refinement": (i32, [P, C.c_void_p, C.c_void_p]),
        "ybus_hip_batch_solutions": (i32, [P, C.c_void_p]),
        "ybus_hip_batch_solve": (i32, [P, C.c_void_p, u64p, C.c_void_p]),
        "ybus_hip_batch_kernel_stats": (None, [P, C.POINTER(_KernelStats)]),
        "ybus_hip_batch_reset_kernel_stats": (None, [P]),
        "ybus_hip_factor_refactor_lu": (i32, [P, U64, u64p, u64p, f64p, C.c_double, C.POINTER(C.c_int32)]),
        "ybus_hip_factor_update_lu_values": (i32, [P, f64p, f64p]),
        "ybus_hip_factor_refactor_lu_complex": (i32, [P, U64, u64p, u64p, f64p, C.c_double, C.POINTER(C.c_int32)]),
        "ybus_hip_factor_update_lu_values_complex": (i32, [P, f64p, f64p]),
        "ybus_hip_factor_refactor_stats": (i32, [P, C.POINTER(_RefactorStats)]),
        "ybus_hip_factor_lu_values": (i32, [P, C.c_void_p, C.c_void_p]),
        "ybus_hip_analyze_cholesky": (i32, [U64, u64p, u64p, P, P, P, P]),
        "ybus_hip_analyze_lu": (i32, [U64, u64p, u64p, u64p, u64p, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _message() -> str:
    buf = C.create_string_buffer(1024)
    library().ybus_hip_last_error_message(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(status: int):
    if status != OK:
        raise _ERRORS.get(status, GpuError)(status, _message())


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- value types (pairs.rs:16-50, factor.rs:21-31) ----


@dataclass(frozen=True)
class Pair:
    """A right-hand side e_plus - e_minus over the unknowns (None for no entry)."""

    plus: Optional[int] = None
    minus: Optional[int] = None


@dataclass(frozen=True)
class PairUpdate:
    """One column of GpuFactor.pair_updates (pairs.rs:24-37)."""

    pair: Pair
    alpha: float
    inverse_weight: float


@dataclass(frozen=True)
class ColumnMax:
    """A column's largest |value| * weight, its row (None if every product is zero) and
    whether any value isn't finite (pairs.rs:39-50)."""

    row: Optional[int]
    value: float
    not_finite: bool


@dataclass(frozen=True)
class Timings:
    upload: float = 0.0
    solve: float = 0.0
    download: float = 0.0
    blocks: int = 0


@dataclass(frozen=True)
class KernelStats:
    launches: int
    milliseconds: float
    bytes: int


# ---- host-side symbolic analysis (no device needed) ----


def _ptr(a):
    return a.ctypes.data if a is not None else None


def analyze_cholesky(n, l_ptr, l_idx):
    """The level of each row of the forward (L) and backward (L^T) sweeps, and the fundamental
    supernodes' first columns (with n appended) of a Cholesky factor's L in CSR form."""
    fwd, bwd = np.zeros(n, np.int32), np.zeros(n, np.int32)
    starts, count = np.zeros(n + 1, np.uint64), U64()
    _check(library().ybus_hip_analyze_cholesky(int(n), _u64(l_ptr), _u64(l_idx), _ptr(fwd), _ptr(bwd), _ptr(starts),
                                               C.addressof(count)))
    return dict(forward=fwd, backward=bwd, supernodes=starts[: int(count.value) + 1].astype(np.int64))


def analyze_lu(n, l_ptr, l_idx, u_ptr, u_idx):
    """The level of each row of an LU factor's four sweeps: L, U, U^T and L^T."""
    out = [np.zeros(n, np.int32) for _ in range(4)]
    _check(library().ybus_hip_analyze_lu(int(n), _u64(l_ptr), _u64(l_idx), _u64(u_ptr), _u64(u_idx),
                                         *[_ptr(o) for o in out]))
    return dict(zip(["forward", "backward", "transpose_forward", "transpose_backward"], out))


# ---- Gpu (lib.rs:166-219) ----


class Gpu:
    """The first HIP device and the one stream all of this library's work runs on in order."""

    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def new() -> Optional["Gpu"]:
        """The first device, or None on a machine without one (Gpu::new's Ok(None))."""
        h = P()
        status = library().ybus_hip_ctx_create(C.byref(h))
        if status == NO_DEVICE:
            return None
        _check(status)
        return Gpu(h)

    def name(self) -> str:
        needed = C.c_size_t()
        buf = C.create_string_buffer(256)
        _check(library().ybus_hip_ctx_name(self._h, buf, 256, C.byref(needed)))
        return buf.value.decode()

    def memory(self):
        free, total = U64(), U64()
        _check(library().ybus_hip_ctx_memory(self._h, C.byref(free), C.byref(total)))
        return int(free.value), int(total.value)

    # Handles still open when the interpreter ends are left to the process's exit, which
    # releases the device memory: the module's globals are gone by then, and the objects' order
    # of collection is arbitrary (a factor could outlive its context).
    def __del__(self, _ending=sys.is_finalizing):
        if getattr(self, "_h", None) and not _ending():
            library().ybus_hip_ctx_free(self._h)
            self._h = None


# A result of this size or more is copied into transparent huge pages where the kernel offers them
HUGE_RESULT_BYTES = 64 << 20


def _result_array(shape, dtype=np.float64) -> np.ndarray:
    """A zeroed array for a result the library copies in from the device. From HUGE_RESULT_BYTES
    up it is an anonymous mapping advised to use transparent huge pages: the copy's first touch
    of each page then faults in 2 MiB rather than 4 KiB, which over a gigabyte of solutions is
    several milliseconds (the mapping is released with the array)."""
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    if nbytes < HUGE_RESULT_BYTES or not hasattr(mmap, "MADV_HUGEPAGE"):
        return np.zeros(shape, dtype)
    m = mmap.mmap(-1, nbytes, flags=mmap.MAP_PRIVATE | mmap.MAP_ANONYMOUS)
    try:
        m.madvise(mmap.MADV_HUGEPAGE)
    except OSError:
        pass  # no transparent huge pages here: small pages, zeroed all the same
    return np.frombuffer(m, dtype=dtype).reshape(shape)


def _rhs_array(rhs, n: int) -> np.ndarray:
    """The right-hand sides as a (count, n) array, checked like solve_many (factor.rs:234-244)."""
    if isinstance(rhs, np.ndarray) and rhs.ndim == 2:
        if rhs.shape[1] != n and rhs.shape[0] > 0:
            raise ShapeMismatch(n, rhs.shape[1])
        return _f64(rhs)
    rows = list(rhs)
    for b in rows:
        if len(b) != n:
            raise ShapeMismatch(n, len(b))
    if not rows:
        return np.zeros((0, n))
    return _f64(np.array(rows, dtype=np.float64).reshape(len(rows), n))


def _complex_rhs(rhs, n: int) -> np.ndarray:
    """Complex right-hand sides as a (count, n) complex128 array."""
    b = np.ascontiguousarray(rhs if isinstance(rhs, np.ndarray) else np.array(list(rhs)), dtype=np.complex128)
    if b.ndim == 1 and b.size == 0:
        b = b.reshape(0, n)
    if b.ndim != 2:
        raise ValueError("complex right-hand sides are a (count, n) array")
    if b.shape[1] != n and b.shape[0] > 0:
        raise ShapeMismatch(n, b.shape[1])
    return b


def _complex_compressed(a):
    """(row_major, ptr, idx, complex data) of a complex matrix: a SciPy CSR/CSC matrix, or an
    object with row_major, ptr, idx and complex data (or re and im)."""
    if hasattr(a, "indptr"):
        fmt = getattr(a, "format", "csr")
        if fmt not in ("csr", "csc"):
            a = a.tocsr()
            fmt = "csr"
        return fmt == "csr", np.asarray(a.indptr), np.asarray(a.indices), np.asarray(a.data, dtype=np.complex128)
    data = getattr(a, "data", None)
    if data is None or not np.iscomplexobj(data):
        data = np.asarray(a.re) + 1j * np.asarray(a.im)
    return bool(a.row_major), np.asarray(a.ptr), np.asarray(a.idx), np.asarray(data, dtype=np.complex128)


def _interleaved(values) -> np.ndarray:
    """Complex values as interleaved (re, im) doubles, Complex64's layout."""
    return np.ascontiguousarray(values, dtype=np.complex128).view(np.float64)


def _pairs(pairs: Sequence[Pair]):
    plus = np.array([-1 if p.plus is None else int(p.plus) for p in pairs], dtype=np.int64)
    minus = np.array([-1 if p.minus is None else int(p.minus) for p in pairs], dtype=np.int64)
    return plus, minus


# ---- GpuFactor (factor.rs, pairs.rs) ----


class GpuFactor:
    """A factorization uploaded once, for solving many right-hand sides against it."""

    def __init__(self, gpu: Gpu, handle, n: int, is_complex: bool = False):
        self._gpu = gpu  # keeps the context alive
        self._h = handle
        self._n = n
        self._complex = is_complex
        self._nnz_a = -1
        self._block: Optional[GpuBlock] = None
        self._lu_sizes = (0, 0)
        self._generation = 0

    @property
    def generation(self) -> int:
        """Counts the changes to the factor's values or its attached matrix (a refactor that
        reused its pivots, update_lu_values, set_matrix): work that began before a change, such as
        an autograd graph's forward solve, can tell that its backward would see other values."""
        return self._generation

    @staticmethod
    def from_cholesky(gpu: Gpu, cholesky) -> "GpuFactor":
        """Uploads P A P^T = L L^T (factor.rs:103-119). `cholesky` has ``l`` (CSR: ptr, idx,
        data, each row ending with its diagonal) and ``perm``, as ybus_sparse's Cholesky."""
        l = cholesky.l
        return GpuFactor.from_cholesky_arrays(gpu, len(cholesky.perm), l.ptr, l.idx, l.data, cholesky.perm)

    @staticmethod
    def from_cholesky_arrays(gpu: Gpu, n, l_ptr, l_idx, l_val, perm) -> "GpuFactor":
        h = P()
        _check(library().ybus_hip_factor_upload_cholesky(gpu._h, int(n), _u64(l_ptr), _u64(l_idx), _f64(l_val),
                                                         _u64(perm), C.byref(h)))
        return GpuFactor(gpu, h, int(n))

    @staticmethod
    def from_lu(gpu: Gpu, lu) -> "GpuFactor":
        """Uploads P A Q = L U (factor.rs:123-133). `lu` has ``l`` and ``u`` (CSC) and
        ``row_perm``/``col_perm``, as ybus_sparse's Lu."""
        return GpuFactor.from_lu_arrays(gpu, len(lu.row_perm), lu.l.ptr, lu.l.idx, lu.l.data, lu.u.ptr, lu.u.idx,
                                        lu.u.data, lu.row_perm, lu.col_perm)

    @staticmethod
    def from_lu_arrays(gpu: Gpu, n, l_ptr, l_idx, l_val, u_ptr, u_idx, u_val, row_perm, col_perm) -> "GpuFactor":
        h = P()
        _check(library().ybus_hip_factor_upload_lu(gpu._h, int(n), _u64(l_ptr), _u64(l_idx), _f64(l_val),
                                                   _u64(u_ptr), _u64(u_idx), _f64(u_val), _u64(row_perm),
                                                   _u64(col_perm), C.byref(h)))
        factor = GpuFactor(gpu, h, int(n))
        factor._lu_sizes = (int(np.asarray(l_ptr)[-1]), int(np.asarray(u_ptr)[-1]))
        return factor

    @staticmethod
    def from_cholesky_complex(gpu: Gpu, cholesky) -> "GpuFactor":
        """Uploads a complex P A P^T = L L^H (Cholesky<Complex64>, cholesky.rs:206-311):
        ``l`` in CSR with complex ``data``, and ``perm``."""
        l = cholesky.l
        return GpuFactor.from_cholesky_complex_arrays(gpu, len(cholesky.perm), l.ptr, l.idx, l.data, cholesky.perm)

    @staticmethod
    def from_cholesky_complex_arrays(gpu: Gpu, n, l_ptr, l_idx, l_val, perm) -> "GpuFactor":
        h = P()
        _check(library().ybus_hip_factor_upload_cholesky_complex(gpu._h, int(n), _u64(l_ptr), _u64(l_idx),
                                                                 _interleaved(l_val), _u64(perm), C.byref(h)))
        return GpuFactor(gpu, h, int(n), True)

    @staticmethod
    def from_lu_complex(gpu: Gpu, lu) -> "GpuFactor":
        """Uploads a complex P A Q = L U (Lu<Complex64>): ``l``, ``u`` (CSC, complex ``data``),
        ``row_perm`` and ``col_perm``."""
        return GpuFactor.from_lu_complex_arrays(gpu, len(lu.row_perm), lu.l.ptr, lu.l.idx, lu.l.data, lu.u.ptr,
                                                lu.u.idx, lu.u.data, lu.row_perm, lu.col_perm)

    @staticmethod
    def from_lu_complex_arrays(gpu: Gpu, n, l_ptr, l_idx, l_val, u_ptr, u_idx, u_val, row_perm,
                               col_perm) -> "GpuFactor":
        h = P()
        _check(library().ybus_hip_factor_upload_lu_complex(gpu._h, int(n), _u64(l_ptr), _u64(l_idx),
                                                           _interleaved(l_val), _u64(u_ptr), _u64(u_idx),
                                                           _interleaved(u_val), _u64(row_perm), _u64(col_perm),
                                                           C.byref(h)))
        factor = GpuFactor(gpu, h, int(n), True)
        factor._lu_sizes = (int(np.asarray(l_ptr)[-1]), int(np.asarray(u_ptr)[-1]))
        return factor

    def is_complex(self) -> bool:
        return bool(library().ybus_hip_factor_is_complex(self._h))

    def __del__(self, _ending=sys.is_finalizing):
        if getattr(self, "_h", None) and not _ending():
            if self._block is not None:
                self._block.free()
            library().ybus_hip_factor_free(self._h)
            self._h = None

    def dim(self) -> int:
        return self._n

    def block_width(self) -> int:
        return int(library().ybus_hip_factor_block_width(self._h))

    def set_refine_stop(self, rule: int):
        """The FP32-refined solves' stop rule: STOP_NORMWISE (refine_normwise's, the default) or
        STOP_AT_FLOOR (also stop once the ratio is within 2x of the FP64 residual's rounding
        floor, 2 u || |A| |x| ||inf / ||b||inf, saving the correction that would not halve it)."""
        _check(library().ybus_hip_factor_set_refine_stop(self._h, int(rule)))

    def set_block_width(self, width: int):
        library().ybus_hip_factor_set_block_width(self._h, int(width))

    def timings(self) -> Timings:
        t = _Timings()
        library().ybus_hip_factor_timings(self._h, C.byref(t))
        return Timings(t.upload, t.solve, t.download, int(t.blocks))

    def reset_timings(self):
        library().ybus_hip_factor_reset_timings(self._h)

    def path(self, columns: int, pairs: bool = False) -> str:
        """The kernels a block of `columns` right-hand sides (a pair block when `pairs`) takes:
        'columns' (one workgroup per right-hand side, its solution in LDS) or 'wide' (column
        tiles in global memory)."""
        return "wide" if library().ybus_hip_factor_path(self._h, columns, int(bool(pairs))) == 1 else "columns"

    def autotune(self, columns: int, pairs=False) -> dict:
        """Times the kernel configurations for blocks of `columns` right-hand sides and keeps the
        fastest for this factor's blocks of that width and kind: pair blocks when `pairs` (True
        for synthetic pairs, or the caller's own list of Pair to time on), else blocks of vectors.
        Returns {'path', 'threads', 'tile', 'parts', 'top_tile', 'top_threads', 'parts_lds_tile',
        'parts_lds_threads', 'window', 'milliseconds'} (top_*: the partitioned solve's top-of-tree
        kernel; parts_lds_*: its parts swept with their values in LDS, tile 0 where they stay in
        global memory; window: a one-schedule wide solve with its last positions in LDS)."""
        self._release_block()
        t = _Tuning()
        plus = minus = None
        if pairs is not True and pairs:
            plus, minus = _pairs(pairs)
            columns = len(pairs)
        _check(library().ybus_hip_factor_autotune(self._h, int(columns), int(bool(pairs)),
                                                  plus.ctypes.data if plus is not None else None,
                                                  minus.ctypes.data if minus is not None else None, C.byref(t)))
        return dict(path="wide" if t.path == 1 else "columns", threads=int(t.threads), tile=int(t.tile),
                    parts=bool(t.parts), top_tile=int(t.top_tile), top_threads=int(t.top_threads),
                    parts_lds_tile=int(t.parts_lds_tile), parts_lds_threads=int(t.parts_lds_threads),
                    window=bool(t.window), milliseconds=float(t.milliseconds))

    def kernel_stats(self) -> KernelStats:
        s = _KernelStats()
        library().ybus_hip_factor_kernel_stats(self._h, C.byref(s))
        return KernelStats(int(s.launches), float(s.milliseconds), int(s.bytes))

    def reset_kernel_stats(self):
        library().ybus_hip_factor_reset_kernel_stats(self._h)

    def refactor(self, a, pivot_threshold: float = 0.1) -> str:
        """Lu::refactor (lu.rs:662-757) on the device: `a` (CSC: ``ptr``, ``idx``, ``data``, as
        ybus_sparse's CscMatrix) factored with this LU factor's pivots and patterns. Returns
        "reused" (the factor now holds the new values, the CPU refactor's bits) or
        "needs_pivoting" (a reused pivot failed the threshold; the factor is unchanged, and the
        reference's fresh factorization is the caller's, followed by a new upload). Raises
        PatternMismatch for another pattern than the first call's."""
        if getattr(a, "row_major", False):
            raise InvalidArgument(INVALID_ARGUMENT, "refactor takes the matrix in CSC form")
        outcome = C.c_int32()
        if self._complex:
            # a complex matrix: data complex128, or ``re`` and ``im`` arrays beside the pattern
            values = getattr(a, "data", None)
            if values is None or not np.iscomplexobj(values):
                values = np.asarray(a.re) + 1j * np.asarray(a.im)
            _check(library().ybus_hip_factor_refactor_lu_complex(
                self._h, self._n, _u64(a.ptr), _u64(a.idx), _interleaved(values), float(pivot_threshold),
                C.byref(outcome)))
        else:
            _check(library().ybus_hip_factor_refactor_lu(self._h, self._n, _u64(a.ptr), _u64(a.idx), _f64(a.data),
                                                         float(pivot_threshold), C.byref(outcome)))
        if outcome.value == 0:
            self._generation += 1
        return "reused" if outcome.value == 0 else "needs_pivoting"

    def update_lu_values(self, l_val, u_val):
        """New L and U values (the uploaded CSC order, the same pattern and permutations), as a host
        refactor that reused its pivots makes them: the device sweeps are rewritten in place."""
        if self._complex:
            _check(library().ybus_hip_factor_update_lu_values_complex(self._h, _interleaved(l_val),
                                                                      _interleaved(u_val)))
        else:
            _check(library().ybus_hip_factor_update_lu_values(self._h, _f64(l_val), _f64(u_val)))
        self._generation += 1

    def refactor_stats(self) -> dict:
        s = _RefactorStats()
        _check(library().ybus_hip_factor_refactor_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in _RefactorStats._fields_}

    def lu_values(self):
        """The factor's current (L values, U values) in the uploaded CSC order."""
        parts = 2 if self._complex else 1
        l_val, u_val = np.zeros(parts * int(self._lu_sizes[0])), np.zeros(parts * int(self._lu_sizes[1]))
        _check(library().ybus_hip_factor_lu_values(self._h, l_val.ctypes.data, u_val.ctypes.data))
        if self._complex:
            return l_val[0::2] + 1j * l_val[1::2], u_val[0::2] + 1j * u_val[1::2]
        return l_val, u_val

    def set_matrix(self, a):
        """The matrix itself (CSR when ``a.row_major`` else CSC), for FP32_REFINED residuals. A
        complex factor takes a complex matrix: a SciPy sparse CSR or CSC matrix, or an object with
        ``row_major``, ``ptr``, ``idx`` and complex ``data``."""
        if self._complex:
            row_major, ptr, idx, data = _complex_compressed(a)
            _check(library().ybus_hip_factor_set_matrix_complex(self._h, int(row_major), _u64(ptr), _u64(idx),
                                                                _interleaved(data)))
        else:
            _check(library().ybus_hip_factor_set_matrix(self._h, int(bool(a.row_major)), _u64(a.ptr), _u64(a.idx),
                                                        _f64(a.data)))
        self._nnz_a = int(np.asarray(a.indptr if hasattr(a, "indptr") else a.ptr)[-1])
        self._generation += 1

    def _release_block(self):
        if self._block is not None:
            self._block.free()

    def _solve_complex(self, rhs, transpose: bool) -> np.ndarray:
        b = _complex_rhs(rhs, self._n)
        x = _result_array(b.shape, b.dtype)
        if b.shape[0]:
            fn = (library().ybus_hip_factor_solve_transpose_many_complex if transpose
                  else library().ybus_hip_factor_solve_many_complex)
            _check(fn(self._h, b.view(np.float64), b.shape[0], x.view(np.float64)))
        return x

    def solve_many(self, rhs) -> np.ndarray:
        """A x = b for each right-hand side, FP64, as a (count, n) array (factor.rs:234-289);
        complex128 arrays for a complex factor."""
        if self._complex:
            return self._solve_complex(rhs, False)
        b = _rhs_array(rhs, self._n)
        self._release_block()
        x = _result_array(b.shape)
        if b.shape[0]:
            _check(library().ybus_hip_factor_solve_many(self._h, b, b.shape[0], x))
        return x

    def solve_transpose_many(self, rhs) -> np.ndarray:
        """A^T x = b for each right-hand side, FP64 (Lu::solve_transpose, lu.rs:502-533); for a
        complex factor the plain transpose, complex128 arrays."""
        if self._complex:
            return self._solve_complex(rhs, True)
        b = _rhs_array(rhs, self._n)
        self._release_block()
        x = _result_array(b.shape)
        if b.shape[0]:
            _check(library().ybus_hip_factor_solve_transpose_many(self._h, b, b.shape[0], x))
        return x

    def solve_many_refined(self, rhs, tolerance: float = 1e-12, max_steps: int = 10):
        """FP32 solves refined in FP64 to `tolerance` (refine_normwise's rule,
        accuracy.rs:83-122): (solutions, steps per column, relative residual per column); complex
        arrays for a complex factor (its complex matrix given to set_matrix)."""
        if self._complex:
            bz = _complex_rhs(rhs, self._n)
            xz = _result_array(bz.shape, bz.dtype)
            steps = np.zeros(bz.shape[0], np.uint64)
            ratio = np.zeros(bz.shape[0])
            if bz.shape[0]:
                _check(library().ybus_hip_factor_solve_many_refined_complex(
                    self._h, bz.view(np.float64), bz.shape[0], float(tolerance), int(max_steps), xz.view(np.float64),
                    steps, ratio))
            return xz, steps.astype(np.int64), ratio
        b = _rhs_array(rhs, self._n)
        self._release_block()
        x = _result_array(b.shape)
        steps = np.zeros(b.shape[0], np.uint64)
        ratio = np.zeros(b.shape[0])
        if b.shape[0]:
            _check(library().ybus_hip_factor_solve_many_refined(self._h, b, b.shape[0], float(tolerance),
                                                                int(max_steps), x, steps, ratio))
        return x, steps.astype(np.int64), ratio

    def adjoint(self, x_bar, x, mode: int = FP64, tolerance: float = 1e-12, max_steps: int = 10):
        """The adjoint of A x = b (needs set_matrix): for vectors x_bar_s (a loss's gradient with
        respect to the solutions x_s), returns (b_bar, a_bar): b_bar_s = A^-T x_bar_s (FP64, or
        FP32_REFINED: FP32 sweeps refined in FP64), and a_bar = -(sum over s of b_bar_s x_s^T) on
        A's stored pattern, in the matrix's storage order."""
        xb = _rhs_array(x_bar, self._n)
        xs = _rhs_array(x, self._n)
        if xb.shape != xs.shape:
            raise ShapeMismatch(xb.shape[0], xs.shape[0])
        self._release_block()
        b_bar = np.zeros_like(xb)
        a_bar = np.zeros(max(self._nnz_a, 0))
        _check(library().ybus_hip_factor_adjoint(self._h, int(mode), float(tolerance), int(max_steps), xb, xs,
                                                 xb.shape[0], b_bar, a_bar))
        return b_bar, a_bar

    # ---- vectors already in device memory: torch tensors (count, dim) of float64 on the GPU ----

    def _device_vectors(self, t, count=None):
        import torch

        want = torch.complex128 if self._complex else torch.float64
        if not isinstance(t, torch.Tensor) or t.dtype != want or not t.is_cuda or t.dim() != 2:
            raise InvalidArgument(INVALID_ARGUMENT, f"a (count, dim) {want} tensor on the GPU")
        if t.shape[1] != self._n:
            raise ShapeMismatch(self._n, t.shape[1])
        if count is not None and t.shape[0] != count:
            raise ShapeMismatch(count, t.shape[0])
        if t.stride(1) != 1:
            raise InvalidArgument(INVALID_ARGUMENT, "each vector must be contiguous")
        return t.data_ptr(), t.stride(0)

    def solve_device(self, b, x=None, transpose: bool = False, mode: int = None, tolerance: float = 1e-12,
                     max_steps: int = 10):
        """Solves the vectors of the GPU tensor b (count, dim) into x (a new tensor unless given),
        without leaving device memory: bit for bit solve_many (or solve_transpose_many), or with
        mode=FP32_REFINED solve_many_refined's solutions."""
        import torch

        pb, ldb = self._device_vectors(b)
        if x is None:
            x = torch.empty_like(b, memory_format=torch.contiguous_format)
        px, ldx = self._device_vectors(x, b.shape[0])
        self._release_block()
        torch.cuda.current_stream().synchronize()
        if mode is None:
            mode = TRANSPOSE if transpose else FP64
        # a complex tensor's strides count complex values, as the complex entry point's do
        fn = library().ybus_hip_factor_solve_device_complex if self._complex else library().ybus_hip_factor_solve_device
        _check(fn(self._h, int(mode), float(tolerance), int(max_steps), pb, ldb, b.shape[0], px, ldx))
        return x

    def solve_device_noalloc(self, b, x, transpose: bool = False):
        """solve_device (FP64, or the transposed solve) into the given GPU tensor x, allocating no
        device memory: the library's workspace is a fixed array of SCRATCH_BYTES per device,
        shared by every such call on it (calls that overlap on one GPU race)."""
        import torch

        pb, ldb = self._device_vectors(b)
        px, ldx = self._device_vectors(x, b.shape[0])
        self._release_block()
        torch.cuda.current_stream().synchronize()
        _check(library().ybus_hip_factor_solve_device_noalloc(self._h, TRANSPOSE if transpose else FP64, pb, ldb,
                                                                b.shape[0], px, ldx))
        return x

    def adjoint_device(self, x_bar, x, mode: int = FP64, tolerance: float = 1e-12, max_steps: int = 10):
        """adjoint() on GPU tensors: returns (b_bar, a_bar) as new GPU tensors."""
        import torch

        if self._nnz_a < 0:
            raise InvalidArgument(INVALID_ARGUMENT, "the adjoint needs the matrix (set_matrix)")
        x_bar = x_bar.contiguous()
        x = x.contiguous()
        pxb, ld = self._device_vectors(x_bar)
        px, ldx = self._device_vectors(x, x_bar.shape[0])
        b_bar = torch.empty_like(x_bar)
        a_bar = torch.empty(max(self._nnz_a, 1), dtype=torch.float64, device=x_bar.device)
        self._release_block()
        torch.cuda.current_stream().synchronize()
        _check(library().ybus_hip_factor_adjoint_device(self._h, int(mode), float(tolerance), int(max_steps), pxb,
                                                          px, ld, x_bar.shape[0], b_bar.data_ptr(), a_bar.data_ptr()))
        return b_bar, a_bar[: self._nnz_a]

    def stage(self, rhs):
        """Copies right-hand sides (at most block_width()) to the device without solving."""
        b = _rhs_array(rhs, self._n)
        self._release_block()
        _check(library().ybus_hip_factor_stage(self._h, b, b.shape[0]))

    def stage_pairs(self, pairs: Sequence[Pair]):
        """Copies pairs (at most block_width()) to the device; solve_staged(FP64) then solves
        them as pair_solves does."""
        self._release_block()
        plus, minus = _pairs(pairs)
        _check(library().ybus_hip_factor_stage_pairs(self._h, plus, minus, len(plus)))

    def solve_staged(self, mode: int = FP64, tolerance: float = 1e-12, max_steps: int = 10) -> "GpuBlock":
        """Solves the staged block on the device, keeping the solutions there."""
        self._release_block()
        h = P()
        _check(library().ybus_hip_factor_solve_staged(self._h, int(mode), float(tolerance), int(max_steps),
                                                      C.byref(h)))
        return self._new_block(h)

    def _new_block(self, h) -> "GpuBlock":
        block = GpuBlock(self, h)
        self._block = block
        return block

    def pair_solves(self, pairs: Sequence[Pair]) -> "GpuBlock":
        """A y = e_plus - e_minus for each pair, kept on the device (pairs.rs:194-202)."""
        self._release_block()
        plus, minus = _pairs(pairs)
        h = P()
        _check(library().ybus_hip_factor_pair_solves(self._h, plus, minus, len(plus), C.byref(h)))
        return self._new_block(h)

    def pair_updates(self, base, updates: Sequence[PairUpdate]) -> "GpuBlock":
        """pair_solves, then each column's rank-one update of `base` (pairs.rs:209-253)."""
        self._release_block()
        base = _f64(base)
        if base.shape[0] != self._n:
            raise InvalidArgument(INVALID_ARGUMENT, "one base value per unknown")
        plus, minus = _pairs([u.pair for u in updates])
        alpha = _f64([u.alpha for u in updates])
        weight = _f64([u.inverse_weight for u in updates])
        h = P()
        _check(library().ybus_hip_factor_pair_updates(self._h, base, plus, minus, alpha, weight, len(plus),
                                                      C.byref(h)))
        return self._new_block(h)

    def pair_updates_arrays(self, base, plus, minus, alpha, inverse_weight) -> "GpuBlock":
        """pair_updates with the columns as arrays: plus and minus (-1 for none), alpha and
        inverse_weight, one entry per column."""
        self._release_block()
        base = _f64(base)
        if base.shape[0] != self._n:
            raise InvalidArgument(INVALID_ARGUMENT, "one base value per unknown")
        plus = np.ascontiguousarray(plus, dtype=np.int64)
        minus = np.ascontiguousarray(minus, dtype=np.int64)
        alpha, weight = _f64(alpha), _f64(inverse_weight)
        if not (len(plus) == len(minus) == len(alpha) == len(weight)):
            raise InvalidArgument(INVALID_ARGUMENT, "one plus, minus, alpha and inverse_weight per column")
        h = P()
        _check(library().ybus_hip_factor_pair_updates(self._h, base, plus, minus, alpha, weight, len(plus),
                                                      C.byref(h)))
        return self._new_block(h)

    def outage_updates(self, base, sets: Sequence[Sequence[PairUpdate]]) -> "GpuBlock":
        """Each set of branches taken out together (a rank-m update of `base`, m = the set's
        size, at most 64): a block of one column per set. The sets' pairs are solved as one block
        (at most block_width() pairs in all)."""
        flat = [u for group in sets for u in group]
        set_ptr = np.zeros(len(sets) + 1, dtype=np.uint64)
        set_ptr[1:] = np.cumsum([len(group) for group in sets]) if len(sets) else []
        plus, minus = _pairs([u.pair for u in flat])
        return self.outage_updates_arrays(base, set_ptr, plus, minus, [u.alpha for u in flat],
                                          [u.inverse_weight for u in flat])

    def outage_updates_arrays(self, base, set_ptr, plus, minus, alpha, inverse_weight) -> "GpuBlock":
        """outage_updates with the sets as arrays: set s is pairs set_ptr[s] .. set_ptr[s + 1] - 1
        of plus and minus (-1 for none), alpha and inverse_weight."""
        self._release_block()
        base = _f64(base)
        if base.shape[0] != self._n:
            raise InvalidArgument(INVALID_ARGUMENT, "one base value per unknown")
        set_ptr = _u64(set_ptr)
        plus = np.ascontiguousarray(plus, dtype=np.int64)
        minus = np.ascontiguousarray(minus, dtype=np.int64)
        alpha, weight = _f64(alpha), _f64(inverse_weight)
        if len(set_ptr) < 1 or int(set_ptr[-1]) != len(plus) or not (len(plus) == len(minus) == len(alpha) == len(weight)):
            raise InvalidArgument(INVALID_ARGUMENT, "set_ptr must end at the number of pairs")
        h = P()
        _check(library().ybus_hip_factor_outage_updates(self._h, base, set_ptr, len(set_ptr) - 1, plus, minus, alpha,
                                                        weight, C.byref(h)))
        return self._new_block(h)

    def operator(self, matrix) -> "GpuOperator":
        """A CSR matrix with one column per unknown, applied to blocks (pairs.rs:112-149)."""
        if matrix.cols != self._n:
            raise InvalidArgument(INVALID_ARGUMENT, "one column per unknown")
        h = P()
        _check(library().ybus_hip_factor_operator(self._h, int(matrix.rows), _u64(matrix.ptr), _u64(matrix.idx),
                                                  _f64(matrix.data), C.byref(h)))
        return GpuOperator(self, h, int(matrix.rows))


class GpuBlock:
    """Solutions on the device, one column per right-hand side; it borrows its factor until
    freed (pairs.rs:88-96)."""

    def __init__(self, factor: GpuFactor, handle):
        self._factor = factor
        self._h = handle

    def free(self):
        if self._h:
            library().ybus_hip_block_free(self._h)
            self._h = None
            if self._factor._block is self:
                self._factor._block = None

    def __del__(self, _ending=sys.is_finalizing):
        if not _ending():
            self.free()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def _live(self):
        if not self._h:
            raise InvalidArgument(INVALID_ARGUMENT, "the block was freed")
        return self._h

    def columns(self) -> int:
        return int(library().ybus_hip_block_columns(self._live()))

    def singular(self):
        out = np.zeros(self.columns(), np.uint8)
        _check(library().ybus_hip_block_singular(self._live(), out))
        return [bool(v) for v in out]

    def refinement(self):
        k = self.columns()
        steps, ratio = np.zeros(k, np.uint64), np.zeros(k)
        _check(library().ybus_hip_block_refinement(self._live(), steps, ratio))
        return steps.astype(np.int64), ratio

    def solutions(self) -> np.ndarray:
        out = _result_array((self.columns(), self._factor.dim()))
        _check(library().ybus_hip_block_solutions(self._live(), out))
        return out

    @staticmethod
    def _maxima(arr) -> list:
        return [ColumnMax(None if m.row < 0 else int(m.row), float(m.value), bool(m.not_finite)) for m in arr]

    def maxima(self):
        k = self.columns()
        arr = (_ColumnMax * max(k, 1))()
        _check(library().ybus_hip_block_maxima(self._live(), arr))
        return self._maxima(arr[:k])

    def maxima_arrays(self):
        """maxima() as arrays: (rows, -1 where every product is zero; values; not_finite)."""
        k = self.columns()
        arr = np.zeros(max(k, 1), dtype=_COLUMN_MAX)
        _check(library().ybus_hip_block_maxima(self._live(), arr.ctypes.data_as(C.POINTER(_ColumnMax))))
        arr = arr[:k]
        return arr["row"].copy(), arr["value"].copy(), arr["not_finite"] != 0

    def product_maxima_arrays(self, operator: "GpuOperator", zero_rows):
        """product_maxima() with the rows to zero as an array (-1 for none), its maxima as arrays:
        (rows, values, not_finite)."""
        z = np.ascontiguousarray(zero_rows, dtype=np.int64)
        k = self.columns()
        if len(z) != k:
            raise InvalidArgument(INVALID_ARGUMENT, "one entry to zero (or none) per column")
        arr = np.zeros(max(k, 1), dtype=_COLUMN_MAX)
        _check(library().ybus_hip_block_product_maxima(self._live(), operator._h, z,
                                                        arr.ctypes.data_as(C.POINTER(_ColumnMax))))
        arr = arr[:k]
        return arr["row"].copy(), arr["value"].copy(), arr["not_finite"] != 0

    def product(self, operator: "GpuOperator", zero) -> np.ndarray:
        z = np.array([-1 if r is None else int(r) for r in zero], dtype=np.int64)
        if len(z) != self.columns():
            raise InvalidArgument(INVALID_ARGUMENT, "one entry to zero (or none) per column")
        out = _result_array((self.columns(), operator.rows()))
        _check(library().ybus_hip_block_product(self._live(), operator._h, z, out))
        return out

    def product_maxima(self, operator: "GpuOperator", zero):
        z = np.array([-1 if r is None else int(r) for r in zero], dtype=np.int64)
        if len(z) != self.columns():
            raise InvalidArgument(INVALID_ARGUMENT, "one entry to zero (or none) per column")
        k = self.columns()
        arr = (_ColumnMax * max(k, 1))()
        _check(library().ybus_hip_block_product_maxima(self._live(), operator._h, z, arr))
        return self._maxima(arr[:k])


class GpuOperator:
    """M theta + offset applied to a factor's blocks (pairs.rs:52-86)."""

    def __init__(self, factor: GpuFactor, handle, rows: int):
        self._factor = factor
        self._h = handle
        self._rows = rows

    def __del__(self, _ending=sys.is_finalizing):
        if getattr(self, "_h", None) and not _ending():
            library().ybus_hip_operator_free(self._h)
            self._h = None

    def rows(self) -> int:
        return self._rows

    def set_offset(self, offset):
        offset = _f64(offset)
        if offset.shape[0] != self._rows:
            raise InvalidArgument(INVALID_ARGUMENT, "one offset per row")
        _check(library().ybus_hip_operator_set_offset(self._h, offset))

    def set_weights(self, weights):
        weights = _f64(weights)
        if weights.shape[0] != self._rows:
            raise InvalidArgument(INVALID_ARGUMENT, "one weight per row")
        _check(library().ybus_hip_operator_set_weights(self._h, weights))


# ---- batches of factors (new) ----


class GpuBatch:
    """Many factors solved together: each factor's right-hand sides, all in one persistent launch
    per LDS size class, replayed from a captured HIP graph. The batch keeps its factors alive."""

    def __init__(self, gpu: Gpu, handle):
        self._gpu = gpu
        self._h = handle
        self._factors = []
        self._staged = None

    @staticmethod
    def new(gpu: Gpu) -> "GpuBatch":
        h = P()
        _check(library().ybus_hip_batch_create(gpu._h, C.byref(h)))
        return GpuBatch(gpu, h)

    def __del__(self, _ending=sys.is_finalizing):
        if getattr(self, "_h", None) and not _ending():
            library().ybus_hip_batch_free(self._h)
            self._h = None

    def add(self, factor: GpuFactor):
        _check(library().ybus_hip_batch_add(self._h, factor._h))
        self._factors.append(factor)
        self._staged = None

    def factors(self) -> int:
        return int(library().ybus_hip_batch_factors(self._h))

    def set_refine_stop(self, rule: int):
        """The stop rule of the batch's FP32-refined solves (GpuFactor.set_refine_stop's rules)."""
        _check(library().ybus_hip_batch_set_refine_stop(self._h, int(rule)))

    def stage(self, rhs):
        """rhs: one (count_i, dim_i) array per factor, in the order added."""
        if len(rhs) != len(self._factors):
            raise InvalidArgument(INVALID_ARGUMENT, "one block of right-hand sides per factor")
        blocks = [_rhs_array(r, f.dim()) for r, f in zip(rhs, self._factors)]
        counts = np.array([b.shape[0] for b in blocks], dtype=np.uint64)
        flat = np.ascontiguousarray(np.concatenate([b.ravel() for b in blocks])) if blocks else np.zeros(0)
        _check(library().ybus_hip_batch_stage(self._h, flat.ctypes.data if flat.size else None, counts))
        self._staged = [(int(c), f.dim()) for c, f in zip(counts, self._factors)]

    def solve_staged(self, mode: int = FP64, tolerance: float = 1e-12, max_steps: int = 10):
        """FP64 (the CPU's bits) or FP32_REFINED (each column refined in FP64 to `tolerance`)."""
        _check(library().ybus_hip_batch_solve_staged(self._h, int(mode), float(tolerance), int(max_steps)))

    def refinement(self):
        """After an FP32_REFINED solve: per factor, (steps, ratio) arrays of its columns."""
        if self._staged is None:
            raise InvalidArgument(INVALID_ARGUMENT, "nothing staged")
        total = sum(c for c, _ in self._staged)
        steps = np.zeros(total, np.uint64)
        ratio = np.zeros(total)
        _check(library().ybus_hip_batch_refinement(self._h, steps.ctypes.data if total else None,
                                                   ratio.ctypes.data if total else None))
        out, at = [], 0
        for c, _ in self._staged:
            out.append((steps[at: at + c].astype(np.int64), ratio[at: at + c]))
            at += c
        return out

    def solutions(self):
        if self._staged is None:
            raise InvalidArgument(INVALID_ARGUMENT, "nothing staged")
        total = sum(c * n for c, n in self._staged)
        flat = _result_array((total,))
        _check(library().ybus_hip_batch_solutions(self._h, flat.ctypes.data if total else None))
        out, at = [], 0
        for c, n in self._staged:
            out.append(flat[at: at + c * n].reshape(c, n))
            at += c * n
        return out

    def solve(self, rhs):
        self.stage(rhs)
        self.solve_staged()
        return self.solutions()

    def kernel_stats(self) -> KernelStats:
        s = _KernelStats()
        library().ybus_hip_batch_kernel_stats(self._h, C.byref(s))
        return KernelStats(int(s.launches), float(s.milliseconds), int(s.bytes))

    def reset_kernel_stats(self):
        library().ybus_hip_batch_reset_kernel_stats(self._h)
