"""ctypes binding of the gfx950 tile-kernel C ABI (include/starpu_amd_kernels.h).

The library is loaded from the package directory; a missing library raises
instead of falling back to anything else.  Device memory is handled through
the library's own helpers (``DeviceArray``), so no other GPU runtime is
involved.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_double_p = C.c_void_p  # device pointers are passed as raw addresses


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libstarpu_amd_kernels.so")
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(path)
        vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
        sig = {
            "sa_dgemm_nt": (i, [vp, i, i, i, d, vp, i, vp, i, d, vp, i]),
            "sa_dsyrk_ln": (i, [vp, i, i, d, vp, i, d, vp, i]),
            "sa_dtrsm_rltn": (i, [vp, i, i, d, vp, i, vp, i, vp, sz]),
            "sa_dtrsm_rltn_worksize": (sz, [i]),
            "sa_dpotrf_l": (i, [vp, i, vp, i, vp, vp, sz]),
            "sa_dpotrf_tile": (i, [vp, i, vp, i, vp, i, vp, i, vp, vp, sz]),
            "sa_dpotrf_tile_wg": (i, [vp, i, vp, i, vp, i, vp, i, vp, vp, sz, i]),
            "sa_tile_outputs_f32": (i, [vp, i, vp, i, vp, i, vp, i, vp, i]),
            "sa_dpotrf_tile_worksize": (sz, [i]),
            "sa_dpotrf_tile_noalloc": (i, [vp, i, vp, i, vp, i, vp, i, vp]),
            "sa_dpotrf_tile_noalloc_maxn": (i, []),
            "sa_dpotrf_tile_maxn": (i, []),
            "sa_dtrtri_tile": (i, [vp, i, vp, i, vp, i, vp, i, vp, sz]),
            "sa_dpotrf_l_worksize": (sz, [i]),
            "sa_dtrtri_diag64": (i, [vp, i, vp, i, vp]),
            "sa_dtrsm_llnn": (i, [vp, i, i, d, vp, i, vp, i]),
            "sa_dtrsm_lltn": (i, [vp, i, i, d, vp, i, vp, i, vp, sz]),
            "sa_dtrsm_lltn_worksize": (sz, [i]),
            "sa_dgemm_nn": (i, [vp, i, i, i, d, vp, i, vp, i, d, vp, i, vp, sz]),
            "sa_dgemm_tn": (i, [vp, i, i, i, d, vp, i, vp, i, d, vp, i, vp, sz]),
            "sa_dgemm_thin_worksize": (sz, [i, i, i]),
            "sa_dtrtri_lt": (i, [vp, i, vp, i, vp, i, vp, sz]),
            "sa_dtrtri_lt_worksize": (sz, [i]),
            "sa_dtrmm_lu": (i, [vp, i, i, i, d, vp, i, vp, i, vp, sz]),
            "sa_dtrmm_lu_worksize": (sz, [i, i]),
            "sa_dlaset": (i, [vp, i, i, d, d, vp, i]),
            "sa_dtranspose": (i, [vp, i, i, vp, i, vp, i]),
            "sa_sgemm_tn": (i, [vp, i, i, i, i, i, C.c_float, vp, i, vp, i, C.c_float, vp, i]),
            "sa_sgemm_tn_split": (i, [vp, i, i, i, i, C.c_float, vp, vp, i, vp, vp, i, C.c_float, vp, i]),
            "sa_sgemm_tn_split_ok": (i, [i, i, i, vp, vp, i, vp, vp, i]),
            "sa_set_sgemm_direct": (i, [i]),
            "sa_sgemm_tn_utri": (i, [vp, i, i, i, C.c_float, vp, i, vp, i, C.c_float, vp, i]),
            "sa_dgemm_nt_ltri": (i, [vp, i, i, d, vp, i, vp, i, d, vp, i]),
            "sa_split_bf16x2": (i, [vp, i, i, vp, i, vp, vp, i]),
            "sa_dlag2s_lower_to_upper_fro_worksize": (sz, [i]),
            "sa_dlag2s_lower_to_upper_fro": (i, [vp, i, vp, i, vp, i, vp, vp, sz]),
            "sa_dlag2s_lower_to_upper": (i, [vp, i, vp, i, vp, i]),
            "sa_slag2d_upper_to_lower": (i, [vp, i, vp, i, vp, i]),
            "sa_dlag2s": (i, [vp, i, i, vp, i, vp, i]),
            "sa_daxpy2d": (i, [vp, i, i, d, vp, i, vp, i]),
            "sa_gemm_nn_sd": (i, [vp, i, i, i, d, vp, i, vp, i, d, vp, i, vp, sz]),
            "sa_gemm_tn_sd": (i, [vp, i, i, i, d, vp, i, vp, i, d, vp, i, vp, sz]),
            "sa_dsymm_lower_resid_worksize": (sz, [i, i]),
            "sa_dsymm_lower_resid": (i, [vp, i, i, vp, i, vp, i, vp, i, vp, i, vp, sz]),
            "sa_trmm_lu_sd": (i, [vp, i, i, i, d, vp, i, vp, i, vp, sz]),
            "sa_set_launch_concurrency": (i, [i]),
            "sa_set_gemm_shape": (i, [i]),
            "sa_gemm_tile_shape": (i, [i, i, i]),
            "sa_device_count": (i, []),
            "sa_set_device": (i, [i]),
            "sa_dev_malloc": (vp, [sz]),
            "sa_dev_free": (i, [vp]),
            "sa_memcpy_h2d": (i, [vp, vp, sz]),
            "sa_memcpy_d2h": (i, [vp, vp, sz]),
            "sa_memcpy_d2d": (i, [vp, vp, sz]),
            "sa_stream_create": (i, [C.POINTER(vp)]),
            "sa_stream_destroy": (i, [vp]),
            "sa_stream_sync": (i, [vp]),
            "sa_device_sync": (i, []),
            "sa_error_string": (C.c_char_p, [i]),
            "sa_timer_start": (i, [vp, C.POINTER(vp)]),
            "sa_timer_stop": (i, [vp, C.POINTER(C.c_float)]),
            "sa_fill_random": (i, [vp, i, i, vp, i, C.c_uint64, i]),
            "sa_add_diagonal": (i, [vp, i, vp, i, d]),
            "sa_symmetrize_lower": (i, [vp, i, vp, i]),
            "sa_dlacpy": (i, [vp, i, i, vp, i, vp, i]),
            "sa_slacpy": (i, [vp, i, i, vp, i, vp, i]),
            "sa_dcolsumsq": (i, [vp, i, i, vp, i, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int, what: str = "call") -> None:
    if rc != 0:
        msg = lib().sa_error_string(rc).decode() if rc > 0 else os.strerror(-rc)
        raise RuntimeError(f"{what} failed with {rc}: {msg}")


def device_count() -> int:
    return lib().sa_device_count()


class DeviceArray:
    """Column-major device matrix (m x n, leading dimension ld) of one dtype."""

    def __init__(self, m: int, n: int, dtype=np.float64, ld: int | None = None):
        self.m, self.n = int(m), int(n)
        self.ld = int(ld) if ld is not None else max(1, self.m)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.ld * max(1, self.n) * self.dtype.itemsize
        p = lib().sa_dev_malloc(self.nbytes)
        if not p:
            raise MemoryError(f"device allocation of {self.nbytes} bytes failed")
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray, ld: int | None = None) -> "DeviceArray":
        a = np.asarray(a)
        m, n = (a.shape[0], a.shape[1]) if a.ndim == 2 else (a.shape[0], 1)
        d = cls(m, n, a.dtype, ld)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.asarray(a, dtype=self.dtype)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        host = np.zeros((self.ld, self.n), dtype=self.dtype, order="F")
        host[: a.shape[0], : a.shape[1]] = a
        check(lib().sa_memcpy_h2d(self.ptr, host.ctypes.data, self.nbytes), "h2d")

    def download(self) -> np.ndarray:
        host = np.empty((self.ld, self.n), dtype=self.dtype, order="F")
        check(lib().sa_memcpy_d2h(host.ctypes.data, self.ptr, self.nbytes), "d2h")
        return np.asfortranarray(host[: self.m, :])

    def at(self, i: int, j: int) -> int:
        """Device address of element (i, j)."""
        return self.ptr + (i + j * self.ld) * self.dtype.itemsize

    def free(self) -> None:
        if self.ptr:
            lib().sa_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceInt:
    """One device-resident int (e.g. a POTRF info flag)."""

    def __init__(self, value: int = -1):
        self.ptr = lib().sa_dev_malloc(4)
        if not self.ptr:
            raise MemoryError("device allocation failed")
        self.set(value)

    def set(self, value: int) -> None:
        v = C.c_int(value)
        check(lib().sa_memcpy_h2d(self.ptr, C.addressof(v), 4), "h2d")

    def get(self) -> int:
        v = C.c_int(0)
        check(lib().sa_memcpy_d2h(C.addressof(v), self.ptr, 4), "d2h")
        return v.value

    def __del__(self):
        try:
            if self.ptr:
                lib().sa_dev_free(self.ptr)
        except Exception:
            pass


def sync() -> None:
    check(lib().sa_device_sync(), "device sync")


# -- kernel entry points (stream None = default stream) ---------------------

def set_launch_concurrency(n: int) -> int:
    """Launches kept in flight per device (tile-shape hint); returns the old value."""
    return lib().sa_set_launch_concurrency(n)


def set_gemm_shape(edge: int) -> int:
    """Force the GEMM workgroup tile (64 or 128), 0 for automatic; returns the old value."""
    r = lib().sa_set_gemm_shape(edge)
    if r < 0:
        check(r, "sa_set_gemm_shape")
    return r


def gemm_tile_shape(lower: bool, m: int, n: int) -> int:
    """Workgroup tile edge an m x n product (lower triangle if ``lower``) runs with now."""
    r = lib().sa_gemm_tile_shape(1 if lower else 0, m, n)
    if r < 0:
        check(r, "sa_gemm_tile_shape")
    return r


def dgemm_nt(m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, stream=None):
    check(lib().sa_dgemm_nt(stream, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc), "sa_dgemm_nt")


def dsyrk_ln(n, k, alpha, A, lda, beta, Cm, ldc, stream=None):
    check(lib().sa_dsyrk_ln(stream, n, k, alpha, A, lda, beta, Cm, ldc), "sa_dsyrk_ln")


class Workspace:
    """Device scratch buffer grown on demand (the kernels never allocate)."""

    def __init__(self):
        self.ptr, self.nbytes = None, 0

    def get(self, nbytes: int) -> int:
        if nbytes > self.nbytes:
            if self.ptr:
                lib().sa_dev_free(self.ptr)
            self.ptr = lib().sa_dev_malloc(nbytes)
            if not self.ptr:
                raise MemoryError("workspace allocation failed")
            self.nbytes = nbytes
        return self.ptr


_WS = None


def _ws(nbytes: int) -> int:
    global _WS
    if _WS is None:
        _WS = Workspace()
    return _WS.get(nbytes)


def dtrsm_rltn(m, n, alpha, L, ldl, B, ldb, stream=None, work=None, work_bytes=None):
    if work is None:
        work_bytes = lib().sa_dtrsm_rltn_worksize(n)
        work = _ws(work_bytes)
    check(lib().sa_dtrsm_rltn(stream, m, n, alpha, L, ldl, B, ldb, work, work_bytes), "sa_dtrsm_rltn")


def dpotrf_l(n, A, lda, info_ptr, stream=None, work=None, work_bytes=None):
    if work is None:
        work_bytes = lib().sa_dpotrf_l_worksize(n)
        work = _ws(work_bytes)
    check(lib().sa_dpotrf_l(stream, n, A, lda, info_ptr, work, work_bytes), "sa_dpotrf_l")


def dpotrf_tile(n, A, lda, Linv, ldl, W, ldw, info_ptr, stream=None, workgroups=0):
    """Factor (and, with Linv and W, invert) one diagonal tile, n <= 2048;
    workgroups > 0 sets the bulk workgroup count (sa_dpotrf_tile_wg)."""
    work_bytes = lib().sa_dpotrf_tile_worksize(n)
    check(lib().sa_dpotrf_tile_wg(stream, n, A, lda, Linv, ldl, W, ldw, info_ptr, _ws(work_bytes), work_bytes,
                                  workgroups), "sa_dpotrf_tile_wg")


def dtrtri_tile(n, L, ldl, Linv, ldx, W, ldw, stream=None):
    work_bytes = lib().sa_dpotrf_tile_worksize(n)
    check(lib().sa_dtrtri_tile(stream, n, L, ldl, Linv, ldx, W, ldw, _ws(work_bytes), work_bytes), "sa_dtrtri_tile")


def dtrsm_llnn(m, n, alpha, L, ldl, B, ldb, stream=None):
    check(lib().sa_dtrsm_llnn(stream, m, n, alpha, L, ldl, B, ldb), "sa_dtrsm_llnn")


def dtrsm_lltn(m, n, alpha, L, ldl, B, ldb, stream=None, work=None, work_bytes=None):
    if work is None:
        work_bytes = lib().sa_dtrsm_lltn_worksize(m)
        work = _ws(work_bytes)
    check(lib().sa_dtrsm_lltn(stream, m, n, alpha, L, ldl, B, ldb, work, work_bytes), "sa_dtrsm_lltn")


def dgemm_nn(m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, stream=None, work=-1, work_bytes=0):
    """work=-1: use the module workspace; work=None: run without a k split."""
    if work == -1:
        work_bytes = lib().sa_dgemm_thin_worksize(m, n, k)
        work = _ws(work_bytes) if work_bytes else None
    check(lib().sa_dgemm_nn(stream, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, work, work_bytes),
          "sa_dgemm_nn")


def dgemm_tn(m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, stream=None, work=-1, work_bytes=0):
    if work == -1:
        work_bytes = lib().sa_dgemm_thin_worksize(m, n, k)
        work = _ws(work_bytes) if work_bytes else None
    check(lib().sa_dgemm_tn(stream, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, work, work_bytes),
          "sa_dgemm_tn")


def dtrtri_lt(n, L, ldl, W, ldw, stream=None):
    work_bytes = lib().sa_dtrtri_lt_worksize(n)
    check(lib().sa_dtrtri_lt(stream, n, L, ldl, W, ldw, _ws(work_bytes), work_bytes), "sa_dtrtri_lt")


def dtrmm_lu(trans, m, n, alpha, U, ldu, B, ldb, stream=None):
    work_bytes = lib().sa_dtrmm_lu_worksize(m, n)
    check(lib().sa_dtrmm_lu(stream, trans, m, n, alpha, U, ldu, B, ldb, _ws(work_bytes), work_bytes),
          "sa_dtrmm_lu")


def colsumsq(m, n, X, ld, stream=None) -> np.ndarray:
    """Squared column norms of a device matrix, returned on the host."""
    out = DeviceArray(max(1, n), 1)
    check(lib().sa_dcolsumsq(stream, m, n, X, ld, out.ptr), "sa_dcolsumsq")
    v = out.download()[:n, 0].copy()
    out.free()
    return v


def sgemm_tn(mode, upper, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, stream=None):
    check(lib().sa_sgemm_tn(stream, mode, upper, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc), "sa_sgemm_tn")
