"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, always as the checker.  The product
path never imports this package.

Each wrapper takes and returns numpy arrays in column-major (Fortran)
order; see chol_oracle.h for the reference lines each routine restates.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "or_dgemm_nt": (None, [i, i, i, d, vp, i, vp, i, d, vp, i]),
            "or_dsyrk_ln": (None, [i, i, d, vp, i, d, vp, i]),
            "or_dtrsm_rltn": (None, [i, i, d, vp, i, vp, i]),
            "or_dpotrf_l": (i, [i, vp, i]),
            "or_tiled_cholesky": (i, [i, i, vp, i]),
            "or_dpotrs_l": (None, [i, i, vp, i, vp, i]),
            "or_dtrsm_llnn": (None, [i, i, d, vp, i, vp, i]),
            "or_dtrsm_lltn": (None, [i, i, d, vp, i, vp, i]),
            "or_dgemm_nn": (None, [i, i, i, d, vp, i, vp, i, d, vp, i]),
            "or_dgemm_tn": (None, [i, i, i, d, vp, i, vp, i, d, vp, i]),
            "or_tiled_potrs": (None, [i, i, i, vp, i, vp, i]),
            "or_residual_sym_l": (None, [i, i, vp, i, vp, i, vp, i, vp, i]),
            "or_sposv_ir": (i, [i, i, vp, i, vp, i, vp, i, i, d, vp, vp]),
            "or_backward_error_llt": (d, [i, vp, i, vp, i]),
            "or_rel_diff_lower": (d, [i, vp, i, vp, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _f(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _ld(a: np.ndarray) -> int:
    return max(1, a.shape[0])


def dgemm_nt(alpha, A, B, beta, Cm):
    A, B, Cm = _f(A), _f(B), _f(Cm).copy(order="F")
    m, k = A.shape
    n = B.shape[0]
    lib().or_dgemm_nt(m, n, k, alpha, A.ctypes.data, _ld(A), B.ctypes.data, _ld(B),
                      beta, Cm.ctypes.data, _ld(Cm))
    return Cm


def dsyrk_ln(alpha, A, beta, Cm):
    A, Cm = _f(A), _f(Cm).copy(order="F")
    n, k = A.shape
    lib().or_dsyrk_ln(n, k, alpha, A.ctypes.data, _ld(A), beta, Cm.ctypes.data, _ld(Cm))
    return Cm


def dtrsm_rltn(alpha, L, B):
    L, B = _f(L), _f(B).copy(order="F")
    m, n = B.shape
    lib().or_dtrsm_rltn(m, n, alpha, L.ctypes.data, _ld(L), B.ctypes.data, _ld(B))
    return B


def dpotrf_l(A):
    A = _f(A).copy(order="F")
    info = lib().or_dpotrf_l(A.shape[0], A.ctypes.data, _ld(A))
    return A, info


def tiled_cholesky(A, nblocks):
    A = _f(A).copy(order="F")
    info = lib().or_tiled_cholesky(A.shape[0], nblocks, A.ctypes.data, _ld(A))
    return A, info


def dpotrs_l(L, B):
    L, B = _f(L), _f(B).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    lib().or_dpotrs_l(L.shape[0], B.shape[1], L.ctypes.data, _ld(L), B.ctypes.data, _ld(B))
    return B


def dtrsm_llnn(alpha, L, B):
    L, B = _f(L), _f(B).copy(order="F")
    m, n = B.shape
    lib().or_dtrsm_llnn(m, n, alpha, L.ctypes.data, _ld(L), B.ctypes.data, _ld(B))
    return B


def dtrsm_lltn(alpha, L, B):
    L, B = _f(L), _f(B).copy(order="F")
    m, n = B.shape
    lib().or_dtrsm_lltn(m, n, alpha, L.ctypes.data, _ld(L), B.ctypes.data, _ld(B))
    return B


def dgemm_nn(alpha, A, B, beta, Cm):
    A, B, Cm = _f(A), _f(B), _f(Cm).copy(order="F")
    m, k = A.shape
    lib().or_dgemm_nn(m, B.shape[1], k, alpha, A.ctypes.data, _ld(A), B.ctypes.data, _ld(B),
                      beta, Cm.ctypes.data, _ld(Cm))
    return Cm


def dgemm_tn(alpha, A, B, beta, Cm):
    A, B, Cm = _f(A), _f(B), _f(Cm).copy(order="F")
    k, m = A.shape
    lib().or_dgemm_tn(m, B.shape[1], k, alpha, A.ctypes.data, _ld(A), B.ctypes.data, _ld(B),
                      beta, Cm.ctypes.data, _ld(Cm))
    return Cm


def tiled_potrs(L, B, nblocks):
    L, B = _f(L), _f(B).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    lib().or_tiled_potrs(L.shape[0], nblocks, B.shape[1], L.ctypes.data, _ld(L), B.ctypes.data, _ld(B))
    return B


def residual_sym_l(A, X, B):
    A, X, B = _f(A), _f(X), _f(B)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
        B = B.reshape(-1, 1, order="F")
    R = np.empty_like(B, order="F")
    lib().or_residual_sym_l(A.shape[0], X.shape[1], A.ctypes.data, _ld(A), X.ctypes.data, _ld(X),
                            B.ctypes.data, _ld(B), R.ctypes.data, _ld(R))
    return R


def sposv_ir(A, B, maxiter: int = 10, tol: float = 1e-14):
    """(X, info, iters, berr): FP32 factor + FP64 refinement of A X = B."""
    A, B = _f(A), _f(B)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, r = A.shape[0], B.shape[1]
    X = np.zeros((n, r), order="F")
    it = np.zeros(r, dtype=np.int32)
    be = np.zeros(r)
    info = lib().or_sposv_ir(n, r, A.ctypes.data, _ld(A), B.ctypes.data, _ld(B), X.ctypes.data, max(1, n),
                             maxiter, tol, it.ctypes.data, be.ctypes.data)
    return X, info, it, be


def backward_error_llt(A, L):
    A, L = _f(A), _f(L)
    return lib().or_backward_error_llt(A.shape[0], A.ctypes.data, _ld(A), L.ctypes.data, _ld(L))


def rel_diff_lower(L, Lref):
    L, Lref = _f(L), _f(Lref)
    return lib().or_rel_diff_lower(L.shape[0], L.ctypes.data, _ld(L), Lref.ctypes.data, _ld(Lref))
