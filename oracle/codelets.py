"""Addresses of the CPU tile codelets, for installing them as the Cholesky
DAG's CPU implementations.  TEST INFRASTRUCTURE ONLY (tests and the
CPU-baseline leg of bench.py).

* ``addresses()``: the oracle's plain-C codelets (oracle/oracle_codelets.c).
* ``blas_addresses()``: the reference's CPU codelets on an external BLAS
  (oracle/blas_codelets.c), resolved from scipy's bundled OpenBLAS."""
import ctypes as C
import glob
import os

from . import lib

KINDS = ("potrf", "trsm", "syrk", "gemm", "solve_fwd", "solve_bwd", "trtri")


def addresses() -> dict:
    L = lib()
    return {k: C.cast(getattr(L, f"or_cl_{k}"), C.c_void_p).value for k in KINDS}


def blas_library() -> str | None:
    """Path of scipy's bundled OpenBLAS (LP64, scipy_-prefixed entry points), or None."""
    try:
        import scipy
    except ImportError:
        return None
    base = os.path.dirname(os.path.dirname(scipy.__file__))
    hits = sorted(glob.glob(os.path.join(base, "scipy.libs", "libscipy_openblas*.so*")))
    return hits[0] if hits else None


def blas_addresses(path: str | None = None) -> dict:
    """Loads the BLAS library (one thread) and returns the BLAS codelets' addresses."""
    L = lib()
    path = path or blas_library()
    if not path:
        raise RuntimeError("no BLAS library found (scipy's bundled OpenBLAS)")
    L.orb_load.restype = C.c_int
    L.orb_load.argtypes = [C.c_char_p]
    if L.orb_load(path.encode()) != 0:
        raise RuntimeError(f"{path}: dgemm/dsyrk/dtrsm/dpotrf not found")
    return {k: C.cast(getattr(L, f"orb_cl_{k}"), C.c_void_p).value for k in KINDS}


def blas_dgemm_nt(alpha, A, B, beta, Cm):
    """C <- beta C + alpha A B^T through the loaded BLAS (after blas_addresses)."""
    import numpy as np
    A, B = np.asfortranarray(A), np.asfortranarray(B)
    Cm = np.array(Cm, order="F", dtype=float)
    L = lib()
    L.orb_dgemm_nt.restype = C.c_int
    L.orb_dgemm_nt.argtypes = [C.c_int] * 3 + [C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                               C.c_double, C.c_void_p, C.c_int]
    m, k = A.shape
    r = L.orb_dgemm_nt(m, B.shape[0], k, alpha, A.ctypes.data, max(1, m), B.ctypes.data, max(1, B.shape[0]),
                       beta, Cm.ctypes.data, max(1, m))
    if r:
        raise RuntimeError("BLAS not loaded")
    return Cm
