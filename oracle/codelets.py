"""Addresses of the oracle's CPU tile codelets (oracle/oracle_codelets.c),
for installing them as the Cholesky DAG's CPU implementations.
TEST INFRASTRUCTURE ONLY (tests and the CPU-baseline leg of bench.py)."""
import ctypes as C

from . import lib


def addresses() -> dict:
    L = lib()
    return {k: C.cast(getattr(L, f"or_cl_{k}"), C.c_void_p).value
            for k in ("potrf", "trsm", "syrk", "gemm", "solve_fwd", "solve_bwd", "trtri")}
