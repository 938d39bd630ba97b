"""Writes the LAPACK/BLAS golden vectors the oracle and the GPU kernels are
pinned against (tests/golden/*.npz).

The reference holds no per-kernel numeric vectors for its Cholesky tile
kernels (its only known-answer test is the closed-form Hilbert + N*I
reconstruction of examples/cholesky/cholesky_implicit.c:474-484,568-583), so
the per-kernel expectations come from the LAPACK/BLAS routines the
reference's codelets call (cholesky_kernels.c:75,169,254,321), as bundled
with scipy (OpenBLAS).  Run from the repo root:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np
import scipy
import scipy.linalg as sla
from scipy.linalg import blas

HERE = os.path.dirname(os.path.abspath(__file__))


def spd(rng, n, shift):
    B = rng.uniform(-1.0, 1.0, size=(n, n))
    return np.asfortranarray(B @ B.T + shift * np.eye(n))


def main() -> None:
    rng = np.random.default_rng(20241202)
    out = {}
    # GEMM NT: C <- beta C + alpha A B^T (dgemm 'N','T')
    for (m, n, k) in [(7, 5, 3), (64, 48, 33), (67, 65, 35)]:
        A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
        B = np.asfortranarray(rng.uniform(-1, 1, (n, k)))
        Cm = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        R = blas.dgemm(-1.0, A, B, beta=1.0, c=Cm.copy(order="F"), trans_b=True)
        out[f"gemm_{m}_{n}_{k}_A"], out[f"gemm_{m}_{n}_{k}_B"] = A, B
        out[f"gemm_{m}_{n}_{k}_C"], out[f"gemm_{m}_{n}_{k}_R"] = Cm, R
    # SYRK lower, no trans: lower(C) <- C - A A^T
    for (n, k) in [(5, 3), (70, 40), (65, 33)]:
        A = np.asfortranarray(rng.uniform(-1, 1, (n, k)))
        Cm = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
        R = blas.dsyrk(-1.0, A, beta=1.0, c=Cm.copy(order="F"), lower=1)
        out[f"syrk_{n}_{k}_A"], out[f"syrk_{n}_{k}_C"], out[f"syrk_{n}_{k}_R"] = A, Cm, R
    # TRSM right, lower, transpose, non-unit: B <- B L^{-T}
    for (m, n) in [(4, 3), (65, 64), (30, 70)]:
        L = np.linalg.cholesky(spd(rng, n, float(n)))
        B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        R = blas.dtrsm(1.0, L, B.copy(order="F"), side=1, lower=1, trans_a=1, diag=0)
        out[f"trsm_{m}_{n}_L"], out[f"trsm_{m}_{n}_B"], out[f"trsm_{m}_{n}_R"] = L, B, R
    # POTRF lower and POTRS
    for n in [1, 2, 17, 64, 65, 100]:
        A = spd(rng, n, float(n))
        L = sla.cholesky(A, lower=True)
        b = np.asfortranarray(rng.uniform(-1, 1, (n, 3)))
        x = sla.cho_solve((L, True), b)
        out[f"potrf_{n}_A"], out[f"potrf_{n}_L"] = A, np.asfortranarray(L)
        out[f"potrs_{n}_b"], out[f"potrs_{n}_x"] = b, np.asfortranarray(x)
    # Not positive definite: LAPACK's info for a matrix whose 4th leading
    # minor is negative.
    A = spd(rng, 8, 8.0)
    A[3, 3] = -50.0
    _, info = sla.lapack.dpotrf(A, lower=1)
    out["notspd_A"], out["notspd_info"] = A, np.array(info)
    np.savez(os.path.join(HERE, "lapack_vectors.npz"), **out)
    with open(os.path.join(HERE, "lapack_vectors.txt"), "w") as f:
        f.write(f"written by make_golden.py with scipy {scipy.__version__}, numpy {np.__version__}\n")
        for k in sorted(out):
            f.write(f"{k} {np.asarray(out[k]).shape}\n")


if __name__ == "__main__":
    main()
