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
    write(out, "lapack_vectors")
    write(solve_vectors(), "solve_vectors")


def write(out: dict, name: str) -> None:
    np.savez(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".txt"), "w") as f:
        f.write(f"written by make_golden.py with scipy {scipy.__version__}, numpy {np.__version__}\n")
        for k in sorted(out):
            f.write(f"{k} {np.asarray(out[k]).shape}\n")


def solve_vectors() -> dict:
    """Tile operations of the solve phase, from the BLAS routines LAPACK
    dpotrs calls: dtrsm Left/Lower/NoTrans and Left/Lower/Trans (non-unit),
    dgemm N,N and T,N; plus whole dpotrs solves with many right-hand sides."""
    rng = np.random.default_rng(20261202)
    out = {}
    for (m, n, alpha) in [(5, 3, 1.0), (64, 16, 1.0), (100, 7, -0.5), (130, 9, 2.0)]:
        L = np.linalg.cholesky(spd(rng, m, float(m)))
        B = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        key = f"{m}_{n}"
        out[f"trsm_llnn_{key}_L"], out[f"trsm_llnn_{key}_B"] = L, B
        out[f"trsm_llnn_{key}_alpha"] = np.array(alpha)
        out[f"trsm_llnn_{key}_R"] = blas.dtrsm(alpha, L, B.copy(order="F"), side=0, lower=1, trans_a=0, diag=0)
        out[f"trsm_lltn_{key}_R"] = blas.dtrsm(alpha, L, B.copy(order="F"), side=0, lower=1, trans_a=1, diag=0)
    for (m, n, k) in [(7, 3, 5), (64, 16, 64), (70, 9, 33)]:
        A = np.asfortranarray(rng.uniform(-1, 1, (m, k)))
        At = np.asfortranarray(rng.uniform(-1, 1, (k, m)))
        B = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
        Cm = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        key = f"{m}_{n}_{k}"
        out[f"gemm_nn_{key}_A"], out[f"gemm_tn_{key}_A"] = A, At
        out[f"gemm_{key}_B"], out[f"gemm_{key}_C"] = B, Cm
        out[f"gemm_nn_{key}_R"] = blas.dgemm(-1.0, A, B, beta=1.0, c=Cm.copy(order="F"))
        out[f"gemm_tn_{key}_R"] = blas.dgemm(-1.0, At, B, beta=1.0, c=Cm.copy(order="F"), trans_a=True)
    for (n, nrhs) in [(96, 16), (130, 5)]:
        A = spd(rng, n, float(n))
        b = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)))
        out[f"potrs_{n}_{nrhs}_A"], out[f"potrs_{n}_{nrhs}_b"] = A, b
        out[f"potrs_{n}_{nrhs}_x"] = np.asfortranarray(sla.cho_solve((sla.cholesky(A, lower=True), True), b))
    return out


if __name__ == "__main__":
    main()
