/*
 * CPU oracle for the tiled Cholesky factor-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product path (libstarpu_amd*.so) never links or calls it.
 *
 * It restates, in FP64 and plain C, the arithmetic of the reference's CPU
 * codelets (examples/cholesky/cholesky_kernels.c) and the submission order of
 * its DAG (examples/cholesky/cholesky_implicit.c).  All matrices are
 * column-major ("fortran ordering", cholesky_implicit.c:24-25); `ld` is in
 * elements.
 *
 * Parity pins (see tests/test_oracle.py): LAPACK/BLAS fixtures written by
 * tests/golden/make_golden.py, closed forms (2x2 Cholesky, diagonal and
 * identity cases, Hilbert + N*I known-answer test of cholesky_implicit.c:
 * 474-484,568-583), integer-exact GEMM/SYRK and brute-force invariants.
 */
#ifndef CHOL_ORACLE_H
#define CHOL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C <- beta*C + alpha*A*B^T ; A is m x k, B is n x k, C is m x n.
 * Semantics of STARPU_SGEMM("N","T",...) at cholesky_kernels.c:75-76. */
void or_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc);

/* Lower triangle of C (n x n) <- beta*C + alpha*A*A^T ; A is n x k.
 * Semantics of STARPU_SSYRK("L","N",...) at cholesky_kernels.c:169-170.
 * The strict upper triangle of C is not referenced. */
void or_dsyrk_ln(int n, int k, double alpha, const double *A, int lda,
                 double beta, double *C, int ldc);

/* B (m x n) <- alpha * B * L^{-T}, L n x n lower, non-unit diagonal.
 * Semantics of STARPU_STRSM("R","L","T","N",...) at cholesky_kernels.c:254. */
void or_dtrsm_rltn(int m, int n, double alpha, const double *L, int ldl,
                   double *B, int ldb);

/* In-place lower Cholesky of the n x n matrix A; returns 0 or the 1-based
 * column at which a non-positive pivot was met (LAPACK dpotrf's info).
 * Restates the unblocked sqrt/SCAL/SYR loop of cholesky_kernels.c:329-342. */
int or_dpotrf_l(int n, double *A, int lda);

/* Tiled right-looking Cholesky of the N x N matrix A (lower) split into
 * nblocks x nblocks tiles, executing the POTRF/TRSM/SYRK/GEMM tasks in the
 * order cholesky_implicit.c:244-318 submits them.  Tile sizes follow the
 * block filter's chunk rule (src/datawizard/filters.c:1655-1679): the first
 * N % nblocks tiles are one larger.
 * Returns 0 or the global 1-based column of the first non-positive pivot. */
int or_tiled_cholesky(int N, int nblocks, double *A, int lda);

/* Solve A X = B with A = L L^T (LAPACK dpotrs 'L'): B (n x nrhs) <- X. */
void or_dpotrs_l(int n, int nrhs, const double *L, int ldl, double *B, int ldb);

/* ---- solve phase (no counterpart in the reference, whose Cholesky
 * example stops after the factorisation): the BLAS operations LAPACK
 * dpotrs applies, restated from the published reference BLAS loops. ---- */

/* B (m x n) <- alpha * L^{-1} B (dtrsm Left, Lower, NoTrans, NonUnit). */
void or_dtrsm_llnn(int m, int n, double alpha, const double *L, int ldl, double *B, int ldb);
/* B (m x n) <- alpha * L^{-T} B (dtrsm Left, Lower, Trans, NonUnit). */
void or_dtrsm_lltn(int m, int n, double alpha, const double *L, int ldl, double *B, int ldb);
/* C (m x n) <- beta*C + alpha*A*B, A m x k, B k x n (dgemm N,N). */
void or_dgemm_nn(int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc);
/* C (m x n) <- beta*C + alpha*A^T*B, A k x m, B k x n (dgemm T,N). */
void or_dgemm_tn(int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc);
/* Tiled solve of L L^T X = B with the factor of or_tiled_cholesky (same
 * nblocks tiling of the rows): B (N x nrhs) <- X.  Executes the tile tasks
 * in the order the solve DAG submits them: forward, for k = 0..nt-1,
 * TRSM_LN(L_kk, b_k) then GEMM_NN b_m -= L_mk b_k (m > k); backward, for
 * k = nt-1..0, TRSM_LT(L_kk, b_k) then GEMM_TN b_m -= L_km^T b_k (m < k). */
void or_tiled_potrs(int N, int nblocks, int nrhs, const double *L, int ldl, double *B, int ldb);

/* FP32 factor with FP64 iterative refinement (the scheme of LAPACK
 * dsposv; Langou et al., SC'06): lower(A) rounded to FP32 and factored by
 * the unblocked loop of cholesky_kernels.c:329-342 in FP32 -- the
 * precision of the reference's own codelets; then per right-hand side
 * x = 0, r = b, repeat d = (L L^T)^{-1} fl32(r) (FP32 substitutions),
 * x += d, r = b - A x in FP64, until ||r||_2 / (||A||_F ||x||_2 + ||b||_2)
 * <= tol or maxiter steps.  X (n x nrhs) receives x, iters[c] / berr[c]
 * the steps and final backward error of column c.  Returns 0 or the
 * 1-based column of a non-positive FP32 pivot (X untouched, iters 0,
 * berr +inf). */
int or_sposv_ir(int n, int nrhs, const double *A, int lda, const double *B, int ldb,
                double *X, int ldx, int maxiter, double tol, int *iters, double *berr);

/* R <- B - A X for symmetric A given by its lower triangle (n x n). */
void or_residual_sym_l(int n, int nrhs, const double *A, int lda,
                       const double *X, int ldx, const double *B, int ldb,
                       double *R, int ldr);

/* ||A - L L^T||_F / ||A||_F over the lower triangle (A given by its lower). */
double or_backward_error_llt(int n, const double *A, int lda,
                             const double *L, int ldl);
/* ||L - Lref||_F / ||Lref||_F over the lower triangle. */
double or_rel_diff_lower(int n, const double *L, int ldl,
                         const double *Lref, int ldr);

#ifdef __cplusplus
}
#endif
#endif
