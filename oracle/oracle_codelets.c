/*
 * CPU tile codelets built on the oracle (TEST INFRASTRUCTURE ONLY): the
 * StarPU codelet signature around or_dpotrf_l / or_dtrsm_rltn /
 * or_dsyrk_ln / or_dgemm_nt, i.e. the reference's CPU codelets
 * (examples/cholesky/cholesky_kernels.c:49-126,151-203,228-278,303-343)
 * in FP64, and around the solve-phase operations (or_dtrsm_llnn,
 * or_dgemm_nn, or_dtrsm_lltn, or_dgemm_tn).  Used to run the tiled
 * Cholesky DAG on CPU workers for the CPU baseline and for CPU-only
 * runtime tests.
 */
#include <starpu.h>

#include "chol_oracle.h"

/* W <- L^{-T} (ld ldw) and Li <- L^{-1} = W^T (ld ldi) for the n x n lower
 * L: the identity solved against L from the right, I L^{-T} */
static void inverse_lt(int n, const double *L, int ldl, double *W, int ldw, double *Li, int ldi)
{
	for (int j = 0; j < n; j++)
		for (int i = 0; i < n; i++)
			W[i + (size_t)j * ldw] = (i == j) ? 1.0 : 0.0;
	or_dtrsm_rltn(n, n, 1.0, L, ldl, W, ldw);
	for (int j = 0; j < n; j++)
		for (int i = 0; i < n; i++)
			Li[i + (size_t)j * ldi] = W[j + (size_t)i * ldw];
}

/* A_kk <- L_kk, with W_k <- L_kk^{-T} and Linv_k <- L_kk^{-1} for the TRSM
 * and solve tasks (RW A_kk, W info_k, W W_k, W Linv_k, scratch) */
void or_cl_potrf(void *descr[], void *arg)
{
	(void)arg;
	double *A = (double *)STARPU_MATRIX_GET_PTR(descr[0]);
	int *info = (int *)STARPU_VECTOR_GET_PTR(descr[1]);
	const int n = (int)STARPU_MATRIX_GET_NY(descr[0]), lda = (int)STARPU_MATRIX_GET_LD(descr[0]);
	*info = or_dpotrf_l(n, A, lda);
	inverse_lt(n, A, lda, (double *)STARPU_MATRIX_GET_PTR(descr[2]), (int)STARPU_MATRIX_GET_LD(descr[2]),
	           (double *)STARPU_MATRIX_GET_PTR(descr[3]), (int)STARPU_MATRIX_GET_LD(descr[3]));
}

/* A_mk <- A_mk L_kk^{-T} by substitution (R L_kk, R Linv_k, RW A_mk, scratch) */
void or_cl_trsm(void *descr[], void *arg)
{
	(void)arg;
	or_dtrsm_rltn((int)STARPU_MATRIX_GET_NX(descr[2]), (int)STARPU_MATRIX_GET_NY(descr[2]), 1.0,
	              (const double *)STARPU_MATRIX_GET_PTR(descr[0]), (int)STARPU_MATRIX_GET_LD(descr[0]),
	              (double *)STARPU_MATRIX_GET_PTR(descr[2]), (int)STARPU_MATRIX_GET_LD(descr[2]));
}

void or_cl_syrk(void *descr[], void *arg)
{
	(void)arg;
	or_dsyrk_ln((int)STARPU_MATRIX_GET_NY(descr[1]), (int)STARPU_MATRIX_GET_NY(descr[0]), -1.0,
	            (const double *)STARPU_MATRIX_GET_PTR(descr[0]), (int)STARPU_MATRIX_GET_LD(descr[0]), 1.0,
	            (double *)STARPU_MATRIX_GET_PTR(descr[1]), (int)STARPU_MATRIX_GET_LD(descr[1]));
}

void or_cl_gemm(void *descr[], void *arg)
{
	(void)arg;
	or_dgemm_nt((int)STARPU_MATRIX_GET_NX(descr[2]), (int)STARPU_MATRIX_GET_NY(descr[2]),
	            (int)STARPU_MATRIX_GET_NY(descr[0]), -1.0, (const double *)STARPU_MATRIX_GET_PTR(descr[0]),
	            (int)STARPU_MATRIX_GET_LD(descr[0]), (const double *)STARPU_MATRIX_GET_PTR(descr[1]),
	            (int)STARPU_MATRIX_GET_LD(descr[1]), 1.0, (double *)STARPU_MATRIX_GET_PTR(descr[2]),
	            (int)STARPU_MATRIX_GET_LD(descr[2]));
}

#define MPTR(i) ((double *)STARPU_MATRIX_GET_PTR(descr[i]))
#define MNX(i) ((int)STARPU_MATRIX_GET_NX(descr[i]))
#define MNY(i) ((int)STARPU_MATRIX_GET_NY(descr[i]))
#define MLD(i) ((int)STARPU_MATRIX_GET_LD(descr[i]))

/* W_k <- L_kk^{-T} and Linv_k <- L_kk^{-1} = W_k^T (R L_kk, W W_k, W Linv_k,
 * scratch) */
void or_cl_trtri(void *descr[], void *arg)
{
	(void)arg;
	inverse_lt(MNY(0), MPTR(0), MLD(0), MPTR(1), MLD(1), MPTR(2), MLD(2));
}

/* Forward sweep of block row k (R L_kk, R W_k, RW b_k, then pairs R L_mk,
 * RW b_m, scratch; cl_arg -> number of pairs): b_k <- L_kk^{-1} b_k by
 * substitution, then b_m -= L_mk b_k tile by tile in increasing m, the
 * order of the oracle's sequential tiled solve. */
void or_cl_solve_fwd(void *descr[], void *arg)
{
	const int np = *(const int *)arg;
	or_dtrsm_llnn(MNX(2), MNY(2), 1.0, MPTR(0), MLD(0), MPTR(2), MLD(2));
	for (int q = 0; q < np; q++) {
		const int t = 3 + 2 * q;
		or_dgemm_nn(MNX(t + 1), MNY(t + 1), MNY(t), -1.0, MPTR(t), MLD(t), MPTR(2), MLD(2), 1.0, MPTR(t + 1),
		            MLD(t + 1));
	}
}

/* Backward sweep of block row k (R L_kk, R W_k, RW b_k, then pairs R L_km,
 * RW b_m, scratch; cl_arg -> number of pairs): b_k <- L_kk^{-T} b_k by
 * substitution, then b_m -= L_km^T b_k in increasing m. */
void or_cl_solve_bwd(void *descr[], void *arg)
{
	const int np = *(const int *)arg;
	or_dtrsm_lltn(MNX(2), MNY(2), 1.0, MPTR(0), MLD(0), MPTR(2), MLD(2));
	for (int q = 0; q < np; q++) {
		const int t = 3 + 2 * q;
		or_dgemm_tn(MNX(t + 1), MNY(t + 1), MNX(t), -1.0, MPTR(t), MLD(t), MPTR(2), MLD(2), 1.0, MPTR(t + 1),
		            MLD(t + 1));
	}
}
