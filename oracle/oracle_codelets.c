/*
 * CPU tile codelets built on the oracle (TEST INFRASTRUCTURE ONLY): the
 * StarPU codelet signature around or_dpotrf_l / or_dtrsm_rltn /
 * or_dsyrk_ln / or_dgemm_nt, i.e. the reference's CPU codelets
 * (examples/cholesky/cholesky_kernels.c:49-126,151-203,228-278,303-343)
 * in FP64.  Used to run the tiled Cholesky DAG on CPU workers for the
 * CPU baseline and for CPU-only runtime tests.
 */
#include <starpu.h>

#include "chol_oracle.h"

void or_cl_potrf(void *descr[], void *arg)
{
	(void)arg;
	double *A = (double *)STARPU_MATRIX_GET_PTR(descr[0]);
	int *info = (int *)STARPU_VECTOR_GET_PTR(descr[1]);
	*info = or_dpotrf_l((int)STARPU_MATRIX_GET_NY(descr[0]), A, (int)STARPU_MATRIX_GET_LD(descr[0]));
}

void or_cl_trsm(void *descr[], void *arg)
{
	(void)arg;
	or_dtrsm_rltn((int)STARPU_MATRIX_GET_NX(descr[1]), (int)STARPU_MATRIX_GET_NY(descr[1]), 1.0,
	              (const double *)STARPU_MATRIX_GET_PTR(descr[0]), (int)STARPU_MATRIX_GET_LD(descr[0]),
	              (double *)STARPU_MATRIX_GET_PTR(descr[1]), (int)STARPU_MATRIX_GET_LD(descr[1]));
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
