/*
 * Tiled Cholesky factorisation on the starpu_amd runtime.
 *
 * FP64 restatement of the reference's implicit-dependency Cholesky
 * (examples/cholesky/cholesky_implicit.c): register A, split it into
 * nblocks x nblocks tiles with the block and vertical-block filters
 * (:416-450), submit for each k a POTRF on A_kk, TRSMs on the panel, then
 * the SYRK/GEMM trailing updates (:244-318) with the base priorities of
 * :125-213, and time from the first submission to wait_for_all
 * (:233,325-326).
 *
 * The HIP codelets replace the reference's hipBLAS/hipSOLVER calls
 * (examples/cholesky/cholesky_kernels.c:112,190,267,433) with the
 * hand-written gfx950 kernels (starpu_amd_kernels.h), enqueued on the
 * worker's stream (STARPU_HIP_ASYNC).  The reference calls TRSM with
 * (m, n) = (NY, NX) of the panel tile, correct only for square tiles
 * (cholesky_kernels.c:241-242,254); here m is the panel's row count (NX)
 * and n its column count (NY).  POTRF reports its pivot status into a
 * per-step info word, which the reference's HIP path ignores (:430).
 */
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <starpu.h>
#include "starpu_amd_chol.h"
#include "starpu_amd_kernels.h"

static struct chol_cpu_kernels cpu_k;

static void hip_check(int e)
{
	if (e)
		STARPU_HIP_REPORT_ERROR(e);
}

static void hip_potrf(void *descr[], void *arg)
{
	(void)arg;
	double *A = (double *)STARPU_MATRIX_GET_PTR(descr[0]);
	int *info = (int *)STARPU_VECTOR_GET_PTR(descr[1]);
	void *work = (void *)STARPU_VARIABLE_GET_PTR(descr[2]);
	hip_check(sa_dpotrf_l(starpu_hip_get_local_stream(), (int)STARPU_MATRIX_GET_NY(descr[0]), A,
	                      (int)STARPU_MATRIX_GET_LD(descr[0]), info, work,
	                      STARPU_VARIABLE_GET_ELEMSIZE(descr[2])));
}

static void hip_trsm(void *descr[], void *arg)
{
	(void)arg;
	const double *L = (const double *)STARPU_MATRIX_GET_PTR(descr[0]);
	double *B = (double *)STARPU_MATRIX_GET_PTR(descr[1]);
	void *work = (void *)STARPU_VARIABLE_GET_PTR(descr[2]);
	hip_check(sa_dtrsm_rltn(starpu_hip_get_local_stream(), (int)STARPU_MATRIX_GET_NX(descr[1]),
	                        (int)STARPU_MATRIX_GET_NY(descr[1]), 1.0, L, (int)STARPU_MATRIX_GET_LD(descr[0]),
	                        B, (int)STARPU_MATRIX_GET_LD(descr[1]), work,
	                        STARPU_VARIABLE_GET_ELEMSIZE(descr[2])));
}

static void hip_syrk(void *descr[], void *arg)
{
	(void)arg;
	const double *A = (const double *)STARPU_MATRIX_GET_PTR(descr[0]);
	double *C = (double *)STARPU_MATRIX_GET_PTR(descr[1]);
	hip_check(sa_dsyrk_ln(starpu_hip_get_local_stream(), (int)STARPU_MATRIX_GET_NY(descr[1]),
	                      (int)STARPU_MATRIX_GET_NY(descr[0]), -1.0, A, (int)STARPU_MATRIX_GET_LD(descr[0]),
	                      1.0, C, (int)STARPU_MATRIX_GET_LD(descr[1])));
}

static void hip_gemm(void *descr[], void *arg)
{
	(void)arg;
	const double *A = (const double *)STARPU_MATRIX_GET_PTR(descr[0]);
	const double *B = (const double *)STARPU_MATRIX_GET_PTR(descr[1]);
	double *C = (double *)STARPU_MATRIX_GET_PTR(descr[2]);
	hip_check(sa_dgemm_nt(starpu_hip_get_local_stream(), (int)STARPU_MATRIX_GET_NX(descr[2]),
	                      (int)STARPU_MATRIX_GET_NY(descr[2]), (int)STARPU_MATRIX_GET_NY(descr[0]), -1.0, A,
	                      (int)STARPU_MATRIX_GET_LD(descr[0]), B, (int)STARPU_MATRIX_GET_LD(descr[1]), 1.0, C,
	                      (int)STARPU_MATRIX_GET_LD(descr[2])));
}

/* History-based models named as the reference's (cholesky_kernels.c:504-523),
 * with a _d suffix: the FP64 kernels are different codelets. */
static struct starpu_perfmodel model_potrf = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_potrf_d"};
static struct starpu_perfmodel model_trsm = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_trsm_d"};
static struct starpu_perfmodel model_syrk = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_syrk_d"};
static struct starpu_perfmodel model_gemm = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_gemm_d"};

static struct starpu_codelet cl_potrf = {
	.type = STARPU_SEQ,
	.hip_funcs = {hip_potrf},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = 3,
	.modes = {STARPU_RW, STARPU_W | STARPU_NOFOOTPRINT, STARPU_SCRATCH | STARPU_NOFOOTPRINT},
	.model = &model_potrf,
	.name = "cl_POTRF",
	.color = 0xffff00,
};

static struct starpu_codelet cl_trsm = {
	.type = STARPU_SEQ,
	.hip_funcs = {hip_trsm},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = 3,
	.modes = {STARPU_R, STARPU_RW, STARPU_SCRATCH | STARPU_NOFOOTPRINT},
	.model = &model_trsm,
	.name = "cl_TRSM",
	.color = 0x8080ff,
};

static struct starpu_codelet cl_syrk = {
	.type = STARPU_SEQ,
	.max_parallelism = INT_MAX,
	.hip_funcs = {hip_syrk},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = 2,
	.modes = {STARPU_R, STARPU_RW},
	.model = &model_syrk,
	.name = "cl_SYRK",
	.color = 0x00ff00,
};

static struct starpu_codelet cl_gemm = {
	.type = STARPU_SEQ,
	.max_parallelism = INT_MAX,
	.hip_funcs = {hip_gemm},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = 3,
	.modes = {STARPU_R, STARPU_R, STARPU_RW},
	.model = &model_gemm,
	.name = "cl_GEMM",
	.color = 0x00c000,
};

void chol_set_cpu_kernels(const struct chol_cpu_kernels *k)
{
	if (k)
		cpu_k = *k;
	else
		memset(&cpu_k, 0, sizeof(cpu_k));
	cl_potrf.cpu_funcs[0] = cpu_k.potrf;
	cl_trsm.cpu_funcs[0] = cpu_k.trsm;
	cl_syrk.cpu_funcs[0] = cpu_k.syrk;
	cl_gemm.cpu_funcs[0] = cpu_k.gemm;
}

/* FMULS_POTRF + FADDS_POTRF of examples/cholesky/cholesky.h:95-98. */
double chol_flops_potrf(double n)
{
	const double fmuls = n * (((1. / 6.) * n + 0.5) * n + (1. / 3.));
	const double fadds = n * (((1. / 6.) * n) * n - (1. / 6.));
	return fmuls + fadds;
}

/* FLOPS_STRSM(m, n) = FMULS_TRSM + FADDS_TRSM, both FMULS_TRMM (cholesky.h:100-109) */
static double flops_trsm(double m, double n) { return m * n * (n + 1.0); }
/* FLOPS_SSYRK(k, n), cholesky.h:111-114 */
static double flops_syrk(double k, double n) { return k * n * (n + 1.0); }
/* FLOPS_SGEMM(m, n, k), cholesky.h:116-119 */
static double flops_gemm(double m, double n, double k) { return 2.0 * m * n * k; }

/* Base priorities (priority_attribution_p == 0) of cholesky_implicit.c:133-191. */
static int prio_potrf(int nb, int k) { return 2 * nb - 2 * k; }
static int prio_trsm(int nb, int k, int m) { return 2 * nb - 2 * k - m; }
static int prio_gemm(int nb, int k, int m, int n) { return (n == m) ? 2 * nb - 2 * k - n : 2 * nb - 2 * k - m - n; }

int chol_factor(int home_node, double *A, int lda, int N, int nblocks, int prio, struct chol_stats *st)
{
	if (N <= 0 || nblocks <= 0 || nblocks > N || lda < N)
		return -EINVAL;
	if (!starpu_is_initialized())
		return -ENODEV;
	starpu_data_handle_t dataA, dataInfo, wsPotrf, wsTrsm;
	int *infos = calloc((size_t)nblocks, sizeof(int));
	if (!infos)
		return -ENOMEM;
	starpu_matrix_data_register(&dataA, home_node, (uintptr_t)A, (size_t)lda, (size_t)N, (size_t)N, sizeof(double));
	starpu_data_set_name(dataA, "A");
	starpu_vector_data_register(&dataInfo, STARPU_MAIN_RAM, (uintptr_t)infos, (size_t)nblocks, sizeof(int));
	struct starpu_data_filter f = {.filter_func = starpu_matrix_filter_block, .nchildren = (unsigned)nblocks};
	struct starpu_data_filter f2 = {.filter_func = starpu_matrix_filter_vertical_block, .nchildren = (unsigned)nblocks};
	starpu_data_map_filters(dataA, 2, &f, &f2);
	struct starpu_data_filter fi = {.filter_func = starpu_vector_filter_block, .nchildren = (unsigned)nblocks};
	starpu_data_partition(dataInfo, &fi);
	/* per-worker device scratch, as the reference's POTRF codelet takes its
	 * hipSOLVER workspace (cholesky_implicit.c:254-255) */
	const int nbmax = (N + nblocks - 1) / nblocks;
	starpu_variable_data_register(&wsPotrf, -1, 0, sa_dpotrf_l_worksize(nbmax));
	starpu_variable_data_register(&wsTrsm, -1, 0, sa_dtrsm_rltn_worksize(nbmax));

	const int prof_prev = starpu_profiling_status_set(1);
	starpu_trace_reset();
	int ret = 0;
	const double start = starpu_timing_now();
	for (int k = 0; k < nblocks && !ret; k++) {
		starpu_iteration_push((unsigned long)k);
		starpu_data_handle_t akk = starpu_data_get_sub_data(dataA, 2, k, k);
		const double nk = (double)starpu_matrix_get_nx(akk);
		ret = starpu_task_insert(&cl_potrf, STARPU_PRIORITY, prio ? prio_potrf(nblocks, k) : 0,
		                         STARPU_RW, akk, STARPU_W, starpu_data_get_child(dataInfo, (unsigned)k),
		                         STARPU_SCRATCH, wsPotrf, STARPU_FLOPS, chol_flops_potrf(nk), STARPU_NAME, "POTRF", 0);
		for (int m = k + 1; m < nblocks && !ret; m++) {
			starpu_data_handle_t amk = starpu_data_get_sub_data(dataA, 2, m, k);
			ret = starpu_task_insert(&cl_trsm, STARPU_PRIORITY, prio ? prio_trsm(nblocks, k, m) : 0,
			                         STARPU_R, akk, STARPU_RW, amk, STARPU_SCRATCH, wsTrsm,
			                         STARPU_FLOPS, flops_trsm((double)starpu_matrix_get_nx(amk), nk),
			                         STARPU_NAME, "TRSM", 0);
		}
		starpu_data_wont_use(akk);
		for (int n = k + 1; n < nblocks && !ret; n++) {
			starpu_data_handle_t ank = starpu_data_get_sub_data(dataA, 2, n, k);
			starpu_data_handle_t ann = starpu_data_get_sub_data(dataA, 2, n, n);
			const double nn = (double)starpu_matrix_get_nx(ann);
			ret = starpu_task_insert(&cl_syrk, STARPU_PRIORITY, prio ? prio_gemm(nblocks, k, n, n) : 0,
			                         STARPU_R, ank, STARPU_RW, ann,
			                         STARPU_FLOPS, flops_syrk(nk, nn), STARPU_NAME, "SYRK", 0);
			for (int m = n + 1; m < nblocks && !ret; m++) {
				starpu_data_handle_t amk = starpu_data_get_sub_data(dataA, 2, m, k);
				starpu_data_handle_t amn = starpu_data_get_sub_data(dataA, 2, m, n);
				ret = starpu_task_insert(&cl_gemm, STARPU_PRIORITY, prio ? prio_gemm(nblocks, k, m, n) : 0,
				                         STARPU_R, amk, STARPU_R, ank, STARPU_RW, amn,
				                         STARPU_FLOPS,
				                         flops_gemm((double)starpu_matrix_get_nx(amn), nn, nk),
				                         STARPU_NAME, "GEMM", 0);
			}
			starpu_data_wont_use(ank);
		}
		starpu_iteration_pop();
	}
	starpu_task_wait_for_all();
	const double end = starpu_timing_now();

	if (st) {
		memset(st, 0, sizeof(*st));
		st->elapsed_us = end - start;
		st->flops = chol_flops_potrf((double)N);
		static const char *kinds[4] = {"POTRF", "TRSM", "SYRK", "GEMM"};
		const size_t nrec = starpu_trace_count();
		st->ntasks = (long)nrec;
		for (size_t i = 0; i < nrec; i++) {
			struct starpu_trace_record r;
			if (starpu_trace_get(i, &r))
				continue;
			if (starpu_worker_get_type(r.workerid) != STARPU_HIP_WORKER)
				continue;
			for (int q = 0; q < 4; q++)
				if (!strcmp(r.name, kinds[q])) {
					st->kernel_us[q] += r.kernel_us;
					st->kernel_n[q]++;
				}
		}
	}
	starpu_profiling_status_set(prof_prev);

	starpu_data_unpartition(dataInfo, STARPU_MAIN_RAM);
	starpu_data_unregister(dataInfo);
	starpu_data_unregister(wsPotrf);
	starpu_data_unregister(wsTrsm);
	starpu_data_unpartition(dataA, (unsigned)home_node);
	starpu_data_unregister(dataA);
	int info = 0;
	for (int k = 0; k < nblocks && !info; k++)
		if (infos[k]) {
			size_t sz, off;
			starpu_filter_nparts_compute_chunk_size_and_offset((unsigned)N, (unsigned)nblocks, 1, (unsigned)k, 1,
			                                                   &sz, &off);
			info = (int)off + infos[k];
		}
	if (st)
		st->info = info;
	free(infos);
	return ret;
}
