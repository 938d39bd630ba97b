/*
 * Tiled Cholesky factorisation and solve on the starpu_amd runtime.
 *
 * The factorisation is the FP64 restatement of the reference's
 * implicit-dependency Cholesky (examples/cholesky/cholesky_implicit.c):
 * register A, split it into nblocks x nblocks tiles with the block and
 * vertical-block filters (:416-450), submit for each k a POTRF on A_kk,
 * TRSMs on the panel, then the SYRK/GEMM trailing updates (:244-318) with
 * the priority attributions of :125-213, and time from the first submission to
 * wait_for_all (:233,325-326).
 *
 * The solve has no counterpart in the reference (its example stops at the
 * factorisation).  It is the tile form of LAPACK dpotrs on the same tile
 * handles: the right-hand sides B (N x nrhs) are split into the same row
 * blocks, then forward substitution (TRSM_LN on b_k, GEMM_NN b_m -= L_mk
 * b_k) and backward substitution (TRSM_LT on b_k, GEMM_TN b_m -= L_km^T
 * b_k) are submitted as tasks.  chol_factor_solve submits both DAGs
 * without a barrier between them: implicit dependencies let the forward
 * solve of the first block rows run while the factorisation proceeds.
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
#include "chol_prio.h"

static struct chol_cpu_kernels cpu_k;

static void hip_check(int e)
{
	if (e)
		STARPU_HIP_REPORT_ERROR(e);
}

#define MPTR(i) ((double *)STARPU_MATRIX_GET_PTR(descr[i]))
#define MNX(i) ((int)STARPU_MATRIX_GET_NX(descr[i]))
#define MNY(i) ((int)STARPU_MATRIX_GET_NY(descr[i]))
#define MLD(i) ((int)STARPU_MATRIX_GET_LD(descr[i]))
#define VPTR(i) ((void *)STARPU_VARIABLE_GET_PTR(descr[i]))
#define VSZ(i) ((size_t)STARPU_VARIABLE_GET_ELEMSIZE(descr[i]))

/* A_kk <- L_kk, W_k <- L_kk^{-T}, Linv_k <- L_kk^{-1} (RW A_kk, W info_k,
 * W W_k, W Linv_k, SCRATCH).  Tiles up to sa_dpotrf_tile_maxn() are factored
 * and inverted in one kernel launch, the inverse's items overlapping the
 * factorisation's critical chain; larger tiles take the multi-launch
 * factorisation and then the inverse by substitution.  A solve on a resident
 * factor forms the inverses with sa_dtrtri_tile (hip_trtri), the same
 * kernels on the same factor, so both give the same bits. */
static void hip_potrf(void *descr[], void *arg)
{
	(void)arg;
	hipStream_t s = starpu_hip_get_local_stream();
	const int n = MNY(0);
	int *info = (int *)STARPU_VECTOR_GET_PTR(descr[1]);
	if (n <= sa_dpotrf_tile_maxn()) {
		hip_check(sa_dpotrf_tile(s, n, MPTR(0), MLD(0), MPTR(3), MLD(3), MPTR(2), MLD(2), info, VPTR(4), VSZ(4)));
		return;
	}
	hip_check(sa_dpotrf_l(s, n, MPTR(0), MLD(0), info, VPTR(4), VSZ(4)));
	hip_check(sa_dtrtri_tile(s, n, MPTR(0), MLD(0), MPTR(3), MLD(3), MPTR(2), MLD(2), VPTR(4), VSZ(4)));
}

/* A_mk <- A_mk L_kk^{-T} as one tile GEMM with the inverse formed by the
 * TRTRI task: A_mk = T Linv_k^T with T a copy of A_mk (R L_kk, R Linv_k,
 * RW A_mk, SCRATCH T).  The flops move from a latency-bound substitution
 * into the MFMA GEMM. */
static void hip_trsm(void *descr[], void *arg)
{
	(void)arg;
	hipStream_t s = starpu_hip_get_local_stream();
	const int m = MNX(2), n = MNY(2);
	double *T = (double *)STARPU_VARIABLE_GET_PTR(descr[3]);
	hip_check(sa_dlacpy(s, m, n, MPTR(2), MLD(2), T, m));
	hip_check(sa_dgemm_nt_ltri(s, m, n, 1.0, T, m, MPTR(1), MLD(1), 0.0, MPTR(2), MLD(2)));
}

static void hip_syrk(void *descr[], void *arg)
{
	(void)arg;
	hip_check(sa_dsyrk_ln(starpu_hip_get_local_stream(), MNY(1), MNY(0), -1.0, MPTR(0), MLD(0), 1.0,
	                      MPTR(1), MLD(1)));
}

static void hip_gemm(void *descr[], void *arg)
{
	(void)arg;
	hip_check(sa_dgemm_nt(starpu_hip_get_local_stream(), MNX(2), MNY(2), MNY(0), -1.0, MPTR(0), MLD(0),
	                      MPTR(1), MLD(1), 1.0, MPTR(2), MLD(2)));
}


/* W_k <- L_kk^{-T}, Linv_k <- L_kk^{-1} = W_k^T (R L_kk, W W_k, W Linv_k,
 * SCRATCH) */
static void hip_trtri(void *descr[], void *arg)
{
	(void)arg;
	hip_check(sa_dtrtri_tile(starpu_hip_get_local_stream(), MNY(0), MPTR(0), MLD(0), MPTR(2), MLD(2), MPTR(1),
	                         MLD(1), VPTR(3), VSZ(3)));
}

/* The solve tasks' buffers: 0 L_kk, 1 W_k, 2 b_k, then pairs (L tile, b_m)
 * from 3, the scratch last; cl_arg points to the number of pairs (see
 * starpu_amd_chol.h). */
static unsigned solve_pairs(const void *arg) { return (unsigned)*(const int *)arg; }

/* True when the pairs' L tiles are consecutive views of one matrix, stacked
 * vertically (`rows`) or side by side, and the b_m consecutive row blocks
 * of one right-hand-side matrix: then one product covers them all. */
static int pairs_contiguous(void *descr[], unsigned np, int rows)
{
	for (unsigned q = 1; q < np; q++) {
		const unsigned t0 = 3 + 2 * (q - 1), t1 = 3 + 2 * q;
		if (MLD(t0) != MLD(t1) || MLD(t0 + 1) != MLD(t1 + 1))
			return 0;
		if (MPTR(t1) != MPTR(t0) + (rows ? (long)MNX(t0) : (long)MNY(t0) * MLD(t0)))
			return 0;
		if (MPTR(t1 + 1) != MPTR(t0 + 1) + MNX(t0 + 1))
			return 0;
	}
	return 1;
}

/* Forward sweep of block row k: b_k <- L_kk^{-1} b_k = W_k^T b_k, then
 * b_m -= L_mk b_k for every m > k.  The column panel below L_kk is one
 * matrix when the tiles are views of a resident A, so one thin GEMM
 * updates every b_m. */
static void hip_solve_fwd(void *descr[], void *arg)
{
	hipStream_t s = starpu_hip_get_local_stream();
	const unsigned np = solve_pairs(arg), sc = 3 + 2 * np;
	const int r = MNY(2), nk = MNX(2);
	hip_check(sa_dtrmm_lu(s, 1, nk, r, 1.0, MPTR(1), MLD(1), MPTR(2), MLD(2), VPTR(sc), VSZ(sc)));
	if (!np)
		return;
	if (pairs_contiguous(descr, np, 1)) {
		int rows = 0;
		for (unsigned q = 0; q < np; q++)
			rows += MNX(3 + 2 * q);
		hip_check(sa_dgemm_nn(s, rows, r, nk, -1.0, MPTR(3), MLD(3), MPTR(2), MLD(2), 1.0, MPTR(4), MLD(4),
		                      VPTR(sc), VSZ(sc)));
		return;
	}
	for (unsigned q = 0; q < np; q++) {
		const unsigned t = 3 + 2 * q;
		hip_check(sa_dgemm_nn(s, MNX(t + 1), r, nk, -1.0, MPTR(t), MLD(t), MPTR(2), MLD(2), 1.0, MPTR(t + 1),
		                      MLD(t + 1), VPTR(sc), VSZ(sc)));
	}
}

/* Backward sweep of block row k: x_k = L_kk^{-T} b_k = W_k b_k, then
 * b_m -= L_km^T x_k for every m < k, over the row panel left of L_kk (one
 * matrix when the tiles are consecutive views). */
static void hip_solve_bwd(void *descr[], void *arg)
{
	hipStream_t s = starpu_hip_get_local_stream();
	const unsigned np = solve_pairs(arg), sc = 3 + 2 * np;
	const int r = MNY(2), nk = MNX(2);
	hip_check(sa_dtrmm_lu(s, 0, nk, r, 1.0, MPTR(1), MLD(1), MPTR(2), MLD(2), VPTR(sc), VSZ(sc)));
	if (!np)
		return;
	if (pairs_contiguous(descr, np, 0)) {
		int cols = 0;
		for (unsigned q = 0; q < np; q++)
			cols += MNY(3 + 2 * q);
		hip_check(sa_dgemm_tn(s, cols, r, nk, -1.0, MPTR(3), MLD(3), MPTR(2), MLD(2), 1.0, MPTR(4), MLD(4),
		                      VPTR(sc), VSZ(sc)));
		return;
	}
	for (unsigned q = 0; q < np; q++) {
		const unsigned t = 3 + 2 * q;
		hip_check(sa_dgemm_tn(s, MNX(t + 1), r, nk, -1.0, MPTR(t), MLD(t), MPTR(2), MLD(2), 1.0, MPTR(t + 1),
		                      MLD(t + 1), VPTR(sc), VSZ(sc)));
	}
}

/* History-based models named as the reference's (cholesky_kernels.c:504-523),
 * with a _d suffix: the FP64 kernels are different codelets. */
static struct starpu_perfmodel model_potrf = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_potrf_d"};
static struct starpu_perfmodel model_trsm = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_trsm_d"};
static struct starpu_perfmodel model_syrk = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_syrk_d"};
static struct starpu_perfmodel model_gemm = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_gemm_d"};
static struct starpu_perfmodel model_trtri = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_trtri_d"};
static struct starpu_perfmodel model_solve_fwd = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_solve_fwd_d"};
static struct starpu_perfmodel model_solve_bwd = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_solve_bwd_d"};

static struct starpu_codelet cl_potrf = {
	.type = STARPU_SEQ,
	.hip_funcs = {hip_potrf},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = 5,
	.modes = {STARPU_RW, STARPU_W | STARPU_NOFOOTPRINT, STARPU_W | STARPU_NOFOOTPRINT, STARPU_W | STARPU_NOFOOTPRINT,
	          STARPU_SCRATCH | STARPU_NOFOOTPRINT},
	.model = &model_potrf,
	.name = "cl_POTRF",
	.color = 0xffff00,
};

static struct starpu_codelet cl_trsm = {
	.type = STARPU_SEQ,
	.hip_funcs = {hip_trsm},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = 4,
	.modes = {STARPU_R, STARPU_R, STARPU_RW, STARPU_SCRATCH | STARPU_NOFOOTPRINT},
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

static struct starpu_codelet cl_trtri = {
	.type = STARPU_SEQ,
	.hip_funcs = {hip_trtri},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = 4,
	.modes = {STARPU_R, STARPU_W, STARPU_W, STARPU_SCRATCH | STARPU_NOFOOTPRINT},
	.model = &model_trtri,
	.name = "cl_TRTRI",
	.color = 0xffc080,
};

static struct starpu_codelet cl_solve_fwd = {
	.type = STARPU_SEQ,
	.hip_funcs = {hip_solve_fwd},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = STARPU_VARIABLE_NBUFFERS,
	.model = &model_solve_fwd,
	.name = "cl_SOLVE_FWD",
	.color = 0xff8080,
};

static struct starpu_codelet cl_solve_bwd = {
	.type = STARPU_SEQ,
	.hip_funcs = {hip_solve_bwd},
	.hip_flags = {STARPU_HIP_ASYNC},
	.nbuffers = STARPU_VARIABLE_NBUFFERS,
	.model = &model_solve_bwd,
	.name = "cl_SOLVE_BWD",
	.color = 0x80ffff,
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
	cl_trtri.cpu_funcs[0] = cpu_k.trtri;
	cl_solve_fwd.cpu_funcs[0] = cpu_k.solve_fwd;
	cl_solve_bwd.cpu_funcs[0] = cpu_k.solve_bwd;
}

/* FMULS_POTRF + FADDS_POTRF of examples/cholesky/cholesky.h:95-98. */
double chol_flops_potrf(double n)
{
	const double fmuls = n * (((1. / 6.) * n + 0.5) * n + (1. / 3.));
	const double fadds = n * (((1. / 6.) * n) * n - (1. / 6.));
	return fmuls + fadds;
}

/* Forward plus backward substitution with an N x N factor: 2 N^2 per
 * right-hand side (FMULS_POTRS + FADDS_POTRS of the LAPACK flop counts). */
double chol_flops_potrs(double n, double nrhs) { return 2.0 * n * n * nrhs; }

/* FLOPS_STRSM(m, n) = FMULS_TRSM + FADDS_TRSM, both FMULS_TRMM (cholesky.h:100-109) */
static double flops_trsm(double m, double n) { return m * n * (n + 1.0); }
/* left-side solve of n right-hand sides against an m x m triangle */
static double flops_trsm_left(double m, double n) { return n * m * m; }
/* FLOPS_SSYRK(k, n), cholesky.h:111-114 */
static double flops_syrk(double k, double n) { return k * n * (n + 1.0); }
/* FLOPS_SGEMM(m, n, k), cholesky.h:116-119 */
static double flops_gemm(double m, double n, double k) { return 2.0 * m * n * k; }

/* Base priorities of the solve (no reference counterpart): the forward
 * sweep ahead of the backward one, block row k ahead of later ones. */
static int prio_solve(const struct chol_prio *p, int nb, int k, int m)
{
	return p->scheme ? 2 * nb - 2 * k - m : 0;
}

static const char *kind_names[CHOL_NKINDS] = {"POTRF", "TRSM", "SYRK", "GEMM", "SOLVE_FWD", "SOLVE_BWD", "TRTRI"};

struct chol_ctx {
	int N, nblocks, nrhs;
	starpu_data_handle_t A, info, B, wsPotrf, wsTrsm, wsTrtri, wsSolve;
	starpu_data_handle_t *W;     /* W_k = L_kk^{-T}, one per diagonal tile */
	starpu_data_handle_t *Linv;  /* Linv_k = L_kk^{-1} */
	int *pairs;                  /* the solve tasks' cl_arg: k + 1 entries 0 .. nblocks-1 */
	int factor;
	int *infos;
	int *dinfos;                 /* device copy of infos when A lives on a GPU */
	int home;
	struct crit_route route;     /* the chain's worker, when the GPU sets one aside */
};

static int ctx_register(struct chol_ctx *c, int home, double *A, int lda, int N, int nblocks,
                        double *B, int ldb, int nrhs, int factor)
{
	memset(c, 0, sizeof(*c));
	c->N = N;
	c->nblocks = nblocks;
	c->nrhs = nrhs;
	c->home = home;
	starpu_matrix_data_register(&c->A, home, (uintptr_t)A, (size_t)lda, (size_t)N, (size_t)N, sizeof(double));
	starpu_data_set_name(c->A, "A");
	struct starpu_data_filter f = {.filter_func = starpu_matrix_filter_block, .nchildren = (unsigned)nblocks};
	struct starpu_data_filter f2 = {.filter_func = starpu_matrix_filter_vertical_block,
	                                .nchildren = (unsigned)nblocks};
	starpu_data_map_filters(c->A, 2, &f, &f2);
	const int nbmax = (N + nblocks - 1) / nblocks;
	if (factor) {
		c->infos = calloc((size_t)nblocks, sizeof(int));
		if (!c->infos)
			return -ENOMEM;
		/* with A on a GPU the tile factorisations' pivot reports stay there
		 * and come back in one copy at the end (a gather to host memory
		 * copies each block's report on its own) */
		if (home != STARPU_MAIN_RAM) {
			c->dinfos = (int *)starpu_malloc_on_node((unsigned)home, (size_t)nblocks * sizeof(int));
			if (!c->dinfos || sa_memcpy_h2d(c->dinfos, c->infos, (size_t)nblocks * sizeof(int)))
				return -ENOMEM;
		}
		starpu_vector_data_register(&c->info, c->dinfos ? home : STARPU_MAIN_RAM,
		                            (uintptr_t)(c->dinfos ? c->dinfos : c->infos), (size_t)nblocks, sizeof(int));
		struct starpu_data_filter fi = {.filter_func = starpu_vector_filter_block, .nchildren = (unsigned)nblocks};
		starpu_data_partition(c->info, &fi);
		/* per-worker device scratch, as the reference's POTRF codelet takes
		 * its hipSOLVER workspace (cholesky_implicit.c:254-255) */
		size_t wp = sa_dpotrf_l_worksize(nbmax);
		if (sa_dpotrf_tile_worksize(nbmax) > wp)
			wp = sa_dpotrf_tile_worksize(nbmax);
		starpu_variable_data_register(&c->wsPotrf, -1, 0, wp);
		starpu_variable_data_register(&c->wsTrsm, -1, 0, (size_t)nbmax * nbmax * sizeof(double));
	}
	c->factor = factor;
	starpu_variable_data_register(&c->wsTrtri, -1, 0, sa_dpotrf_tile_worksize(nbmax));
	c->W = calloc((size_t)nblocks, sizeof(*c->W));
	c->Linv = calloc((size_t)nblocks, sizeof(*c->Linv));
	if (!c->W || !c->Linv)
		return -ENOMEM;
	for (int k = 0; k < nblocks; k++) {
		starpu_data_handle_t akk = starpu_data_get_sub_data(c->A, 2, k, k);
		const size_t nk = starpu_matrix_get_nx(akk);
		starpu_matrix_data_register(&c->W[k], -1, 0, nk, nk, nk, sizeof(double));
		starpu_matrix_data_register(&c->Linv[k], -1, 0, nk, nk, nk, sizeof(double));
	}
	if (B) {
		starpu_matrix_data_register(&c->B, home, (uintptr_t)B, (size_t)ldb, (size_t)N, (size_t)nrhs,
		                            sizeof(double));
		starpu_data_set_name(c->B, "B");
		starpu_data_partition(c->B, &f);
		size_t ws = sa_dtrmm_lu_worksize(nbmax, nrhs);
		/* the split-k partial sums of a panel product, for any panel height
		 * (a smaller workspace would only make the product unsplit) */
		size_t wg = sa_dgemm_thin_worksize(N, nrhs, nbmax);
		for (int rows = 64; rows < N; rows += 64)
			if (sa_dgemm_thin_worksize(rows, nrhs, nbmax) > wg)
				wg = sa_dgemm_thin_worksize(rows, nrhs, nbmax);
		if (wg > ws)
			ws = wg;
		starpu_variable_data_register(&c->wsSolve, -1, 0, ws);
	}
	return 0;
}

/* Expected length of a built task, harmonic mean over the workers that can
 * run it (extract_task_duration_for_model, cholesky_implicit.c:50-66); the
 * task is destroyed.  NaN while some worker's model is uncalibrated. */
static double model_time(struct starpu_task *t)
{
	if (!t)
		return NAN;
	double speed = 0.0;
	int n = 0;
	for (unsigned w = 0; w < starpu_worker_get_count(); w++) {
		if (!starpu_worker_can_execute_task(w, t, 0))
			continue;
		speed += 1.0 / starpu_task_expected_length(t, starpu_worker_get_perf_archtype((int)w, 0), 0);
		n++;
	}
	t->destroy = 0;
	starpu_task_destroy(t);
	return n ? (double)n / speed : NAN;
}

/* Priority attribution for one run; attribution 3 measures the four task
 * kinds on tiles (0,0), (1,0), (1,1) as time_init does (:79-110). */
static void prio_init(struct chol_prio *p, const struct chol_ctx *c, int prio)
{
	memset(p, 0, sizeof(*p));
	p->scheme = (prio < 0 || prio > 4) ? 1 : prio;
	if (p->scheme != 3)
		return;
	if (c->nblocks < 2 || !c->info) {
		p->scheme = 1;
		return;
	}
	starpu_data_handle_t a00 = starpu_data_get_sub_data(c->A, 2, 0, 0);
	starpu_data_handle_t a10 = starpu_data_get_sub_data(c->A, 2, 1, 0);
	starpu_data_handle_t a11 = starpu_data_get_sub_data(c->A, 2, 1, 1);
	p->t_potrf = model_time(starpu_task_build(&cl_potrf, STARPU_RW, a00, STARPU_W,
	                                          starpu_data_get_child(c->info, 0), STARPU_W, c->W[0], STARPU_W,
	                                          c->Linv[0], STARPU_SCRATCH, c->wsPotrf, 0));
	p->t_trsm = model_time(starpu_task_build(&cl_trsm, STARPU_R, a00, STARPU_R, c->Linv[0], STARPU_RW, a10,
	                                         STARPU_SCRATCH, c->wsTrsm, 0));
	p->t_gemm = model_time(starpu_task_build(&cl_gemm, STARPU_R, a10, STARPU_R, a10, STARPU_RW, a11, 0));
	p->t_syrk = model_time(starpu_task_build(&cl_syrk, STARPU_R, a10, STARPU_RW, a11, 0));
	if (!(p->t_potrf > 0) || !(p->t_trsm > 0) || !(p->t_gemm > 0) || !(p->t_syrk > 0))
		p->scheme = 1;
}

/* W_k = L_kk^{-T} and Linv_k = L_kk^{-1} of an already factored tile, for
 * a solve without the factorisation (the POTRF task forms them otherwise). */
static int submit_trtri(struct chol_ctx *c, int k, int prio)
{
	starpu_data_handle_t lkk = starpu_data_get_sub_data(c->A, 2, k, k);
	const double nk = (double)starpu_matrix_get_nx(lkk);
	return starpu_task_insert(&cl_trtri, STARPU_PRIORITY, prio, STARPU_R, lkk, STARPU_W, c->W[k], STARPU_W,
	                          c->Linv[k], STARPU_SCRATCH, c->wsTrtri, STARPU_FLOPS, nk * nk * nk / 3.0,
	                          STARPU_NAME, "TRTRI", 0);
}

static int submit_factor(struct chol_ctx *c, const struct chol_prio *pr)
{
	const int nb = c->nblocks;
	const int crit = c->route.crit;
	uint32_t *others = c->route.others;
	const unsigned olen = c->route.len;
	int ret = 0;
	for (int k = 0; k < nb && !ret; k++) {
		starpu_iteration_push((unsigned long)k);
		starpu_data_handle_t akk = starpu_data_get_sub_data(c->A, 2, k, k);
		const double nk = (double)starpu_matrix_get_nx(akk);
		ret = starpu_task_insert(&cl_potrf, STARPU_PRIORITY, prio_potrf(pr, nb, k),
		                         STARPU_RW, akk, STARPU_W, starpu_data_get_child(c->info, (unsigned)k),
		                         STARPU_W, c->W[k], STARPU_W, c->Linv[k], STARPU_SCRATCH, c->wsPotrf,
		                         STARPU_FLOPS, chol_flops_potrf(nk), STARPU_NAME, "POTRF",
		                         STARPU_EXECUTE_ON_WORKER, crit, 0);
		for (int m = k + 1; m < nb && !ret; m++) {
			starpu_data_handle_t amk = starpu_data_get_sub_data(c->A, 2, m, k);
			const int on_chain = m == k + 1;
			ret = starpu_task_insert(&cl_trsm, STARPU_PRIORITY, prio_trsm(pr, nb, k, m),
			                         STARPU_R, akk, STARPU_R, c->Linv[k], STARPU_RW, amk,
			                         STARPU_SCRATCH, c->wsTrsm,
			                         STARPU_FLOPS, flops_trsm((double)starpu_matrix_get_nx(amk), nk),
			                         STARPU_NAME, "TRSM", STARPU_EXECUTE_ON_WORKER, on_chain ? crit : -1,
			                         STARPU_TASK_WORKERIDS, on_chain ? 0u : olen, others, 0);
		}
		starpu_data_wont_use(akk);
		for (int n = k + 1; n < nb && !ret; n++) {
			starpu_data_handle_t ank = starpu_data_get_sub_data(c->A, 2, n, k);
			starpu_data_handle_t ann = starpu_data_get_sub_data(c->A, 2, n, n);
			const double nn = (double)starpu_matrix_get_nx(ann);
			const int on_chain = n == k + 1;
			ret = starpu_task_insert(&cl_syrk, STARPU_PRIORITY, prio_gemm(pr, nb, k, n, n),
			                         STARPU_R, ank, STARPU_RW, ann,
			                         STARPU_FLOPS, flops_syrk(nk, nn), STARPU_NAME, "SYRK",
			                         STARPU_EXECUTE_ON_WORKER, on_chain ? crit : -1,
			                         STARPU_TASK_WORKERIDS, on_chain ? 0u : olen, others, 0);
			for (int m = n + 1; m < nb && !ret; m++) {
				starpu_data_handle_t amk = starpu_data_get_sub_data(c->A, 2, m, k);
				starpu_data_handle_t amn = starpu_data_get_sub_data(c->A, 2, m, n);
				ret = starpu_task_insert(&cl_gemm, STARPU_PRIORITY, prio_gemm(pr, nb, k, m, n),
				                         STARPU_R, amk, STARPU_R, ank, STARPU_RW, amn,
				                         STARPU_FLOPS,
				                         flops_gemm((double)starpu_matrix_get_nx(amn), nn, nk),
				                         STARPU_NAME, "GEMM", STARPU_TASK_WORKERIDS, olen, others, 0);
			}
			starpu_data_wont_use(ank);
		}
		starpu_iteration_pop();
	}
	return ret;
}

/* Priorities of the solve tasks: forward substitution follows the
 * factorisation's panel order (the TRSM_LN of block k ranks with the first
 * TRSM of step k, its TRTRI just above), backward substitution ranks the
 * later blocks first. */
static int submit_solve(struct chol_ctx *c, const struct chol_prio *pr)
{
	const int nb = c->nblocks;
	const double r = (double)c->nrhs;
	int ret = 0;
	struct starpu_data_descr *d = malloc(sizeof(*d) * (size_t)(2 * nb + 4));
	c->pairs = malloc(sizeof(int) * (size_t)nb);
	if (!d || !c->pairs) {
		free(d);
		return -ENOMEM;
	}
	for (int k = 0; k < nb; k++)
		c->pairs[k] = k;
	/* forward: b_k <- L_kk^{-1} b_k, then b_m -= L_mk b_k below */
	for (int k = 0; k < nb && !ret; k++) {
		starpu_data_handle_t lkk = starpu_data_get_sub_data(c->A, 2, k, k);
		const double nk = (double)starpu_matrix_get_nx(lkk);
		if (!c->factor) {            /* the factorisation forms W_k itself */
			ret = submit_trtri(c, k, pr->scheme ? prio_solve(pr, nb, k, k + 1) + 1 : 0);
			if (ret)
				break;
		}
		int n = 0;
		double below = 0.0;
		d[n++] = (struct starpu_data_descr){lkk, STARPU_R};
		d[n++] = (struct starpu_data_descr){c->W[k], STARPU_R};
		d[n++] = (struct starpu_data_descr){starpu_data_get_sub_data(c->B, 1, k), STARPU_RW};
		for (int m = k + 1; m < nb; m++) {
			starpu_data_handle_t lmk = starpu_data_get_sub_data(c->A, 2, m, k);
			d[n++] = (struct starpu_data_descr){lmk, STARPU_R};
			d[n++] = (struct starpu_data_descr){starpu_data_get_sub_data(c->B, 1, m), STARPU_RW};
			below += (double)starpu_matrix_get_nx(lmk);
		}
		d[n++] = (struct starpu_data_descr){c->wsSolve, STARPU_SCRATCH | STARPU_NOFOOTPRINT};
		ret = starpu_task_insert(&cl_solve_fwd, STARPU_PRIORITY, prio_solve(pr, nb, k, k + 1),
		                         STARPU_DATA_MODE_ARRAY, d, n, STARPU_CL_ARGS_NFREE, &c->pairs[nb - 1 - k], sizeof(int),
		                         STARPU_FLOPS, flops_trsm_left(nk, r) + flops_gemm(below, r, nk),
		                         STARPU_NAME, "SOLVE_FWD", STARPU_TASK_WORKERIDS, c->route.len, c->route.others, 0);
	}
	/* backward: x_k = L_kk^{-T} b_k, then b_m -= L_km^T x_k above */
	for (int k = nb - 1; k >= 0 && !ret; k--) {
		starpu_data_handle_t lkk = starpu_data_get_sub_data(c->A, 2, k, k);
		const double nk = (double)starpu_matrix_get_nx(lkk);
		int n = 0;
		double above = 0.0;
		d[n++] = (struct starpu_data_descr){lkk, STARPU_R};
		d[n++] = (struct starpu_data_descr){c->W[k], STARPU_R};
		d[n++] = (struct starpu_data_descr){starpu_data_get_sub_data(c->B, 1, k), STARPU_RW};
		for (int m = 0; m < k; m++) {
			starpu_data_handle_t bm = starpu_data_get_sub_data(c->B, 1, m);
			d[n++] = (struct starpu_data_descr){starpu_data_get_sub_data(c->A, 2, k, m), STARPU_R};
			d[n++] = (struct starpu_data_descr){bm, STARPU_RW};
			above += (double)starpu_matrix_get_nx(bm);
		}
		d[n++] = (struct starpu_data_descr){c->wsSolve, STARPU_SCRATCH | STARPU_NOFOOTPRINT};
		ret = starpu_task_insert(&cl_solve_bwd, STARPU_PRIORITY, pr->scheme ? 2 * k + 2 : 0,
		                         STARPU_DATA_MODE_ARRAY, d, n, STARPU_CL_ARGS_NFREE, &c->pairs[k], sizeof(int),
		                         STARPU_FLOPS, flops_trsm_left(nk, r) + flops_gemm(above, r, nk),
		                         STARPU_NAME, "SOLVE_BWD", STARPU_TASK_WORKERIDS, c->route.len, c->route.others, 0);
	}
	free(d);
	return ret;
}

static void collect_stats(struct chol_stats *st, double elapsed_us, double flops)
{
	memset(st, 0, sizeof(*st));
	st->elapsed_us = elapsed_us;
	st->flops = flops;
	const size_t nrec = starpu_trace_count();
	st->ntasks = (long)nrec;
	for (size_t i = 0; i < nrec; i++) {
		struct starpu_trace_record r;
		if (starpu_trace_get(i, &r))
			continue;
		if (starpu_worker_get_type(r.workerid) != STARPU_HIP_WORKER)
			continue;
		for (int q = 0; q < CHOL_NKINDS; q++)
			if (!strcmp(r.name, kind_names[q])) {
				st->kernel_us[q] += r.kernel_us;
				st->kernel_n[q]++;
			}
	}
}

/* Unpartitions and unregisters everything; returns POTRF's global info. */
static int ctx_release(struct chol_ctx *c)
{
	int info = 0;
	if (c->wsTrtri)
		starpu_data_unregister(c->wsTrtri);
	for (int k = 0; k < c->nblocks; k++) {
		if (c->W && c->W[k])
			starpu_data_unregister(c->W[k]);
		if (c->Linv && c->Linv[k])
			starpu_data_unregister(c->Linv[k]);
	}
	free(c->W);
	free(c->Linv);
	free(c->pairs);
	if (c->info) {
		starpu_data_unpartition(c->info, c->dinfos ? (unsigned)c->home : STARPU_MAIN_RAM);
		starpu_data_unregister(c->info);
		if (c->dinfos) {
			if (sa_memcpy_d2h(c->infos, c->dinfos, (size_t)c->nblocks * sizeof(int)))
				info = -EIO;
			starpu_free_on_node((unsigned)c->home, (uintptr_t)c->dinfos, (size_t)c->nblocks * sizeof(int));
			c->dinfos = NULL;
		}
		starpu_data_unregister(c->wsPotrf);
		starpu_data_unregister(c->wsTrsm);
		for (int k = 0; k < c->nblocks && !info; k++)
			if (c->infos[k]) {
				size_t sz, off;
				starpu_filter_nparts_compute_chunk_size_and_offset((unsigned)c->N, (unsigned)c->nblocks, 1,
				                                                   (unsigned)k, 1, &sz, &off);
				info = (int)off + c->infos[k];
			}
		free(c->infos);
	} else {
		if (c->dinfos)
			starpu_free_on_node((unsigned)c->home, (uintptr_t)c->dinfos, (size_t)c->nblocks * sizeof(int));
		free(c->infos);
	}
	if (c->B) {
		starpu_data_unpartition(c->B, (unsigned)c->home);
		starpu_data_unregister(c->B);
		if (c->wsSolve)
			starpu_data_unregister(c->wsSolve);
	}
	starpu_data_unpartition(c->A, (unsigned)c->home);
	starpu_data_unregister(c->A);
	return info;
}

/* Tile kernels that run in parallel on one device: its HIP workers
 * (streams), but no more than the hardware queues HIP spreads streams over
 * (GPU_MAX_HW_QUEUES, 4 unless set), since streams sharing a queue run one
 * after the other. */
static int hip_streams_per_device(void)
{
	int hip = 0, maxdev = -1;
	for (unsigned w = 0; w < starpu_worker_get_count(); w++)
		if (starpu_worker_get_type((int)w) == STARPU_HIP_WORKER) {
			hip++;
			if (starpu_worker_get_devid((int)w) > maxdev)
				maxdev = starpu_worker_get_devid((int)w);
		}
	int streams = hip ? hip / (maxdev + 1) : 1;
	const char *q = getenv("GPU_MAX_HW_QUEUES");
	const int queues = (q && atoi(q) > 0) ? atoi(q) : 4;
	return streams < queues ? streams : queues;
}

/* The factorisation's chain on the GPU's set-aside worker, when A lives on
 * that GPU and it has one (SA_CHOL_CRITICAL_WORKER=0 turns the routing
 * off); the other tasks then avoid that worker. */
static int critical_worker_enabled(void)
{
	const char *e = getenv("SA_CHOL_CRITICAL_WORKER");
	return !(e && atoi(e) == 0);
}

static int run(int home_node, double *A, int lda, int N, int nblocks, double *B, int ldb, int nrhs,
               int factor, int prio, struct chol_stats *st)
{
	if (N <= 0 || nblocks <= 0 || nblocks > N || lda < N)
		return -EINVAL;
	if (B && (nrhs <= 0 || ldb < N))
		return -EINVAL;
	if (!starpu_is_initialized())
		return -ENODEV;
	struct chol_ctx c;
	int ret = ctx_register(&c, home_node, A, lda, N, nblocks, B, ldb, nrhs, factor);
	if (ret) {
		ctx_release(&c);
		return ret;
	}
	crit_route_init(&c.route, home_node, critical_worker_enabled());
	struct chol_prio pr;
	prio_init(&pr, &c, prio);
	const int conc_prev = sa_set_launch_concurrency(hip_streams_per_device());
	const int prof_prev = starpu_profiling_status_set(1);
	starpu_trace_reset();
	const double start = starpu_timing_now();
	if (factor)
		ret = submit_factor(&c, &pr);
	if (!ret && B)
		ret = submit_solve(&c, &pr);
	starpu_task_wait_for_all();
	const double end = starpu_timing_now();
	if (st)
		collect_stats(st, end - start,
		              (factor ? chol_flops_potrf((double)N) : 0.0) +
		                  (B ? chol_flops_potrs((double)N, (double)nrhs) : 0.0));
	starpu_profiling_status_set(prof_prev);
	sa_set_launch_concurrency(conc_prev);
	const int info = ctx_release(&c);
	if (st)
		st->info = info;
	return ret;
}

int chol_factor(int home_node, double *A, int lda, int N, int nblocks, int prio, struct chol_stats *st)
{
	return run(home_node, A, lda, N, nblocks, NULL, 0, 0, 1, prio, st);
}

int chol_solve(int home_node, const double *L, int ldl, int N, int nblocks, double *B, int ldb, int nrhs,
               int prio, struct chol_stats *st)
{
	if (!B)
		return -EINVAL;
	return run(home_node, (double *)L, ldl, N, nblocks, B, ldb, nrhs, 0, prio, st);
}

int chol_factor_solve(int home_node, double *A, int lda, int N, int nblocks, double *B, int ldb, int nrhs,
                      int prio, struct chol_stats *st)
{
	if (!B)
		return -EINVAL;
	return run(home_node, A, lda, N, nblocks, B, ldb, nrhs, 1, prio, st);
}
