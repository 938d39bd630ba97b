/*
 * Mixed-precision Cholesky solve with iterative refinement on the
 * starpu_amd runtime (no counterpart in the reference, whose Cholesky is a
 * single-precision factorisation without a solve).
 *
 * A (N x N, FP64, on the GPU; only its lower triangle is read) is converted into an FP32
 * working copy that is factored as A = U^T U on its upper triangle by the
 * tile DAG of cholesky_implicit.c turned on its side: for each k, POTRF on
 * U_kk, TRSM on the row panel U_kn (n > k), SYRK on A_nn and GEMM on A_mn
 * (k < m < n), with the same base priorities.  Working on the upper
 * triangle makes every operand of the updates k-contiguous (see
 * sa_mixed.hip).  Precision by task:
 *   POTRF  FP64: the FP32 tile is widened, factored and inverted,
 *          W_k = L_kk^{-T} (= U_kk^{-1}), in one sa_dpotrf_tile launch in
 *          FP64, both rounded to FP32;
 *   TRSM   U_kn = W_k^T A_kn, a GEMM in the update precision;
 *   SYRK / GEMM  in the update precision: FP32 MFMA, BF16 MFMA, or BF16x3
 *          (hi/lo split), FP32 accumulation.
 * Then refinement, in FP64 on the GPU: r = b - A x (FP64 residual over the
 * FP64 A's lower triangle, sa_dsymm_lower_resid), d = (U^T U)^{-1} r with the FP32 factor (panel products in FP64
 * arithmetic), x += d, until the normwise backward error
 * max_j ||r_j|| / (||A||_F ||x_j|| + ||b_j||) reaches the tolerance.
 */
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <starpu.h>
#include "starpu_amd_chol.h"
#include "starpu_amd_kernels.h"
#include "chol_prio.h"

static void hip_check(int e)
{
	if (e)
		STARPU_HIP_REPORT_ERROR(e);
}

#define FPTR(i) ((float *)STARPU_MATRIX_GET_PTR(descr[i]))
#define MNX(i) ((int)STARPU_MATRIX_GET_NX(descr[i]))
#define MNY(i) ((int)STARPU_MATRIX_GET_NY(descr[i]))
#define MLD(i) ((int)STARPU_MATRIX_GET_LD(descr[i]))
#define VPTR(i) ((char *)STARPU_VARIABLE_GET_PTR(descr[i]))
#define VSZ(i) ((size_t)STARPU_VARIABLE_GET_ELEMSIZE(descr[i]))

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

/* scratch of the POTRF task: L and W = L^{-T} in FP64, then the larger of
 * the POTRF and TRTRI workspaces */
static size_t potrf_ws(int nb)
{
	size_t w = sa_dpotrf_l_worksize(nb);
	if (sa_dtrtri_lt_worksize(nb) > w)
		w = sa_dtrtri_lt_worksize(nb);
	if (sa_dpotrf_tile_worksize(nb) > w)
		w = sa_dpotrf_tile_worksize(nb);
	return 2 * align256((size_t)nb * nb * sizeof(double)) + w;
}

/* Bulk workgroups of the mixed factorisation's tile POTRF: more than the
 * FP64 graph's 31, since this chain, not the update throughput, bounds the
 * mixed factorisation (N = 32768 BF16x3: 95.7-96.3 ms with 63 against
 * 98.2-100.5 with 31; 47 and 95 between).  SA_MIXED_POTRF_WORKGROUPS
 * overrides (1..255). */
static int mixed_potrf_workgroups(void)
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_MIXED_POTRF_WORKGROUPS");
		const int x = e ? atoi(e) : 0;
		v = (x >= 1 && x <= 255) ? x : 63;
	}
	return v;
}

/* U_kk (FP32 upper, RW), W_k (FP32, W), info_k (W), scratch */
static void hip_potrf_mixed(void *descr[], void *arg)
{
	(void)arg;
	hipStream_t s = starpu_hip_get_local_stream();
	const int n = MNY(0);
	float *U = FPTR(0);
	float *W32 = FPTR(1);
	int *info = (int *)STARPU_VECTOR_GET_PTR(descr[2]);
	char *ws = VPTR(3);
	const size_t t = align256((size_t)n * n * sizeof(double));
	double *L = (double *)ws, *W = (double *)(ws + t);
	void *work = ws + 2 * t;
	const size_t wb = VSZ(3) - 2 * t;
	hip_check(sa_slag2d_upper_to_lower(s, n, U, MLD(0), L, n));
	if (n <= sa_dpotrf_tile_maxn()) {
		/* factor and W = L^-T in one launch */
		hip_check(sa_dpotrf_tile_wg(s, n, L, n, NULL, 0, W, n, info, work, wb, mixed_potrf_workgroups()));
	} else {
		hip_check(sa_dpotrf_l(s, n, L, n, info, work, wb));
		hip_check(sa_dtrtri_lt(s, n, L, n, W, n, work, wb));
	}
	hip_check(sa_tile_outputs_f32(s, n, L, n, U, MLD(0), W, n, W32, MLD(1)));
}

/* W_k (R), A_kn (RW), scratch: A_kn <- W_k^T A_kn */
static void trsm_mixed(void *descr[], int mode)
{
	hipStream_t s = starpu_hip_get_local_stream();
	const int m = MNX(1), n = MNY(1);
	float *T = (float *)VPTR(2);
	hip_check(sa_slacpy(s, m, n, FPTR(1), MLD(1), T, m));
	hip_check(sa_sgemm_tn_utri(s, mode, m, n, 1.0f, FPTR(0), MLD(0), T, m, 0.0f, FPTR(1), MLD(1)));
}

/* U_kn (R), A_nn (RW): upper(A_nn) -= U_kn^T U_kn */
static void syrk_mixed(void *descr[], int mode)
{
	hip_check(sa_sgemm_tn(starpu_hip_get_local_stream(), mode, 1, MNY(1), MNY(1), MNX(0), -1.0f, FPTR(0), MLD(0),
	                      FPTR(0), MLD(0), 1.0f, FPTR(1), MLD(1)));
}

/* U_km (R), U_kn (R), A_mn (RW): A_mn -= U_km^T U_kn */
static void gemm_mixed(void *descr[], int mode)
{
	hip_check(sa_sgemm_tn(starpu_hip_get_local_stream(), mode, 0, MNX(2), MNY(2), MNX(0), -1.0f, FPTR(0), MLD(0),
	                      FPTR(1), MLD(1), 1.0f, FPTR(2), MLD(2)));
}

/* BF16x3 with the panel split ahead of time (opt-in, SA_MIXED_PRESPLIT=1):
 * the TRSM that finishes U_kn also writes its bf16 planes H_kn = bf16(U_kn),
 * L_kn = bf16(U_kn - H_kn), and the SYRK and GEMM updates read the planes
 * instead of splitting U_kn in every product (bit-identical results). */

/* W_k (R), A_kn (RW), scratch, H_kn (W), L_kn (W) */
static void hip_trsm_split(void *descr[], void *arg)
{
	(void)arg;
	trsm_mixed(descr, CHOL_PREC_BF16X3);
	hip_check(sa_split_bf16x2(starpu_hip_get_local_stream(), MNX(1), MNY(1), FPTR(1), MLD(1),
	                          (void *)STARPU_MATRIX_GET_PTR(descr[3]), (void *)STARPU_MATRIX_GET_PTR(descr[4]),
	                          MLD(3)));
}

#define HPTR(i) ((const void *)STARPU_MATRIX_GET_PTR(descr[i]))

/* H_kn, L_kn (R), A_nn (RW): upper(A_nn) -= U_kn^T U_kn */
static void hip_syrk_split(void *descr[], void *arg)
{
	(void)arg;
	hip_check(sa_sgemm_tn_split(starpu_hip_get_local_stream(), 1, MNY(0), MNY(0), MNX(0), -1.0f, HPTR(0), HPTR(1),
	                            MLD(0), HPTR(0), HPTR(1), MLD(0), 1.0f, FPTR(2), MLD(2)));
}

/* H_km, L_km, H_kn, L_kn (R), A_mn (RW): A_mn -= U_km^T U_kn */
static void hip_gemm_split(void *descr[], void *arg)
{
	(void)arg;
	hip_check(sa_sgemm_tn_split(starpu_hip_get_local_stream(), 0, MNX(4), MNY(4), MNX(0), -1.0f, HPTR(0), HPTR(1),
	                            MLD(0), HPTR(2), HPTR(3), MLD(2), 1.0f, FPTR(4), MLD(4)));
}

#define MIXED_FUNCS(MODE, SUF)                                                                  \
	static void hip_trsm_##SUF(void *d[], void *a) { (void)a; trsm_mixed(d, MODE); }            \
	static void hip_syrk_##SUF(void *d[], void *a) { (void)a; syrk_mixed(d, MODE); }            \
	static void hip_gemm_##SUF(void *d[], void *a) { (void)a; gemm_mixed(d, MODE); }
MIXED_FUNCS(CHOL_PREC_FP32, f32)
MIXED_FUNCS(CHOL_PREC_BF16, bf16)
MIXED_FUNCS(CHOL_PREC_BF16X3, bf16x3)

static const char *mode_name[3] = {"f32", "bf16", "bf16x3"};
static starpu_hip_func_t trsm_f[3] = {hip_trsm_f32, hip_trsm_bf16, hip_trsm_bf16x3};
static starpu_hip_func_t syrk_f[3] = {hip_syrk_f32, hip_syrk_bf16, hip_syrk_bf16x3};
static starpu_hip_func_t gemm_f[3] = {hip_gemm_f32, hip_gemm_bf16, hip_gemm_bf16x3};

/* one codelet set and one history model set per precision: the footprint
 * of a tile does not tell the precisions apart */
static struct starpu_perfmodel m_potrf = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_potrf_mixed"};
static struct starpu_perfmodel m_trsm[3], m_syrk[3], m_gemm[3];
static struct starpu_codelet cl_potrf_m, cl_trsm_m[3], cl_syrk_m[3], cl_gemm_m[3];
static struct starpu_perfmodel m_trsm_s = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_trsm_bf16x3_split"};
static struct starpu_perfmodel m_syrk_s = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_syrk_bf16x3_split"};
static struct starpu_perfmodel m_gemm_s = {.type = STARPU_HISTORY_BASED, .symbol = "chol_model_gemm_bf16x3_split"};
static struct starpu_codelet cl_trsm_s, cl_syrk_s, cl_gemm_s;
static char model_names[9][48];
static pthread_once_t codelets_once = PTHREAD_ONCE_INIT;

static void init_codelets_once(void)
{
	memset(&cl_potrf_m, 0, sizeof(cl_potrf_m));
	cl_potrf_m.type = STARPU_SEQ;
	cl_potrf_m.hip_funcs[0] = hip_potrf_mixed;
	cl_potrf_m.hip_flags[0] = STARPU_HIP_ASYNC;
	cl_potrf_m.nbuffers = 4;
	cl_potrf_m.modes[0] = STARPU_RW;
	cl_potrf_m.modes[1] = STARPU_W;
	cl_potrf_m.modes[2] = STARPU_W | STARPU_NOFOOTPRINT;
	cl_potrf_m.modes[3] = STARPU_SCRATCH | STARPU_NOFOOTPRINT;
	cl_potrf_m.model = &m_potrf;
	cl_potrf_m.name = "cl_POTRF_mixed";
	for (int p = 0; p < 3; p++) {
		struct starpu_perfmodel *ms[3] = {&m_trsm[p], &m_syrk[p], &m_gemm[p]};
		const char *kind[3] = {"trsm", "syrk", "gemm"};
		for (int q = 0; q < 3; q++) {
			memset(ms[q], 0, sizeof(*ms[q]));
			ms[q]->type = STARPU_HISTORY_BASED;
			snprintf(model_names[3 * p + q], sizeof(model_names[0]), "chol_model_%s_%s", kind[q], mode_name[p]);
			ms[q]->symbol = model_names[3 * p + q];
		}
		struct starpu_codelet *c = &cl_trsm_m[p];
		memset(c, 0, sizeof(*c));
		c->type = STARPU_SEQ;
		c->hip_funcs[0] = trsm_f[p];
		c->hip_flags[0] = STARPU_HIP_ASYNC;
		c->nbuffers = 3;
		c->modes[0] = STARPU_R;
		c->modes[1] = STARPU_RW;
		c->modes[2] = STARPU_SCRATCH | STARPU_NOFOOTPRINT;
		c->model = &m_trsm[p];
		c->name = "cl_TRSM_mixed";
		c = &cl_syrk_m[p];
		memset(c, 0, sizeof(*c));
		c->type = STARPU_SEQ;
		c->hip_funcs[0] = syrk_f[p];
		c->hip_flags[0] = STARPU_HIP_ASYNC;
		c->nbuffers = 2;
		c->modes[0] = STARPU_R;
		c->modes[1] = STARPU_RW;
		c->model = &m_syrk[p];
		c->name = "cl_SYRK_mixed";
		c = &cl_gemm_m[p];
		memset(c, 0, sizeof(*c));
		c->type = STARPU_SEQ;
		c->hip_funcs[0] = gemm_f[p];
		c->hip_flags[0] = STARPU_HIP_ASYNC;
		c->nbuffers = 3;
		c->modes[0] = STARPU_R;
		c->modes[1] = STARPU_R;
		c->modes[2] = STARPU_RW;
		c->model = &m_gemm[p];
		c->name = "cl_GEMM_mixed";
	}
	struct starpu_codelet *c = &cl_trsm_s;
	memset(c, 0, sizeof(*c));
	c->type = STARPU_SEQ;
	c->hip_funcs[0] = hip_trsm_split;
	c->hip_flags[0] = STARPU_HIP_ASYNC;
	c->nbuffers = 5;
	c->modes[0] = STARPU_R;
	c->modes[1] = STARPU_RW;
	c->modes[2] = STARPU_SCRATCH | STARPU_NOFOOTPRINT;
	c->modes[3] = STARPU_W;
	c->modes[4] = STARPU_W;
	c->model = &m_trsm_s;
	c->name = "cl_TRSM_mixed_split";
	c = &cl_syrk_s;
	memset(c, 0, sizeof(*c));
	c->type = STARPU_SEQ;
	c->hip_funcs[0] = hip_syrk_split;
	c->hip_flags[0] = STARPU_HIP_ASYNC;
	c->nbuffers = 3;
	c->modes[0] = STARPU_R;
	c->modes[1] = STARPU_R;
	c->modes[2] = STARPU_RW;
	c->model = &m_syrk_s;
	c->name = "cl_SYRK_mixed_split";
	c = &cl_gemm_s;
	memset(c, 0, sizeof(*c));
	c->type = STARPU_SEQ;
	c->hip_funcs[0] = hip_gemm_split;
	c->hip_flags[0] = STARPU_HIP_ASYNC;
	c->nbuffers = 5;
	for (int b = 0; b < 4; b++)
		c->modes[b] = STARPU_R;
	c->modes[4] = STARPU_RW;
	c->model = &m_gemm_s;
	c->name = "cl_GEMM_mixed_split";
}

/* thread-safe: several factorisations may start at once (the solve server) */
static void init_codelets(void) { pthread_once(&codelets_once, init_codelets_once); }

/* The priority attributions of the FP64 graph (chol_prio.h) with (row,
 * column) swapped for the upper working copy; the attribution from task
 * times (3) falls back to the base one here. */

struct tile_offsets {
	int *off, *sz;
};

static int submit_mixed_factor(starpu_data_handle_t A32, starpu_data_handle_t *W, starpu_data_handle_t info,
                               starpu_data_handle_t wsP, starpu_data_handle_t wsT, starpu_data_handle_t H,
                               starpu_data_handle_t L, int nb, int mode, int prio, const struct crit_route *rt)
{
	const int split = H && L;
	const int crit = rt->crit;
	uint32_t *others = (uint32_t *)rt->others;
	const unsigned olen = rt->len;
	struct chol_prio pr = {0, 0.0, 0.0, 0.0, 0.0};
	pr.scheme = (prio < 0 || prio > 4 || prio == 3) ? 1 : prio;
	int ret = 0;
	for (int k = 0; k < nb && !ret; k++) {
		starpu_data_handle_t ukk = starpu_data_get_sub_data(A32, 2, k, k);
		const double nk = (double)starpu_matrix_get_nx(ukk);
		ret = starpu_task_insert(&cl_potrf_m, STARPU_PRIORITY, prio_potrf(&pr, nb, k), STARPU_RW, ukk,
		                         STARPU_W, W[k], STARPU_W, starpu_data_get_child(info, (unsigned)k), STARPU_SCRATCH,
		                         wsP, STARPU_FLOPS, chol_flops_potrf(nk) + nk * nk * nk, STARPU_NAME, "POTRF",
		                         STARPU_EXECUTE_ON_WORKER, crit, 0);
		for (int n = k + 1; n < nb && !ret; n++) {
			starpu_data_handle_t akn = starpu_data_get_sub_data(A32, 2, k, n);
			const double fl = 2.0 * nk * nk * (double)starpu_matrix_get_ny(akn);
			const int on_chain = n == k + 1;
			if (split)
				ret = starpu_task_insert(&cl_trsm_s, STARPU_PRIORITY, prio_trsm(&pr, nb, k, n), STARPU_R, W[k],
				                         STARPU_RW, akn, STARPU_SCRATCH, wsT, STARPU_W,
				                         starpu_data_get_sub_data(H, 2, k, n), STARPU_W,
				                         starpu_data_get_sub_data(L, 2, k, n), STARPU_FLOPS, fl, STARPU_NAME, "TRSM",
				                         STARPU_EXECUTE_ON_WORKER, on_chain ? crit : -1, STARPU_TASK_WORKERIDS,
				                         on_chain ? 0u : olen, others, 0);
			else
				ret = starpu_task_insert(&cl_trsm_m[mode], STARPU_PRIORITY, prio_trsm(&pr, nb, k, n), STARPU_R,
				                         W[k], STARPU_RW, akn, STARPU_SCRATCH, wsT, STARPU_FLOPS, fl, STARPU_NAME,
				                         "TRSM", STARPU_EXECUTE_ON_WORKER, on_chain ? crit : -1,
				                         STARPU_TASK_WORKERIDS, on_chain ? 0u : olen, others, 0);
		}
		starpu_data_wont_use(ukk);
		for (int n = k + 1; n < nb && !ret; n++) {
			starpu_data_handle_t ukn = starpu_data_get_sub_data(A32, 2, k, n);
			starpu_data_handle_t ann = starpu_data_get_sub_data(A32, 2, n, n);
			const double nn = (double)starpu_matrix_get_ny(ann);
			starpu_data_handle_t hkn = split ? starpu_data_get_sub_data(H, 2, k, n) : NULL;
			starpu_data_handle_t lkn = split ? starpu_data_get_sub_data(L, 2, k, n) : NULL;
			const int on_chain = n == k + 1;
			if (split)
				ret = starpu_task_insert(&cl_syrk_s, STARPU_PRIORITY, prio_gemm(&pr, nb, k, n, n), STARPU_R, hkn,
				                         STARPU_R, lkn, STARPU_RW, ann, STARPU_FLOPS, nk * nn * (nn + 1.0),
				                         STARPU_NAME, "SYRK", STARPU_EXECUTE_ON_WORKER, on_chain ? crit : -1,
				                         STARPU_TASK_WORKERIDS, on_chain ? 0u : olen, others, 0);
			else
				ret = starpu_task_insert(&cl_syrk_m[mode], STARPU_PRIORITY, prio_gemm(&pr, nb, k, n, n), STARPU_R,
				                         ukn, STARPU_RW, ann, STARPU_FLOPS, nk * nn * (nn + 1.0), STARPU_NAME, "SYRK",
				                         STARPU_EXECUTE_ON_WORKER, on_chain ? crit : -1, STARPU_TASK_WORKERIDS,
				                         on_chain ? 0u : olen, others, 0);
			for (int m = k + 1; m < n && !ret; m++) {
				starpu_data_handle_t ukm = starpu_data_get_sub_data(A32, 2, k, m);
				starpu_data_handle_t amn = starpu_data_get_sub_data(A32, 2, m, n);
				const double fl = 2.0 * nk * nn * (double)starpu_matrix_get_nx(amn);
				if (split)
					ret = starpu_task_insert(&cl_gemm_s, STARPU_PRIORITY, prio_gemm(&pr, nb, k, n, m), STARPU_R,
					                         starpu_data_get_sub_data(H, 2, k, m), STARPU_R,
					                         starpu_data_get_sub_data(L, 2, k, m), STARPU_R, hkn, STARPU_R, lkn,
					                         STARPU_RW, amn, STARPU_FLOPS, fl, STARPU_NAME, "GEMM",
					                         STARPU_TASK_WORKERIDS, olen, others, 0);
				else
					ret = starpu_task_insert(&cl_gemm_m[mode], STARPU_PRIORITY, prio_gemm(&pr, nb, k, n, m),
					                         STARPU_R, ukm, STARPU_R, ukn, STARPU_RW, amn, STARPU_FLOPS, fl,
					                         STARPU_NAME, "GEMM", STARPU_TASK_WORKERIDS, olen, others, 0);
			}
			starpu_data_wont_use(ukn);
			if (split) {
				starpu_data_wont_use(hkn);
				starpu_data_wont_use(lkn);
			}
		}
	}
	return ret;
}

/* R <- (U^T U)^{-1} R with the FP32 factor U (upper, ld ldu) and the FP32
 * inverses W (block k at W + off_k * ldw, nb_k x nb_k, ld ldw), right-looking
 * block by block (FP64 arithmetic, FP64 right-hand sides, Y an N x nrhs
 * workspace):
 *   forward  U^T y = r:  y_k = W_k^T r_k, then r_j -= U_kj^T y_k for all
 *                        j > k in one product with the row panel U_k,(k+1:);
 *   backward U x = y:    x_k = W_k y_k (into R), then y_j -= U_jk x_k for all
 *                        j < k in one product with the column panel U_(0:k),k.
 * Every update spans the whole remaining panel, so it needs no split over k
 * (nor its reduction) once the panel is tall, and the triangular products
 * write their own output block, so no copy of the right-hand sides is made. */
static int solve_fp32_factor(hipStream_t s, int N, int nblocks, const int *off, const float *U, int ldu,
                             const float *W, int ldw, double *R, int ldr, double *Y, int nrhs, void *work, size_t wb)
{
	int e = 0;
	for (int k = 0; k < nblocks && !e; k++) {
		const int o = off[k], w = off[k + 1] - off[k], o2 = off[k + 1];
		e = sa_gemm_tn_sd(s, w, nrhs, w, 1.0, W + (long)o * ldw, ldw, R + o, ldr, 0.0, Y + o, N, work, wb);
		if (!e && o2 < N)
			e = sa_gemm_tn_sd(s, N - o2, nrhs, w, -1.0, U + o + (long)o2 * ldu, ldu, Y + o, N, 1.0, R + o2, ldr,
			                  work, wb);
	}
	for (int k = nblocks - 1; k >= 0 && !e; k--) {
		const int o = off[k], w = off[k + 1] - off[k];
		e = sa_gemm_nn_sd(s, w, nrhs, w, 1.0, W + (long)o * ldw, ldw, Y + o, N, 0.0, R + o, ldr, work, wb);
		if (!e && o > 0)
			e = sa_gemm_nn_sd(s, o, nrhs, w, -1.0, U + (long)o * ldu, ldu, R + o, ldr, 1.0, Y, N, work, wb);
	}
	return e;
}

static double max_berr(int N, int nrhs, double anorm, const double *rr, const double *xx, const double *bb)
{
	double worst = 0.0;
	for (int j = 0; j < nrhs; j++) {
		const double be = sqrt(rr[j]) / (anorm * sqrt(xx[j]) + sqrt(bb[j]));
		if (!(be <= worst))      /* NaN propagates as the worst */
			worst = be;
	}
	(void)N;
	return worst;
}

/* ---- reusable steps: the factor object, its factorisation, refinement ---- */

void chol_mixed_factor_release(struct chol_mixed_factor *f)
{
	if (f->U)
		sa_dev_free(f->U);
	if (f->W)
		sa_dev_free(f->W);
	free(f->off);
	memset(f, 0, sizeof(*f));
}

size_t chol_mixed_factor_bytes(int N, int nblocks)
{
	const size_t nbmax = (size_t)((N + nblocks - 1) / nblocks);
	return (size_t)N * N * sizeof(float) + nbmax * N * sizeof(float);
}

static int mixed_factor(int hip_node, const double *A, int lda, int N, int nblocks, int prec, int prio,
                        hipStream_t s, struct chol_mixed_factor *f, double *convert_us, double *factor_us,
                        int critical_worker);

/* Factorisations run side by side (the solve server) share the streams:
 * the reserved worker is not given to one factorisation's chain. */
int chol_mixed_factor(int hip_node, const double *A, int lda, int N, int nblocks, int prec, int prio,
                      hipStream_t s, struct chol_mixed_factor *f, double *convert_us, double *factor_us)
{
	return mixed_factor(hip_node, A, lda, N, nblocks, prec, prio, s, f, convert_us, factor_us, 0);
}

static int mixed_factor(int hip_node, const double *A, int lda, int N, int nblocks, int prec, int prio,
                        hipStream_t s, struct chol_mixed_factor *f, double *convert_us, double *factor_us,
                        int critical_worker)
{
	if (N <= 0 || nblocks <= 0 || nblocks > N || lda < N || prec < 0 || prec > 2 || !A || !f)
		return -EINVAL;
	if (!starpu_is_initialized())
		return -ENODEV;
	init_codelets();
	memset(f, 0, sizeof(*f));
	f->N = N;
	f->nblocks = nblocks;
	f->node = hip_node;
	f->off = calloc((size_t)nblocks + 1, sizeof(int));
	int *infos = calloc((size_t)nblocks, sizeof(int));
	starpu_data_handle_t *Wh = calloc((size_t)nblocks, sizeof(*Wh));
	int ret = 0;
	if (!f->off || !infos || !Wh) {
		ret = -ENOMEM;
		goto out;
	}
	for (int k = 0; k < nblocks; k++) {
		size_t sz, o;
		starpu_filter_nparts_compute_chunk_size_and_offset((unsigned)N, (unsigned)nblocks, 1, (unsigned)k, 1, &sz, &o);
		f->off[k] = (int)o;
	}
	f->off[nblocks] = N;
	f->nbmax = f->off[1] - f->off[0];
	f->ldu = N;
	f->ldw = f->nbmax;
	f->U = sa_dev_malloc((size_t)f->ldu * N * sizeof(float));
	f->W = sa_dev_malloc((size_t)f->ldw * N * sizeof(float));
	if (!f->U || !f->W) {
		ret = -ENOMEM;
		goto out;
	}
	const double t0 = starpu_timing_now();
	/* the FP32 working copy (upper triangle from A's lower triangle) and
	 * ||A||_F, kept with the factor for refinement, in one read of the
	 * lower triangle */
	{
		const size_t cb = sa_dlag2s_lower_to_upper_fro_worksize(N);
		char *cw = (char *)starpu_malloc_on_node((unsigned)hip_node, cb + 256);
		double a2 = 0.0;
		int e = cw ? 0 : -ENOMEM;
		e = e ? e : sa_dlag2s_lower_to_upper_fro(s, N, A, lda, f->U, f->ldu, (double *)cw, cw + 256, cb);
		e = e ? e : sa_memcpy2d_async(&a2, sizeof(double), cw, sizeof(double), sizeof(double), 1, 2, s);
		e = e ? e : sa_stream_sync(s);
		if (cw)
			starpu_free_on_node((unsigned)hip_node, (uintptr_t)cw, cb + 256);
		if (e) {
			ret = e < 0 ? e : -EIO;
			goto out;
		}
		f->anorm = sqrt(a2);
	}
	const double t1 = starpu_timing_now();
	if (convert_us)
		*convert_us = t1 - t0;

	/* the tile factorisations' pivot reports stay on the GPU and come back
	 * in one copy at the end (a gather to host memory would copy each
	 * block's report on its own) */
	int *dinfos = (int *)starpu_malloc_on_node((unsigned)hip_node, (size_t)nblocks * sizeof(int));
	if (!dinfos || sa_memcpy_h2d(dinfos, infos, (size_t)nblocks * sizeof(int))) {
		if (dinfos)
			starpu_free_on_node((unsigned)hip_node, (uintptr_t)dinfos, (size_t)nblocks * sizeof(int));
		ret = -ENOMEM;
		goto out;
	}
	/* factorisation DAG; unpartitioning waits for exactly this DAG's tasks,
	 * so several factorisations can run at once on one runtime */
	starpu_data_handle_t hA, hW, hInfo, wsP, wsT;
	starpu_matrix_data_register(&hA, hip_node, (uintptr_t)f->U, (size_t)f->ldu, (size_t)N, (size_t)N, sizeof(float));
	struct starpu_data_filter fb = {.filter_func = starpu_matrix_filter_block, .nchildren = (unsigned)nblocks};
	struct starpu_data_filter fv = {.filter_func = starpu_matrix_filter_vertical_block,
	                                .nchildren = (unsigned)nblocks};
	starpu_data_map_filters(hA, 2, &fb, &fv);
	starpu_matrix_data_register(&hW, hip_node, (uintptr_t)f->W, (size_t)f->ldw, (size_t)f->nbmax, (size_t)N,
	                            sizeof(float));
	starpu_data_partition(hW, &fv);
	for (int k = 0; k < nblocks; k++)
		Wh[k] = starpu_data_get_sub_data(hW, 1, k);
	starpu_vector_data_register(&hInfo, hip_node, (uintptr_t)dinfos, (size_t)nblocks, sizeof(int));
	struct starpu_data_filter fi = {.filter_func = starpu_vector_filter_block, .nchildren = (unsigned)nblocks};
	starpu_data_partition(hInfo, &fi);
	starpu_variable_data_register(&wsP, -1, 0, potrf_ws(f->nbmax));
	starpu_variable_data_register(&wsT, -1, 0, (size_t)f->nbmax * f->nbmax * sizeof(float));
	/* SA_MIXED_PRESPLIT=1, BF16x3 on tiles of multiples of 128: the panels'
	 * bf16 planes (N x N each, freed with the DAG), written by the TRSMs and
	 * read by the updates.  Off by default: the update GEMM alone runs 13-17 %
	 * faster on split operands, but the DAG with two more data handles per
	 * update ran slower (93.9 against 84.1 ms at N=32768, tile 1024) */
	int split = prec == CHOL_PREC_BF16X3;
	for (int k = 0; k < nblocks && split; k++)
		if ((f->off[k + 1] - f->off[k]) % 128)
			split = 0;
	const char *ps = getenv("SA_MIXED_PRESPLIT");
	if (!ps || atoi(ps) != 1)
		split = 0;
	void *Hp = split ? sa_dev_malloc((size_t)N * N * 2) : NULL;
	void *Lp = split ? sa_dev_malloc((size_t)N * N * 2) : NULL;
	starpu_data_handle_t hH = NULL, hL = NULL;
	if (Hp && Lp) {
		starpu_matrix_data_register(&hH, hip_node, (uintptr_t)Hp, (size_t)N, (size_t)N, (size_t)N, 2);
		starpu_matrix_data_register(&hL, hip_node, (uintptr_t)Lp, (size_t)N, (size_t)N, (size_t)N, 2);
		starpu_data_map_filters(hH, 2, &fb, &fv);
		starpu_data_map_filters(hL, 2, &fb, &fv);
	}
	struct crit_route route;
	crit_route_init(&route, hip_node, critical_worker);
	ret = submit_mixed_factor(hA, Wh, hInfo, wsP, wsT, hH, hL, nblocks, prec, prio, &route);
	starpu_data_unpartition(hA, (unsigned)hip_node);
	starpu_data_unpartition(hW, (unsigned)hip_node);
	if (hH) {
		starpu_data_unpartition(hH, (unsigned)hip_node);
		starpu_data_unpartition(hL, (unsigned)hip_node);
		starpu_data_unregister(hH);
		starpu_data_unregister(hL);
	}
	if (Hp)
		sa_dev_free(Hp);
	if (Lp)
		sa_dev_free(Lp);
	starpu_data_unpartition(hInfo, (unsigned)hip_node);
	starpu_data_unregister(hInfo);
	/* one copy on the caller's stream, then the buffer back to the node's
	 * cache (a device free waits for the whole GPU) */
	const size_t ib = (size_t)nblocks * sizeof(int);
	if ((sa_memcpy2d_async(infos, ib, dinfos, ib, ib, 1, 2, s) || sa_stream_sync(s)) && !ret)
		ret = -EIO;
	starpu_free_on_node((unsigned)hip_node, (uintptr_t)dinfos, ib);
	starpu_data_unregister(wsP);
	starpu_data_unregister(wsT);
	starpu_data_unregister(hW);
	starpu_data_unregister(hA);
	if (factor_us)
		*factor_us = starpu_timing_now() - t1;
	for (int k = 0; k < nblocks && !f->info; k++)
		if (infos[k])
			f->info = f->off[k] + infos[k];
out:
	free(infos);
	free(Wh);
	if (ret)
		chol_mixed_factor_release(f);
	return ret;
}

size_t chol_mixed_refine_worksize(int N, int nblocks, int nrhs)
{
	if (N <= 0 || nblocks <= 0 || nblocks > N || nrhs <= 0)
		return 0;
	/* split-k room of the largest thin product: the residual (N x N) and
	 * every block's triangular product and panel updates */
	size_t ws = sa_dgemm_thin_worksize(N, nrhs, N);
	for (int k = 0; k < nblocks; k++) {
		size_t sz, o;
		starpu_filter_nparts_compute_chunk_size_and_offset((unsigned)N, (unsigned)nblocks, 1, (unsigned)k, 1, &sz, &o);
		const int w = (int)sz, m[3] = {w, N - (int)(o + sz), (int)o};
		for (int q = 0; q < 3; q++)
			if (m[q] > 0 && sa_dgemm_thin_worksize(m[q], nrhs, w) > ws)
				ws = sa_dgemm_thin_worksize(m[q], nrhs, w);
	}
	/* the residual's partial sums share that room */
	if (sa_dsymm_lower_resid_worksize(N, nrhs) > ws)
		ws = sa_dsymm_lower_resid_worksize(N, nrhs);
	/* residual and solve workspace (N x nrhs each) and 3 nrhs norms */
	return ws + 2 * align256((size_t)N * nrhs * sizeof(double)) + align256((size_t)3 * nrhs * sizeof(double));
}

/* ---- refinement as runtime tasks ------------------------------------------
 * One refinement step is four tasks ordered by their data: SOLVE (R U, R W,
 * RW r, W y: r <- (U^T U)^{-1} r with the FP32 factor, solve_fp32_factor),
 * UPDATE (R r, RW x: x += r), RESID (R A, R b, R x, W r: r = b - A x in
 * FP64) and NORMS (R r, R x, W the host vector of ||r_j||^2, ||x_j||^2).  The
 * split-k room of the thin products is shared by the tasks of one
 * refinement, which never run at the same time (each reads what the
 * previous one wrote). */

struct refine_arg {
	const struct chol_mixed_factor *f;
	int nrhs;
	void *work;
	size_t wb;
};

#define DPTR(i) ((double *)STARPU_MATRIX_GET_PTR(descr[i]))

/* U (R), W (R), r (RW), y (W) */
static void hip_ref_solve(void *descr[], void *arg)
{
	const struct refine_arg *a = arg;
	const struct chol_mixed_factor *f = a->f;
	hip_check(solve_fp32_factor(starpu_hip_get_local_stream(), f->N, f->nblocks, f->off, FPTR(0), MLD(0), FPTR(1),
	                            MLD(1), DPTR(2), MLD(2), DPTR(3), a->nrhs, a->work, a->wb));
}

/* r (R), x (RW) */
static void hip_ref_update(void *descr[], void *arg)
{
	(void)arg;
	hip_check(sa_daxpy2d(starpu_hip_get_local_stream(), MNX(1), MNY(1), 1.0, DPTR(0), MLD(0), DPTR(1), MLD(1)));
}

/* A (R, its lower triangle), b (R), x (R), r (W): r = b - A x */
static void hip_ref_resid(void *descr[], void *arg)
{
	const struct refine_arg *a = arg;
	const int n = MNX(1), r = MNY(1);
	hip_check(sa_dsymm_lower_resid(starpu_hip_get_local_stream(), n, r, DPTR(0), MLD(0), DPTR(2), MLD(2), DPTR(1),
	                               MLD(1), DPTR(3), MLD(3), a->work, a->wb));
}

/* r (R), x (R), norms (W, 2 nrhs doubles: ||r_j||^2 then ||x_j||^2) */
static void hip_ref_norms(void *descr[], void *arg)
{
	(void)arg;
	hipStream_t s = starpu_hip_get_local_stream();
	const int n = MNX(0), r = MNY(0);
	double *out = (double *)STARPU_VECTOR_GET_PTR(descr[2]);
	hip_check(sa_dcolsumsq(s, n, r, DPTR(0), MLD(0), out));
	hip_check(sa_dcolsumsq(s, n, r, DPTR(1), MLD(1), out + r));
}

static struct starpu_codelet cl_ref_solve = {
	.hip_funcs = {hip_ref_solve}, .hip_flags = {STARPU_HIP_ASYNC}, .nbuffers = 4,
	.modes = {STARPU_R, STARPU_R, STARPU_RW, STARPU_W}, .name = "cl_REFINE_solve"};
static struct starpu_codelet cl_ref_update = {
	.hip_funcs = {hip_ref_update}, .hip_flags = {STARPU_HIP_ASYNC}, .nbuffers = 2,
	.modes = {STARPU_R, STARPU_RW}, .name = "cl_REFINE_update"};
static struct starpu_codelet cl_ref_resid = {
	.hip_funcs = {hip_ref_resid}, .hip_flags = {STARPU_HIP_ASYNC}, .nbuffers = 4,
	.modes = {STARPU_R, STARPU_R, STARPU_R, STARPU_W}, .name = "cl_REFINE_residual"};
static struct starpu_codelet cl_ref_norms = {
	.hip_funcs = {hip_ref_norms}, .hip_flags = {STARPU_HIP_ASYNC}, .nbuffers = 3,
	.modes = {STARPU_R, STARPU_R, STARPU_W}, .name = "cl_REFINE_norms"};

int chol_mixed_refine(const struct chol_mixed_factor *f, const double *A, int lda, const double *B, int ldb, int nrhs,
                      double *X, int ldx, double tol, int maxiter, hipStream_t s, void *work, size_t work_bytes,
                      int *iterations, double *final_berr, double *berr_hist, int hist_len)
{
	const int N = f->N;
	if (nrhs <= 0 || ldb < N || ldx < N || lda < N || maxiter < 0)
		return -EINVAL;
	const size_t need = chol_mixed_refine_worksize(N, f->nblocks, nrhs);
	if (!work || work_bytes < need)
		return -EINVAL;
	if (!starpu_is_initialized())
		return -ENODEV;
	char *w = (char *)work;
	double *Rm = (double *)w;
	w += align256((size_t)N * nrhs * sizeof(double));
	double *Ym = (double *)w;
	w += align256((size_t)N * nrhs * sizeof(double));
	double *nrm = (double *)w;
	w += align256((size_t)3 * nrhs * sizeof(double));
	struct refine_arg ra = {f, nrhs, w, work_bytes - (size_t)(w - (char *)work)};
	double *hn = malloc((size_t)3 * nrhs * sizeof(double));
	if (!hn)
		return -ENOMEM;
	/* x = 0, r = b and ||b_j||^2, after the caller's work on its stream */
	int e = sa_dlaset(s, N, nrhs, 0.0, 0.0, X, ldx);
	e = e ? e : sa_dlacpy(s, N, nrhs, B, ldb, Rm, N);
	e = e ? e : sa_dcolsumsq(s, N, nrhs, B, ldb, nrm + 2 * nrhs);
	e = e ? e : sa_stream_sync(s);
	e = e ? e : sa_memcpy_d2h(hn + 2 * nrhs, nrm + 2 * nrhs, (size_t)nrhs * sizeof(double));
	if (e) {
		free(hn);
		return -EIO;
	}
	const unsigned node = (unsigned)f->node;
	starpu_data_handle_t hU, hW, hA, hB, hX, hR, hY, hN;
	starpu_matrix_data_register(&hU, node, (uintptr_t)f->U, (size_t)f->ldu, (size_t)N, (size_t)N, sizeof(float));
	starpu_matrix_data_register(&hW, node, (uintptr_t)f->W, (size_t)f->ldw, (size_t)f->nbmax, (size_t)N, sizeof(float));
	starpu_matrix_data_register(&hA, node, (uintptr_t)A, (size_t)lda, (size_t)N, (size_t)N, sizeof(double));
	starpu_matrix_data_register(&hB, node, (uintptr_t)B, (size_t)ldb, (size_t)N, (size_t)nrhs, sizeof(double));
	starpu_matrix_data_register(&hX, node, (uintptr_t)X, (size_t)ldx, (size_t)N, (size_t)nrhs, sizeof(double));
	starpu_matrix_data_register(&hR, node, (uintptr_t)Rm, (size_t)N, (size_t)N, (size_t)nrhs, sizeof(double));
	starpu_matrix_data_register(&hY, node, (uintptr_t)Ym, (size_t)N, (size_t)N, (size_t)nrhs, sizeof(double));
	starpu_vector_data_register(&hN, STARPU_MAIN_RAM, (uintptr_t)hn, (size_t)2 * nrhs, sizeof(double));
	/* x = 0, r = b; then repeat: d = (U^T U)^{-1} r, x += d, r = b - A x */
	int it = 0;
	double berr = 1.0;
	for (;;) {
		if (it > 0) {
			e = starpu_task_insert(&cl_ref_resid, STARPU_R, hA, STARPU_R, hB, STARPU_R, hX, STARPU_W, hR,
			                       STARPU_CL_ARGS_NFREE, &ra, sizeof(ra), STARPU_NAME, "REFINE_residual", 0);
			e = e ? e : starpu_task_insert(&cl_ref_norms, STARPU_R, hR, STARPU_R, hX, STARPU_W, hN,
			                               STARPU_NAME, "REFINE_norms", 0);
			e = e ? e : starpu_data_acquire(hN, STARPU_R);
			if (e)
				break;
			berr = max_berr(N, nrhs, f->anorm, hn, hn + nrhs, hn + 2 * nrhs);
			starpu_data_release(hN);
			if (berr_hist && it - 1 < hist_len)
				berr_hist[it - 1] = berr;
			if (berr <= tol || it > maxiter)
				break;
		}
		e = starpu_task_insert(&cl_ref_solve, STARPU_R, hU, STARPU_R, hW, STARPU_RW, hR, STARPU_W, hY,
		                       STARPU_CL_ARGS_NFREE, &ra, sizeof(ra), STARPU_NAME, "REFINE_solve", 0);
		e = e ? e : starpu_task_insert(&cl_ref_update, STARPU_R, hR, STARPU_RW, hX, STARPU_NAME, "REFINE_update", 0);
		if (e)
			break;
		it++;
	}
	/* unregistering waits for the tasks still using each handle */
	starpu_data_unregister(hN);
	starpu_data_unregister(hY);
	starpu_data_unregister(hR);
	starpu_data_unregister(hX);
	starpu_data_unregister(hB);
	starpu_data_unregister(hA);
	starpu_data_unregister(hW);
	starpu_data_unregister(hU);
	free(hn);
	if (iterations)
		*iterations = it;
	if (final_berr)
		*final_berr = berr;
	return e ? (e < 0 ? e : -EIO) : 0;
}

/* The chain on the GPU's reserved worker, when it has one
 * (SA_MIXED_CRITICAL_WORKER=0 turns the routing off). */
static int critical_worker_enabled(void)
{
	const char *e = getenv("SA_MIXED_CRITICAL_WORKER");
	return !(e && atoi(e) == 0);
}

int chol_mixed_solve(int hip_node, const double *A, int lda, int N, int nblocks, const double *B, int ldb, int nrhs,
                     double *X, int ldx, int prec, double tol, int maxiter, int prio, struct chol_mixed_stats *st)
{
	if (N <= 0 || nblocks <= 0 || nblocks > N || lda < N || ldb < N || ldx < N || nrhs <= 0 || prec < 0 ||
	    prec > 2 || maxiter < 0 || !A || !B || !X)
		return -EINVAL;
	if (!starpu_is_initialized())
		return -ENODEV;
	struct chol_mixed_stats local;
	if (!st)
		st = &local;
	memset(st, 0, sizeof(*st));
	hipStream_t s;
	if (sa_stream_create(&s))
		return -ENODEV;
	const size_t wb = chol_mixed_refine_worksize(N, nblocks, nrhs);
	void *work = sa_dev_malloc(wb);
	struct chol_mixed_factor f;
	memset(&f, 0, sizeof(f));
	int ret = work ? 0 : -ENOMEM;
	const double t0 = starpu_timing_now();
	const size_t ntasks0 = starpu_trace_count();
	if (!ret)
		ret = mixed_factor(hip_node, A, lda, N, nblocks, prec, prio, s, &f, &st->convert_us, &st->factor_us,
		                   critical_worker_enabled());
	st->ntasks = (long)(starpu_trace_count() - ntasks0);
	st->flops_factor = chol_flops_potrf((double)N);
	const double t2 = starpu_timing_now();
	if (!ret) {
		st->info = f.info;
		if (!f.info) {
			ret = chol_mixed_refine(&f, A, lda, B, ldb, nrhs, X, ldx, tol, maxiter, s, work, wb, &st->iterations,
			                        &st->final_berr, st->berr, CHOL_MIXED_MAXHIST);
			st->converged = st->final_berr <= tol;
		}
	}
	sa_stream_sync(s);
	const double t3 = starpu_timing_now();
	st->refine_us = t3 - t2;
	st->elapsed_us = t3 - t0;
	chol_mixed_factor_release(&f);
	if (work)
		sa_dev_free(work);
	sa_stream_destroy(s);
	return ret;
}
