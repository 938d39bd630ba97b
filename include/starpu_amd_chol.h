/*
 * starpu_amd_chol.h -- the tiled Cholesky application on the starpu_amd
 * runtime: the FP64 POTRF/TRSM/SYRK/GEMM task DAG of the reference's
 * examples/cholesky/cholesky_implicit.c:215-460 over a 2-D tile grid, with
 * HIP codelets that call the gfx950 kernels of starpu_amd_kernels.h.
 *
 * CPU codelets are supplied by the caller (chol_set_cpu_kernels), which is
 * how the CPU baseline runs the same DAG on host cores under dmda.
 */
#ifndef STARPU_AMD_CHOL_H
#define STARPU_AMD_CHOL_H

#include <stddef.h>
#include <stdint.h>

#ifndef STARPU_AMD_HIP_STREAM_T
#define STARPU_AMD_HIP_STREAM_T
typedef struct ihipStream_t *hipStream_t;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* CPU implementations of the tile codelets, with the StarPU codelet
 * signature (void *buffers[], void *cl_arg).  W_k = L_kk^{-T} and
 * Linv_k = L_kk^{-1} are formed once per diagonal tile by TRTRI; a CPU
 * codelet may ignore the inverse buffers it reads and the scratch buffer
 * (the last buffer of POTRF, TRSM, TRTRI and the solve codelets), but
 * POTRF and TRTRI must write W_k and Linv_k, which the HIP codelets read.
 * Buffers of the factorisation: POTRF (RW A_kk, W info_k: a 1-int vector,
 * W W_k, W Linv_k, scratch); TRTRI, in a solve without the factorisation
 * (R L_kk, W W_k, W Linv_k, scratch); TRSM (R L_kk, R Linv_k, RW A_mk, scratch:
 * A_mk <- A_mk L_kk^{-T}); SYRK (R A_nk, RW A_nn); GEMM (R A_mk, R A_nk,
 * RW A_mn).  The solve has one task per block row and sweep (b_k: row
 * block k of the right-hand sides; a variable number of buffers; cl_arg
 * points to an int, the number of (L tile, b_m) pairs):
 * SOLVE_FWD k (R L_kk, R W_k, RW b_k, then R L_mk, RW b_m for each m > k,
 * scratch): b_k <- L_kk^{-1} b_k, then b_m -= L_mk b_k in increasing m;
 * SOLVE_BWD k (R L_kk, R W_k, RW b_k, then R L_km, RW b_m for each m < k,
 * scratch): b_k <- L_kk^{-T} b_k, then b_m -= L_km^T b_k in increasing m.
 * A NULL entry leaves that codelet GPU-only. */
struct chol_cpu_kernels {
	void (*potrf)(void **, void *);
	void (*trsm)(void **, void *);
	void (*syrk)(void **, void *);
	void (*gemm)(void **, void *);
	void (*solve_fwd)(void **, void *);
	void (*solve_bwd)(void **, void *);
	void (*trtri)(void **, void *);
};

/* NULL removes the CPU implementations (HIP-only codelets). */
void chol_set_cpu_kernels(const struct chol_cpu_kernels *k);

/* Task kinds, in the order of chol_stats.kernel_us / kernel_n. */
#define CHOL_NKINDS 7   /* POTRF, TRSM, SYRK, GEMM, SOLVE_FWD, SOLVE_BWD, TRTRI */

struct chol_stats {
	double elapsed_us;       /* first submission to starpu_task_wait_for_all */
	double flops;            /* FLOPS_DPOTRF(N) (cholesky.h:95-98) + 2 N^2 nrhs for a solve */
	int info;                /* 0 or 1-based column of the first bad pivot */
	long ntasks;
	/* summed device time of each kind of tile kernel (HIP events), and
	 * the number of such tasks that ran on a GPU */
	double kernel_us[CHOL_NKINDS];
	long kernel_n[CHOL_NKINDS];
};

/* Factor the N x N column-major SPD matrix A (lower) in place with
 * nblocks x nblocks tiles.  A lives on memory node `home_node`
 * (STARPU_MAIN_RAM = 0 for host memory, or the node of a GPU, see
 * starpu_hip_memory_node; then A is a device pointer and the tiles stay
 * resident in HBM).  prio selects the task priorities of the reference's
 * priority attributions (cholesky_implicit.c:125-213, its -p option):
 * 1 base, 2 bottom levels of the Kut polyhedral tool, 3 bottom levels from
 * the calibrated task times (base while the models are uncalibrated),
 * 4 PaRSEC's; 0 gives every task the default priority. 
 * The runtime must be initialised.  Returns 0, or a negative errno. */
int chol_factor(int home_node, double *A, int lda, int N, int nblocks, int prio,
                struct chol_stats *stats);

/* Solve A X = B with the lower factor L of chol_factor (same N, nblocks):
 * B (N x nrhs, leading dimension ldb, on the same memory node) is
 * overwritten with X.  No counterpart in the reference. */
int chol_solve(int home_node, const double *L, int ldl, int N, int nblocks,
               double *B, int ldb, int nrhs, int prio, struct chol_stats *stats);

/* chol_factor then chol_solve as one task graph, without a barrier between
 * the two: A is overwritten with L and B with X. */
int chol_factor_solve(int home_node, double *A, int lda, int N, int nblocks,
                      double *B, int ldb, int nrhs, int prio,
                      struct chol_stats *stats);

/* 2 N^2 nrhs: forward and backward substitution with an N x N factor. */
double chol_flops_potrs(double n, double nrhs);

/* ---- mixed precision with iterative refinement (no reference counterpart) */

/* Precision of the trailing updates (and of the panel TRSM, done as a
 * product with the FP64-formed inverse of the diagonal tile). */
#define CHOL_PREC_FP32 0     /* FP32 MFMA */
#define CHOL_PREC_BF16 1     /* BF16 MFMA, FP32 accumulation */
#define CHOL_PREC_BF16X3 2   /* BF16 MFMA on hi/lo splits: hi*hi + hi*lo + lo*hi */
#define CHOL_MIXED_MAXHIST 32

struct chol_mixed_stats {
	double elapsed_us;       /* time to solution: conversion + factorisation + refinement */
	double convert_us, factor_us, refine_us;
	double flops_factor;     /* FLOPS_DPOTRF(N) */
	int info;                /* 0 or 1-based column of the first bad pivot */
	int iterations;          /* solves with the low-precision factor */
	int converged;           /* final backward error <= tol */
	long ntasks;
	double final_berr;       /* max_j ||b_j - A x_j|| / (||A||_F ||x_j|| + ||b_j||) */
	double berr[CHOL_MIXED_MAXHIST];  /* the same after each solve */
};

/* Solve A X = B on the GPU node `hip_node`: A (N x N FP64 symmetric,
 * device memory; only its lower triangle is read), B (N x nrhs, device) and
 * X (N x nrhs, device, output).  A is not modified.  The factorisation runs as a task DAG on an
 * FP32 copy (upper triangle) with POTRF in FP64 and the updates in `prec`;
 * refinement stops when the backward error is <= tol or after maxiter
 * refinement steps.  Returns 0 (check stats->converged and stats->info) or
 * a negative errno. */
int chol_mixed_solve(int hip_node, const double *A, int lda, int N, int nblocks,
                     const double *B, int ldb, int nrhs, double *X, int ldx,
                     int prec, double tol, int maxiter, int prio,
                     struct chol_mixed_stats *stats);

/* The steps of chol_mixed_solve, for callers that keep factors resident
 * (the solve server): an FP32 factor of A on the GPU with its FP32
 * diagonal-tile inverses and ||A||_F. */
struct chol_mixed_factor {
	int N, nblocks, nbmax;
	int *off;            /* nblocks + 1 tile offsets */
	float *U;            /* FP32 upper factor, ld = ldu (device) */
	float *W;            /* W_k = U_kk^{-1} at W + off[k] * ldw (device) */
	int ldu, ldw;
	double anorm;        /* ||A||_F of the FP64 matrix */
	int info;            /* 0 or 1-based column of the first bad pivot */
	int node;            /* the GPU's memory node (of U, W and the refinement's tasks) */
};
/* Device bytes a factor holds (U and W).  A BF16x3 factorisation on tiles of
 * multiples of 128 also holds 4 N^2 bytes of bf16 panel planes while its DAG
 * runs. */
size_t chol_mixed_factor_bytes(int N, int nblocks);
/* Factor A (device, lower triangle read) with the task DAG of chol_mixed_solve.
 * `stream` carries the conversion; the call returns when the DAG is done,
 * waiting only for its own tasks (several calls may run concurrently from
 * different threads).  Times in microseconds. */
int chol_mixed_factor(int hip_node, const double *A, int lda, int N, int nblocks, int prec, int prio,
                      hipStream_t stream, struct chol_mixed_factor *f, double *convert_us, double *factor_us);
void chol_mixed_factor_release(struct chol_mixed_factor *f);
/* Device workspace of chol_mixed_refine. */
size_t chol_mixed_refine_worksize(int N, int nblocks, int nrhs);
/* Iterative refinement of A X = B with a factor of A, as runtime tasks on
 * the factor's GPU: per step an FP32-factor solve, the update x += d, the
 * FP64 residual r = b - A x (from A's lower triangle) and its norms (the host waits only for this
 * refinement's norms, so refinements of several requests run at once).
 * `stream` orders the start after the caller's earlier work on it. */
int chol_mixed_refine(const struct chol_mixed_factor *f, const double *A, int lda, const double *B, int ldb,
                      int nrhs, double *X, int ldx, double tol, int maxiter, hipStream_t stream, void *work,
                      size_t work_bytes, int *iterations, double *final_berr, double *berr_hist, int hist_len);

/* ---- solve server (no reference counterpart) ---------------------------- */

/* One SPD solve request.  A (n x n, lower triangle read), B and X live in device
 * memory; requests with the same `matrix` index share A, so a later one is
 * a re-solve that reuses the resident factor. */
struct chol_serve_request {
	int n, nrhs, matrix;
	double arrival_us;           /* offset from the start of the trace, non-decreasing */
	const double *A;
	int lda;
	const double *B;
	int ldb;
	double *X;
	int ldx;
	/* results */
	double start_us, finish_us;  /* service start and completion, from the start of the trace */
	double berr;                 /* final normwise backward error */
	int iterations;              /* solves with the FP32 factor */
	int status;                  /* 0, a pivot column (not SPD) or a negative errno */
	int factored;                /* 1 if this request ran the factorisation */
	double factor_start_us, factor_end_us;  /* its factorisation (when factored), from the start of the trace */
};

struct chol_serve_config {
	double budget_bytes;         /* HBM budget for factors and work areas */
	int slots;                   /* requests in service at once */
	int nb;                      /* tile size */
	int prec;                    /* CHOL_PREC_* of the factorisation's updates */
	int maxiter;
	int prio;
	double tol;                  /* backward error target of the refinement */
};

struct chol_serve_stats {
	double wall_us, peak_bytes;
	int factorizations, cache_hits, evictions, memory_waits, max_inflight;
};

/* Serve the requests (sorted by arrival) on the GPU node `hip_node`: each
 * is released at its arrival time, admitted when a slot and its bytes are
 * free, factored (unless its factor is resident) and refined.  Returns
 * when all are done. */
int chol_serve(int hip_node, const struct chol_serve_config *cfg, struct chol_serve_request *req, int nreq,
               struct chol_serve_stats *stats);

/* FLOPS_SPOTRF(n) of cholesky.h:95-98 (the same count for FP64). */
double chol_flops_potrf(double n);

#ifdef __cplusplus
}
#endif
#endif
