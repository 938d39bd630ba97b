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

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CPU implementations of the tile codelets, with the StarPU codelet
 * signature (void *buffers[], void *cl_arg).  Buffers of the factorisation:
 * POTRF (RW A_kk, W info_k: a 1-int vector); TRSM (R A_kk, RW A_mk); SYRK
 * (R A_nk, RW A_nn); GEMM (R A_mk, R A_nk, RW A_mn).  Buffers of the solve
 * (b_k: row block k of the right-hand sides; W_k = L_kk^{-T}; the last
 * buffer of the solve codelets is a scratch area a CPU codelet may ignore):
 * TRTRI (R L_kk, W W_k, scratch); TRSM_LN (R L_kk, R W_k, RW b_k, scratch:
 * b_k <- L_kk^{-1} b_k); GEMM_NN (R L_mk, R b_k, RW b_m, scratch: b_m -=
 * L_mk b_k); TRSM_LT (R L_kk, R W_k, RW b_k, scratch: b_k <- L_kk^{-T}
 * b_k); GEMM_TN (R L_km, R b_k, RW b_m, scratch: b_m -= L_km^T b_k).  A
 * NULL entry leaves that codelet GPU-only. */
struct chol_cpu_kernels {
	void (*potrf)(void **, void *);
	void (*trsm)(void **, void *);
	void (*syrk)(void **, void *);
	void (*gemm)(void **, void *);
	void (*trsm_ln)(void **, void *);
	void (*gemm_nn)(void **, void *);
	void (*trsm_lt)(void **, void *);
	void (*gemm_tn)(void **, void *);
	void (*trtri)(void **, void *);
};

/* NULL removes the CPU implementations (HIP-only codelets). */
void chol_set_cpu_kernels(const struct chol_cpu_kernels *k);

/* Task kinds, in the order of chol_stats.kernel_us / kernel_n. */
#define CHOL_NKINDS 9   /* POTRF, TRSM, SYRK, GEMM, TRSM_LN, GEMM_NN, TRSM_LT, GEMM_TN, TRTRI */

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
 * resident in HBM).  prio: 1 uses the reference's base priorities
 * (cholesky_implicit.c:125-213), 0 gives every task the default priority.
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

/* Solve A X = B on the GPU node `hip_node`: A (N x N FP64, both triangles
 * stored, device memory), B (N x nrhs, device) and X (N x nrhs, device,
 * output).  A is not modified.  The factorisation runs as a task DAG on an
 * FP32 copy (upper triangle) with POTRF in FP64 and the updates in `prec`;
 * refinement stops when the backward error is <= tol or after maxiter
 * refinement steps.  Returns 0 (check stats->converged and stats->info) or
 * a negative errno. */
int chol_mixed_solve(int hip_node, const double *A, int lda, int N, int nblocks,
                     const double *B, int ldb, int nrhs, double *X, int ldx,
                     int prec, double tol, int maxiter, int prio,
                     struct chol_mixed_stats *stats);

/* FLOPS_SPOTRF(n) of cholesky.h:95-98 (the same count for FP64). */
double chol_flops_potrf(double n);

#ifdef __cplusplus
}
#endif
#endif
