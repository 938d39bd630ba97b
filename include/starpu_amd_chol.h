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

/* CPU implementations of the four tile codelets, with the StarPU codelet
 * signature (void *buffers[], void *cl_arg).  Buffers: POTRF (RW A_kk,
 * W info_k: a 1-int vector); TRSM (R A_kk, RW A_mk); SYRK (R A_nk, RW
 * A_nn); GEMM (R A_mk, R A_nk, RW A_mn). */
struct chol_cpu_kernels {
	void (*potrf)(void **, void *);
	void (*trsm)(void **, void *);
	void (*syrk)(void **, void *);
	void (*gemm)(void **, void *);
};

/* NULL removes the CPU implementations (HIP-only codelets). */
void chol_set_cpu_kernels(const struct chol_cpu_kernels *k);

struct chol_stats {
	double elapsed_us;       /* first submission to starpu_task_wait_for_all */
	double flops;            /* FLOPS_DPOTRF(N), cholesky.h:95-98 */
	int info;                /* 0 or 1-based column of the first bad pivot */
	long ntasks;
	/* summed device time of each kind of tile kernel (HIP events), and
	 * the number of such tasks that ran on a GPU */
	double kernel_us[4];     /* POTRF, TRSM, SYRK, GEMM */
	long kernel_n[4];
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

/* FLOPS_SPOTRF(n) of cholesky.h:95-98 (the same count for FP64). */
double chol_flops_potrf(double n);

#ifdef __cplusplus
}
#endif
#endif
