/*
 * starpu_amd_kernels.h -- C ABI of the hand-written gfx950 tile kernels.
 *
 * These entry points replace, one for one, the vendor-library calls the
 * reference's HIP codelets make on the tiled-Cholesky path
 * (examples/cholesky/cholesky_kernels.c).  They take plain device pointers,
 * sizes and a HIP stream; they never allocate device memory.  Matrices are
 * column-major with leading dimensions in elements, exactly as the
 * reference's STARPU_MATRIX_GET_PTR/NX/NY/LD triplets describe a tile
 * (include/starpu_data_interfaces.h:1370-1408 in the reference).
 *
 * Every function enqueues work on `stream` and returns without waiting.
 * Return value: 0 on success, -EINVAL (-22) for invalid sizes or leading
 * dimensions, or a positive hipError_t code if a launch failed.
 */
#ifndef STARPU_AMD_KERNELS_H
#define STARPU_AMD_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef STARPU_AMD_HIP_STREAM_T
#define STARPU_AMD_HIP_STREAM_T
/* Same type as HIP's hipStream_t, declared without the HIP headers so that
 * plain C (gcc) callers can use this ABI. */
typedef struct ihipStream_t *hipStream_t;
#endif

/* ---------------------------------------------------------------- FP64 --- */

/* Hint for the tile shape of the GEMM-based kernels: the number of such
 * launches the caller keeps in flight on one device (its HIP streams).
 * With several in flight, fewer and larger workgroups (128x128) fill the
 * device at half the operand traffic per flop.  Process-wide; returns the
 * previous value (default 1).  No reference counterpart (the reference
 * leaves tiling to hipBLAS). */
int sa_set_launch_concurrency(int n);

/* Forces the workgroup tile edge of the GEMM-based kernels (64 or 128), or
 * 0 to let the launch-concurrency hint choose (SA_GEMM_SHAPE gives the
 * starting value).  Process-wide; returns the previous value, -EINVAL for
 * another edge.  No reference counterpart (tiling is hipBLAS's in the
 * reference). */
int sa_set_gemm_shape(int edge);

/* Workgroup tile edge (64 or 128) that sa_dgemm_nt (lower = 0) or
 * sa_dsyrk_ln / the lower-triangle update (lower = 1, n ignored) would run
 * an m x n product with now; -EINVAL for an empty shape. */
int sa_gemm_tile_shape(int lower, int m, int n);

/* C(m x n) <- beta*C + alpha * A(m x k) * B(n x k)^T.
 * Replaces hipblasSgemm(handle, HIPBLAS_OP_N, HIPBLAS_OP_T, dy, dx, dz,
 * &m1, left, ld21, right, ld12, &p1, center, ld22) at
 * examples/cholesky/cholesky_kernels.c:112-115 (and the CPU
 * STARPU_SGEMM("N","T") at :75-76). */
int sa_dgemm_nt(hipStream_t stream, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc);
/* C(m x n) <- beta*C + alpha * A(m x n) * L^T with L (n x n) lower
 * triangular and zeros stored above its diagonal (the TRSM's L_kk^-1): the
 * product of sa_dgemm_nt with k = n, each wave's block of columns of C taking
 * only the k up to its last column (about half the flops; the terms left
 * out are exact zeros). */
int sa_dgemm_nt_ltri(hipStream_t stream, int m, int n, double alpha, const double *A, int lda,
                     const double *L, int ldl, double beta, double *C, int ldc);

/* lower(C(n x n)) <- beta*C + alpha * A(n x k) * A^T; the strict upper
 * triangle of C is neither read nor written.
 * Replaces hipblasSsyrk(handle, HIPBLAS_FILL_MODE_LOWER, HIPBLAS_OP_N, dx,
 * dz, &m1, left, ld21, &p1, center, ld22) at cholesky_kernels.c:190-193. */
int sa_dsyrk_ln(hipStream_t stream, int n, int k, double alpha,
                const double *A, int lda, double beta, double *C, int ldc);

/* B(m x n) <- alpha * B * L^{-T}; L is n x n lower triangular, non-unit.
 * `work` (device memory, sa_dtrsm_rltn_worksize(n) bytes) receives the
 * inverses of L's 64x64 diagonal blocks.
 * Replaces hipblasStrsm(handle, HIPBLAS_SIDE_RIGHT, HIPBLAS_FILL_MODE_LOWER,
 * HIPBLAS_OP_T, HIPBLAS_DIAG_NON_UNIT, nx21, ny21, &p1, sub11, ld11, sub21,
 * ld21) at cholesky_kernels.c:267-269. */
int sa_dtrsm_rltn(hipStream_t stream, int m, int n, double alpha,
                  const double *L, int ldl, double *B, int ldb,
                  void *work, size_t work_bytes);
size_t sa_dtrsm_rltn_worksize(int n);

/* In-place lower Cholesky factorisation of the n x n matrix A.  *dev_info
 * (device memory) receives 0, or the 1-based column of the first
 * non-positive pivot, as LAPACK dpotrf reports it.  The strict upper
 * triangle of A is not referenced.  `work` (device memory,
 * sa_dpotrf_l_worksize(n) bytes) is scratch space.
 * Replaces hipsolverDnSpotrf(handle, HIPBLAS_FILL_MODE_LOWER, nx, sub11, ld,
 * workspace, Lwork, d_info) at cholesky_kernels.c:433 (whose workspace and
 * devInfo are the codelet's STARPU_SCRATCH buffers, :551-556), and the
 * per-column hipblasSscal/Ssyr loop of :437-475. */
int sa_dpotrf_l(hipStream_t stream, int n, double *A, int lda, int *dev_info,
                void *work, size_t work_bytes);
size_t sa_dpotrf_l_worksize(int n);

/* Factorisation of one diagonal tile (n <= sa_dpotrf_tile_maxn() = 2048) in
 * a single persistent kernel launch: A <- L in place as sa_dpotrf_l, and,
 * when W is given, W <- L^{-T} (upper, zero below the diagonal) in the same
 * launch, its block items overlapping the factorisation's critical chain,
 * with Linv <- L^{-1} (lower, zero above; optional, needs W).  The inverse
 * has the same bits as sa_dtrtri_tile's on the factor.  `work` holds
 * sa_dpotrf_tile_worksize(n) bytes.  Replaces the reference's POTRF codelet
 * (cholesky_kernels.c:433) together with the inverse that the TRSM and solve
 * tasks here multiply by (no reference counterpart for the inverse).
 * *dev_info = -1 reports a schedule that stopped waiting (it cannot happen
 * with a correct kernel; the launch ends instead of hanging). */
int sa_dpotrf_tile(hipStream_t stream, int n, double *A, int lda, double *Linv, int ldl, double *W, int ldw,
                   int *dev_info, void *work, size_t work_bytes);
/* sa_dpotrf_tile with the number of workgroups that take the bulk items
 * (besides the critical chain's) chosen by the caller: 1..255, or <= 0 for
 * the default (SA_POTRF_WORKGROUPS, else 31).  More finish a tile sooner
 * while the other streams' kernels hold the compute units, fewer leave more
 * of them to those kernels.  The result has the same bits whatever the
 * count.  No reference counterpart (the reference's POTRF is a library call). */
int sa_dpotrf_tile_wg(hipStream_t stream, int n, double *A, int lda, double *Linv, int ldl, double *W, int ldw,
                      int *dev_info, void *work, size_t work_bytes, int workgroups);
size_t sa_dpotrf_tile_worksize(int n);
int sa_dpotrf_tile_maxn(void);

/* Allocation-free variant of sa_dpotrf_tile: the same kernels, with the
 * workspace in a device array the library holds on each GPU for the
 * process's lifetime (about 1 MiB), so the call allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  n is at most
 * sa_dpotrf_tile_noalloc_maxn() (2048); larger tiles return -EOVERFLOW.
 * Calls that overlap on one GPU share the workspace and race: serialise
 * them (one stream, or events). */
int sa_dpotrf_tile_noalloc(hipStream_t stream, int n, double *A, int lda, double *Linv, int ldl, double *W,
                           int ldw, int *dev_info);
int sa_dpotrf_tile_noalloc_maxn(void);

/* Linv <- L^{-1} and W <- L^{-T} of an already factored n x n lower L:
 * for n <= sa_dpotrf_tile_maxn() the inverse items of sa_dpotrf_tile's
 * kernel (the same bits as a factorisation with the inverse gives), larger
 * tiles by sa_dtrtri_lt and a transpose.  `work`:
 * sa_dpotrf_tile_worksize(n) bytes.  No reference counterpart. */
int sa_dtrtri_tile(hipStream_t stream, int n, const double *L, int ldl, double *Linv, int ldx, double *W, int ldw,
                   void *work, size_t work_bytes);

/* Dinv <- inverses of the 64x64 diagonal blocks of the n x n lower
 * triangular L (block b at Dinv + 4096 b, leading dimension 64, identity
 * past n).  The building block of sa_dtrsm_rltn; no reference counterpart. */
int sa_dtrtri_diag64(hipStream_t stream, int n, const double *L, int ldl, double *Dinv);

/* ------------------------------------------------------- FP64 solve --- */
/* The solve phase has no counterpart in the reference, whose Cholesky
 * example stops after the factorisation (cholesky_implicit.c:325-334).
 * These are the tile operations of LAPACK dpotrs on the lower factor:
 * L y = b (dtrsm Left, Lower, NoTrans, NonUnit), L^T x = y (dtrsm Left,
 * Lower, Trans, NonUnit) and the block updates between tiles (dgemm N,N
 * and T,N).  They are written for a thin block of right-hand sides
 * (n up to a few tens): L is streamed once, the right-hand sides stay on
 * chip. */

/* B(m x n) <- alpha * L^{-1} * B; L is m x m lower triangular, non-unit. */
int sa_dtrsm_llnn(hipStream_t stream, int m, int n, double alpha,
                  const double *L, int ldl, double *B, int ldb);

/* B(m x n) <- alpha * L^{-T} * B; L is m x m lower triangular, non-unit.
 * `work` (device memory, sa_dtrsm_lltn_worksize(m) bytes) receives L^T of
 * one diagonal chunk of up to 1024 rows. */
int sa_dtrsm_lltn(hipStream_t stream, int m, int n, double alpha,
                  const double *L, int ldl, double *B, int ldb,
                  void *work, size_t work_bytes);
size_t sa_dtrsm_lltn_worksize(int m);

/* C(m x n) <- beta*C + alpha * A(m x k) * B(k x n).
 * `work` (device memory, sa_dgemm_thin_worksize(m, n, k) bytes) holds the
 * partial sums of a product split over k; with NULL the product runs in one
 * k slice (correct, with fewer loads in flight for small m). */
int sa_dgemm_nn(hipStream_t stream, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc, void *work, size_t work_bytes);

/* C(m x n) <- beta*C + alpha * A(k x m)^T * B(k x n); `work` as sa_dgemm_nn. */
int sa_dgemm_tn(hipStream_t stream, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc, void *work, size_t work_bytes);
size_t sa_dgemm_thin_worksize(int m, int n, int k);

/* W(n x n) <- L^{-T} for the lower triangular non-unit L: an upper
 * triangular matrix, zeros stored below its diagonal.  `work`:
 * sa_dtrtri_lt_worksize(n) bytes.  Lets the tile solves run as thin GEMMs
 * (sa_dtrmm_lu): L^{-1} b = W^T b and L^{-T} b = W b. */
int sa_dtrtri_lt(hipStream_t stream, int n, const double *L, int ldl,
                 double *W, int ldw, void *work, size_t work_bytes);
size_t sa_dtrtri_lt_worksize(int n);

/* B(m x n) <- alpha * op(U) * B with U (m x m) upper triangular and its
 * strict lower part stored as zeros (as sa_dtrtri_lt leaves it); op(U) = U^T
 * for trans = 1, U for trans = 0 (dtrmm Left, Upper, NonUnit).  `work`:
 * sa_dtrmm_lu_worksize(m, n) bytes (a copy of B and the split-k sums). */
int sa_dtrmm_lu(hipStream_t stream, int trans, int m, int n, double alpha,
                const double *U, int ldu, double *B, int ldb,
                void *work, size_t work_bytes);
size_t sa_dtrmm_lu_worksize(int m, int n);

/* ------------------------------------------------ mixed precision --- */
/* No counterpart in the reference.  The FP32 working copy of A is factored
 * as A = U^T U on its upper triangle, so the update operands are
 * k-contiguous. */

/* C(m x n) <- beta*C + alpha * A(k x m)^T * B(k x n), FP32 storage and
 * accumulation.  mode 0: FP32 MFMA; 1: BF16 MFMA on bf16(A), bf16(B);
 * 2: BF16 MFMA on x = hi + lo splits (hi*hi + hi*lo + lo*hi).  upper = 1
 * (needs m == n) writes only C(i, j) with i <= j. */
int sa_sgemm_tn(hipStream_t stream, int mode, int upper, int m, int n, int k,
                float alpha, const float *A, int lda, const float *B, int ldb,
                float beta, float *C, int ldc);

/* C(m x n) <- beta*C + alpha * U^T * B with U (m x m) upper triangular and
 * its strict lower part stored as zeros (the mixed TRSM's W_k = U_kk^-1):
 * sa_sgemm_tn with k = m, the k-range of each 128-row block of C stopped
 * at its last row where the staged kernels run (about half the products;
 * the terms left out are exact zeros). */
int sa_sgemm_tn_utri(hipStream_t stream, int mode, int m, int n, float alpha, const float *U, int ldu,
                     const float *B, int ldb, float beta, float *C, int ldc);

/* The BF16x3 product of sa_sgemm_tn (mode 2) on operands split ahead of
 * time: Ah/Al (and Bh/Bl) are the bf16 planes hi = bf16(x), lo = bf16(x - hi)
 * of A (and B), laid out like the FP32 matrix with leading dimension lda
 * (ldb), as sa_split_bf16x2 writes them.  The result is bit-identical to
 * sa_sgemm_tn mode 2 on the FP32 operands.  Only for m and n multiples of
 * 128, k a multiple of 32, lda and ldb multiples of 8 and 16-byte aligned
 * planes: sa_sgemm_tn_split_ok tells. */
int sa_sgemm_tn_split_ok(int m, int n, int k, const void *Ah, const void *Al, int lda,
                         const void *Bh, const void *Bl, int ldb);
int sa_sgemm_tn_split(hipStream_t stream, int upper, int m, int n, int k, float alpha,
                      const void *Ah, const void *Al, int lda, const void *Bh, const void *Bl,
                      int ldb, float beta, float *C, int ldc);
/* Staging of sa_sgemm_tn_split: 1 (default) copies the planes from memory
 * straight into LDS, 0 through registers; both give the same bits.  A
 * process-wide setting; returns the previous one, -EINVAL for another
 * value. */
int sa_set_sgemm_direct(int on);
/* H <- bf16(A), L <- bf16(A - H) over m x n (the BF16x3 split; H and L
 * share the leading dimension ldh). */
int sa_split_bf16x2(hipStream_t stream, int m, int n, const float *A, int lda, void *H, void *L,
                    int ldh);

/* U(j, i) <- (float) A(i, j) for i >= j: FP64 lower triangle to FP32 upper. */
int sa_dlag2s_lower_to_upper(hipStream_t stream, int n, const double *A, int lda,
                             float *U, int ldu);
/* The same conversion, reading A's lower triangle once, with
 * *fro2 (device) <- ||A||_F^2 of the symmetric A it stands for (diagonal
 * once, the rest twice; summed in a fixed order).  work: the worksize's
 * bytes of device memory. */
size_t sa_dlag2s_lower_to_upper_fro_worksize(int n);
int sa_dlag2s_lower_to_upper_fro(hipStream_t stream, int n, const double *A, int lda, float *U, int ldu,
                                 double *fro2, void *work, size_t work_bytes);
/* L(i, j) <- (double) U(j, i) for i >= j: FP32 upper triangle to FP64 lower. */
int sa_slag2d_upper_to_lower(hipStream_t stream, int n, const float *U, int ldu,
                             double *L, int ldl);
/* B(m x n) <- (float) A (LAPACK dlag2s without the overflow check). */
int sa_dlag2s(hipStream_t stream, int m, int n, const double *A, int lda, float *B, int ldb);
/* The FP32 outputs of a diagonal tile factored in FP64, in one launch:
 * U's upper triangle <- L's lower triangle transposed (as
 * sa_dlag2s_lower_to_upper) and W32 <- W, n x n (as sa_dlag2s); the same
 * bits as those two calls.  No reference counterpart (the mixed tile
 * POTRF's epilogue). */
int sa_tile_outputs_f32(hipStream_t stream, int n, const double *L, int ldl, float *U, int ldu, const double *W,
                        int ldw, float *W32, int ldw32);
/* X(m x n) <- X + alpha * D, FP64. */
int sa_daxpy2d(hipStream_t stream, int m, int n, double alpha, const double *D, int ldd,
               double *X, int ldx);

/* R(n x r) <- B - A X for a symmetric A given by its lower triangle (the
 * strict upper triangle is never read): the FP64 residual of the
 * refinement, each stored element read once (LAPACK dsymm Left, Lower,
 * alpha = -1, beta = 1, into a separate output).  work: the worksize's
 * bytes of device memory (per-block partial sums, added in a fixed
 * order). */
size_t sa_dsymm_lower_resid_worksize(int n, int r);
int sa_dsymm_lower_resid(hipStream_t stream, int n, int r, const double *A, int lda, const double *X, int ldx,
                         const double *B, int ldb, double *R, int ldr, void *work, size_t work_bytes);

/* sa_dgemm_nn / sa_dgemm_tn / sa_dtrmm_lu with an FP32 matrix A (or U)
 * and FP64 right-hand sides and arithmetic: the solves with the FP32
 * factor inside iterative refinement.  Workspaces as for the FP64 forms. */
int sa_gemm_nn_sd(hipStream_t stream, int m, int n, int k, double alpha,
                  const float *A, int lda, const double *B, int ldb,
                  double beta, double *C, int ldc, void *work, size_t work_bytes);
int sa_gemm_tn_sd(hipStream_t stream, int m, int n, int k, double alpha,
                  const float *A, int lda, const double *B, int ldb,
                  double beta, double *C, int ldc, void *work, size_t work_bytes);
int sa_trmm_lu_sd(hipStream_t stream, int trans, int m, int n, double alpha,
                  const float *U, int ldu, double *B, int ldb,
                  void *work, size_t work_bytes);

/* ------------------------------------------------ batched systems --- */
/* No counterpart in the reference.  Many independent SPD systems of
 * uneven sizes, factored and solved by one persistent kernel whose
 * workgroups pull systems from a device-side queue. */

/* Doubles of block-inverse storage one system of size n needs. */
size_t sa_batched_dinv_elems(int n);

/* Panel-update shape of sa_dpotrs_batched: rows blocks of 16 per wave in
 * the left-looking panel update (2 or 4; 0 restores the default 2), a
 * process-wide setting the autotuner sweeps; returns the previous value or
 * -EINVAL. */
int sa_set_batched_rows(int rows_blocks);

/* For each system i: A_i (n[i] x n[i], ld n[i], at A + offA[i]) is factored
 * in place (lower), info[i] = 0 or the 1-based column of the first
 * non-positive pivot, and if nrhs > 0 the nrhs right-hand sides at
 * B + offB[i] (ld n[i]) are overwritten with the solution.  Systems are
 * taken in the order of `order` (a permutation of 0..nsys-1; largest
 * first balances the load).  Dinv (sa_batched_dinv_elems(n[i]) doubles at
 * Dinv + offD[i]) receives the block inverses.  `counter` is one device
 * int, reset by a kernel at the start of the call, so the call can be
 * captured in a graph.  `grid` workgroups of 256 threads stay resident
 * while the queue lasts.  All pointers are device memory. */
int sa_dpotrs_batched(hipStream_t stream, int nsys, const int *order, const int *n, const long *offA,
                      const long *offB, const long *offD, double *A, double *B, int nrhs,
                      double *Dinv, int *info, int *counter, int grid);

/* Mixed-precision variant (the refinement scheme of LAPACK dsposv): per
 * system, lower(A_i) is rounded to FP32 into Af (Af + offA[i], ld n[i]) and
 * factored there in FP32 (block inverses into Dinvf at offD[i]); each
 * right-hand side b (B + offB[i] + c n[i]) is solved by x = 0, r = b, then
 * repeatedly d = (L L^T)^{-1} fl32(r) in FP32, x += d, r = b - A x in FP64,
 * until ||r|| / (||A||_F ||x|| + ||b||) <= tol or maxiter steps.  A and B
 * are only read; X receives x, R (FP64) and Df (FP32) are workspaces at the
 * offsets of B.  iters[i nrhs + c] and berr[i nrhs + c] receive the steps
 * taken and the final backward error.  info[i] > 0 reports a non-positive
 * pivot of the FP32 factor: X of that system is not written (iters 0, berr
 * +inf) and the caller factors it in FP64 (sa_dpotrs_batched over just the
 * failed systems), as dsposv falls back to dpotrf.  Queue, counter and grid
 * as for sa_dpotrs_batched. */
int sa_dposv_batched_ir(hipStream_t stream, int nsys, const int *order, const int *n, const long *offA,
                        const long *offB, const long *offD, const double *A, const double *B, double *X,
                        int nrhs, float *Af, float *Dinvf, double *R, float *Df, int *info, int *iters,
                        double *berr, int maxiter, double tol, int *counter, int grid);

/* Capture the work enqueued on `stream` between sa_graph_begin and
 * sa_graph_end into an executable graph; sa_graph_launch replays it. */
int sa_graph_begin(hipStream_t stream);
int sa_graph_end(hipStream_t stream, void **exec);
int sa_graph_launch(void *exec, hipStream_t stream);
int sa_graph_destroy(void *exec);

/* ------------------------------------------------ device-memory helpers --- */
/* Thin wrappers over the HIP runtime used by tests, the benchmark and the
 * runtime's HIP memory node.  They have no counterpart in the reference's
 * codelets; they mirror the node_ops malloc/free/copy entries of
 * src/datawizard/node_ops.h:91-219. */
int sa_device_count(void);
int sa_set_device(int dev);
void *sa_dev_malloc(size_t bytes);
int sa_dev_free(void *p);
int sa_memcpy_h2d(void *dst, const void *src, size_t bytes);
int sa_memcpy_d2h(void *dst, const void *src, size_t bytes);
int sa_memcpy_d2d(void *dst, const void *src, size_t bytes);
int sa_memcpy2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                      size_t width_bytes, size_t height, int kind,
                      hipStream_t stream);
int sa_stream_create(hipStream_t *stream);
int sa_stream_destroy(hipStream_t stream);
/* Events ordering one stream's work after another's (no timing). */
int sa_event_create(void **event);
int sa_event_record(void *event, hipStream_t stream);
int sa_stream_wait_event(hipStream_t stream, void *event);
int sa_event_destroy(void *event);
int sa_stream_sync(hipStream_t stream);
int sa_device_sync(void);
const char *sa_error_string(int err);
/* Elapsed milliseconds of the work enqueued on `stream` between two calls:
 * sa_timer_start records a start event, sa_timer_stop a stop event and
 * waits for it. */
int sa_timer_start(hipStream_t stream, void **timer);
int sa_timer_stop(void *timer, float *ms);

/* ------------------------------------------------- synthetic matrices --- */
/* Fill the m x n matrix X (ld) with i.i.d. values from a counter-based
 * generator keyed by (seed, i, j): dist 0 = uniform[-1,1], 1 = N(0,1). */
int sa_fill_random(hipStream_t stream, int m, int n, double *X, int ld,
                   uint64_t seed, int dist);
/* Add `shift` to the n diagonal entries of X. */
int sa_add_diagonal(hipStream_t stream, int n, double *X, int ld, double shift);
/* Copy the lower triangle of X into its upper triangle. */
int sa_symmetrize_lower(hipStream_t stream, int n, double *X, int ld);

/* out[j] <- sum_i X(i, j)^2 for the n columns of the m x n matrix X (out is
 * device memory): the squared column norms behind the normwise backward
 * error ||b - A x|| / (||A|| ||x|| + ||b||) of a solve. */
int sa_dcolsumsq(hipStream_t stream, int m, int n, const double *X, int ld, double *out);

/* B(n x m) <- A(m x n)^T on the device. */
int sa_dtranspose(hipStream_t stream, int m, int n, const double *A, int lda, double *B, int ldb);

/* A(i, j) <- diag for i == j, offdiag elsewhere, on the m x n region
 * (LAPACK dlaset). */
int sa_dlaset(hipStream_t stream, int m, int n, double offdiag, double diag, double *A, int lda);

/* B(m x n) <- A(m x n) on the device, any leading dimensions.  The kernel
 * form of the 2-D device-to-device copy the matrix interface performs
 * between replicates (src/datawizard/interfaces/matrix_interface.c:578-603,
 * a cudaMemcpy2DAsync there); also the known-size plain copy that scales
 * the HBM byte counters. */
int sa_dlacpy(hipStream_t stream, int m, int n, const double *A, int lda, double *B, int ldb);
/* The same for FP32 (LAPACK slacpy, all of A). */
int sa_slacpy(hipStream_t stream, int m, int n, const float *A, int lda, float *B, int ldb);

#ifdef __cplusplus
}
#endif
#endif
