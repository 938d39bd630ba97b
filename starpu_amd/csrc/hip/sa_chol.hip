// Triangular FP64 kernels for the tiled Cholesky path on gfx950: POTRF of a
// diagonal tile and TRSM (right, lower, transposed, non-unit) of a panel
// tile, built so that almost all of their flops run on the matrix cores.
//
// k_potrf_inv64: one 4-wave workgroup factors a 64x64 diagonal block held
//   in LDS as 4x4 blocks of 16x16 (right-looking: 16x16 Cholesky in
//   registers, panel = A_IJ * inv(L_JJ)^T and trailing SYRK/GEMM as 16x16x16
//   MFMA products) and also produces inv(L) of the block (block forward
//   substitution, again 16x16x16 MFMA products).  Only the 16x16 pieces run
//   as short dependency chains on the VALU.
// k_apply_inv: P <- P * inv(L)^T for a panel below a 64x64 block (MFMA).
// k_trsm_rows16: B <- B * L^{-T} for n <= 1024, one workgroup per 16 rows;
//   the solved row stripe stays in LDS, each 64-column block is
//   acc = B - X * L^T (MFMA over the already solved columns) followed by
//   X = acc * inv(L_jj)^T (MFMA with the precomputed diagonal-block inverse).
//
// The diagonal-block inverses live in a caller-provided workspace: the C ABI
// never allocates.
#include "sa_block64.h"
#include "starpu_amd_kernels.h"

namespace {
using namespace sa64;

// FACTOR: Cholesky of the w x w (w <= 64) lower block D (in place) and its
// inverse into Dinv (64 x 64, leading dimension 64, identity past w); one
// workgroup, n = w.
// !FACTOR: the n x n lower L is already factored; workgroup b writes the
// inverse of L's 64x64 diagonal block b to Dinv + 4096 b.
template <bool FACTOR>
__global__ __launch_bounds__(256)
void k_potrf_inv64(int n, double *__restrict__ Dbase, long ldd, double *__restrict__ Dinvbase,
                   int *__restrict__ info, int col_base)
{
	__shared__ double S[64 * LD];
	__shared__ double Li[64 * LD];
	// workgroup b handles the diagonal block b (rows/cols 64b .. 64b+63)
	const int j0 = 64 * blockIdx.x;
	const int w = (n - j0) < 64 ? (n - j0) : 64;
	const int bad = block_potrf_inv64<FACTOR>(w, Dbase + j0 + (long)j0 * ldd, ldd, Dinvbase + 4096L * blockIdx.x,
	                                          S, Li);
	if (FACTOR && threadIdx.x == 0) {
		if (col_base == 0)
			*info = (bad && bad <= w) ? bad : 0;
		else if (bad && bad <= w)
			atomicCAS(info, 0, col_base + bad);
	}
}

// P (m x w, w <= 64) <- P * Dinv^T, Dinv 64x64 (ld 64, identity past w);
// one wave per 16 rows.
__global__ __launch_bounds__(256)
void k_apply_inv(int m, int w, const double *__restrict__ Dinv, double *__restrict__ P, long ldp)
{
	rows_apply_inv((blockIdx.x * 4 + (threadIdx.x >> 6)) * 16, m, w, Dinv, P, ldp);
}

// B (m x n, n <= 1024) <- alpha * B * L^{-T}; Dinv holds inv(L_jj) of the
// 64x64 diagonal blocks of L (block j at Dinv + 4096*j, ld 64).
// 8 waves: wave w computes output columns 16*(w%4) .. +15 of the current
// 64-column block over the solved 64-column chunks of parity w/4, so two
// waves share each SIMD and hide each other's L2 latency; the halves are
// summed through LDS.
//
// IDENT: B's input is the identity (B is only written), so X = alpha L^{-T}
// is upper triangular: the stripe's columns left of its 64-column block
// jb0 = row0 / 64 are zero (written as such, not computed), and the sums
// over solved columns start at block jb0.
template <int NMAX, bool IDENT = false>
__global__ __launch_bounds__(512)
void k_trsm_rows16(int m, int n, double alpha, const double *__restrict__ L, long ldl,
                   const double *__restrict__ Dinv, double *__restrict__ B, long ldb)
{
	__shared__ double Xs[NMAX * 16];   // solved stripe, column-major: Xs[c*16 + r]
	__shared__ double Ps[4 * 256];     // partial sums of the second K half
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	const int ob = wave & 3, half = wave >> 2;
	const int r = lane & 15, g = lane >> 4;
	const int row0 = blockIdx.x * 16;
	const int nblk = (n + 63) / 64;
	const int jb0 = IDENT ? row0 / 64 : 0;
	if (IDENT) {
		for (int idx = tid; idx < 16 * 64 * jb0; idx += 512) {
			const int rr = row0 + (idx & 15), cc = idx >> 4;
			if (rr < m && cc < n)
				B[rr + (long)cc * ldb] = 0.0;
		}
	}
	for (int jb = jb0; jb < nblk; jb++) {
		const int c0 = 64 * jb;
		const int col = c0 + 16 * ob + r;                // this lane's C column (as B operand row)
		const int colc = col < n ? col : n - 1;
		// acc = sum over solved 64-chunks k0 = 64*(2i + half) < c0 of
		// X(:, k0..k0+63) * L(col, k0..k0+63)^T; the next chunk's L values
		// are loaded (clamped past the end) while this one computes.
		sa_d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
		const double *Lrow = L + colc;
		double bc[16], bn[16];
		int k0 = 64 * jb0 + 64 * half;
		if (k0 < c0) {
#pragma unroll
			for (int i = 0; i < 16; i++)
				bc[i] = Lrow[(long)(k0 + 4 * i + g) * ldl];
			for (; k0 < c0; k0 += 128) {
#pragma unroll
				for (int i = 0; i < 16; i++) {
					const int kn = k0 + 128 + 4 * i + g;
					bn[i] = Lrow[(long)(kn < c0 ? kn : c0 - 1) * ldl];
				}
#pragma unroll
				for (int i = 0; i < 16; i += 2) {
					acc0 = sa_mfma(Xs[(k0 + 4 * i + g) * 16 + r], bc[i], acc0);
					acc1 = sa_mfma(Xs[(k0 + 4 * i + 4 + g) * 16 + r], bc[i + 1], acc1);
				}
#pragma unroll
				for (int i = 0; i < 16; i++)
					bc[i] = bn[i];
			}
		}
		acc0 += acc1;
		if (half) {
#pragma unroll
			for (int q = 0; q < 4; q++)
				Ps[ob * 256 + q * 64 + lane] = acc0[q];
		}
		__syncthreads();
		if (!half) {
#pragma unroll
			for (int q = 0; q < 4; q++)
				acc0[q] += Ps[ob * 256 + q * 64 + lane];
			// rhs block: T = alpha*B(rows, c0..c0+63) - acc, staged in Xs
#pragma unroll
			for (int q = 0; q < 4; q++) {
				const int rr = row0 + 4 * q + g;
				const int rrc = rr < m ? rr : m - 1;
				const double bv = IDENT ? (rr == col ? 1.0 : 0.0) : (col < n) ? B[rrc + (long)col * ldb] : 0.0;
				Xs[(c0 + 16 * ob + r) * 16 + 4 * q + g] = (col < n) ? alpha * bv - acc0[q] : 0.0;
			}
		}
		__syncthreads();
		// X(:, c0 + 16 ob + r) = sum_t T(:, t) * Dinv(16 ob + r, t)
		sa_d4 x = {0.0, 0.0, 0.0, 0.0};
		if (!half) {
			const double *Dj = Dinv + 4096L * jb;
#pragma unroll
			for (int s = 0; s < 16; s++)
				x = sa_mfma(Xs[(c0 + 4 * s + g) * 16 + r], Dj[(16 * ob + r) + (4 * s + g) * 64], x);
		}
		__syncthreads();
		if (!half) {
#pragma unroll
			for (int q = 0; q < 4; q++) {
				const int rr = row0 + 4 * q + g;
				Xs[(c0 + 16 * ob + r) * 16 + 4 * q + g] = x[q];
				if (rr < m && col < n)
					B[rr + (long)col * ldb] = x[q];
			}
		}
		__syncthreads();
	}
}

} // namespace

// Tiles up to 2048 are factored by the single-launch kernel of sa_potrf.hip;
// larger ones by one launch per 64-column step below.
extern "C" size_t sa_dpotrf_l_worksize(int n)
{
	const size_t step = 64 * 64 * sizeof(double);
	const size_t tile = n <= 2048 ? sa_dpotrf_tile_worksize(n) : 0;
	return tile > step ? tile : step;
}

extern "C" size_t sa_dtrsm_rltn_worksize(int n)
{
	const size_t nb = n > 0 ? (size_t)(n + 63) / 64 : 1;
	return nb * 64 * 64 * sizeof(double);
}

extern "C" int sa_dtrtri_diag64(hipStream_t s, int n, const double *L, int ldl, double *Dinv)
{
	if (n < 0 || ldl < (n > 1 ? n : 1) || (n > 0 && !Dinv))
		return -EINVAL;
	if (n == 0)
		return 0;
	hipLaunchKernelGGL(k_potrf_inv64<false>, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s, n,
	                   const_cast<double *>(L), (long)ldl, Dinv, (int *)nullptr, 0);
	return sa_launch_status();
}

// W = L^{-T} (n <= 1024): the diagonal-block inverses, then the stripe
// kernel on an identity right-hand side it generates itself.
int sa_trtri_lt_rows16(hipStream_t s, int n, const double *L, int ldl, double *W, int ldw, void *work,
                       size_t work_bytes)
{
	if (n <= 0 || n > 1024 || !work || work_bytes < sa_dtrsm_rltn_worksize(n))
		return -EINVAL;
	double *Dinv = (double *)work;
	int e = sa_dtrtri_diag64(s, n, L, ldl, Dinv);
	if (e)
		return e;
	hipLaunchKernelGGL((k_trsm_rows16<1024, true>), dim3((unsigned)((n + 15) / 16)), dim3(512), 0, s, n, n, 1.0, L,
	                   (long)ldl, (const double *)Dinv, W, (long)ldw);
	return sa_launch_status();
}

extern "C" int sa_dtrsm_rltn(hipStream_t s, int m, int n, double alpha, const double *L, int ldl,
                             double *B, int ldb, void *work, size_t work_bytes)
{
	if (m < 0 || n < 0 || ldl < (n > 1 ? n : 1) || ldb < (m > 1 ? m : 1))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	if (!work || work_bytes < sa_dtrsm_rltn_worksize(n < 1024 ? n : 1024))
		return -EINVAL;
	double *Dinv = (double *)work;
	// Column blocks of at most 1024 (the LDS stripe); the columns right of a
	// block are updated by a GEMM with the solved block.
	for (int j = 0; j < n; j += 1024) {
		const int w = (n - j) < 1024 ? (n - j) : 1024;
		int e = sa_dtrtri_diag64(s, w, L + j + (long)j * ldl, ldl, Dinv);
		if (e)
			return e;
		hipLaunchKernelGGL(k_trsm_rows16<1024>, dim3((unsigned)((m + 15) / 16)), dim3(512), 0, s, m, w,
		                   j == 0 ? alpha : 1.0, L + j + (long)j * ldl, (long)ldl, (const double *)Dinv,
		                   B + (long)j * ldb, (long)ldb);
		e = sa_launch_status();
		if (e)
			return e;
		if (j + w < n) {
			if (j == 0 && alpha != 1.0) {
				e = sa_scale_launch(s, m, n - w, B + (long)w * ldb, ldb, alpha, false);
				if (e)
					return e;
			}
			e = sa_gemm_launch(s, false, m, n - j - w, w, -1.0, B + (long)j * ldb, ldb,
			                   L + (j + w) + (long)j * ldl, ldl, 1.0, B + (long)(j + w) * ldb, ldb);
			if (e)
				return e;
		}
	}
	return 0;
}

extern "C" int sa_dpotrf_l(hipStream_t s, int n, double *A, int lda, int *dev_info, void *work,
                           size_t work_bytes)
{
	if (n < 0 || lda < (n > 1 ? n : 1) || dev_info == nullptr || !work ||
	    work_bytes < sa_dpotrf_l_worksize(n))
		return -EINVAL;
	if (n <= 2048)
		return sa_dpotrf_tile(s, n, A, lda, nullptr, 0, nullptr, 0, dev_info, work, work_bytes);
	double *Dinv = (double *)work;
	if (n == 0) {
		hipLaunchKernelGGL(k_potrf_inv64<true>, dim3(1), dim3(256), 0, s, 0, A, (long)lda, Dinv,
		                   dev_info, 0);   // reports info = 0
		return sa_launch_status();
	}
	for (int j = 0; j < n; j += 64) {
		const int w = (n - j) < 64 ? (n - j) : 64;
		const int rows = n - j - w;
		double *Djj = A + j + (long)j * lda;
		hipLaunchKernelGGL(k_potrf_inv64<true>, dim3(1), dim3(256), 0, s, w, Djj, (long)lda, Dinv,
		                   dev_info, j);
		int e = sa_launch_status();
		if (e)
			return e;
		if (rows > 0) {
			hipLaunchKernelGGL(k_apply_inv, dim3((unsigned)((rows + 63) / 64)), dim3(256), 0, s, rows, w,
			                   (const double *)Dinv, Djj + w, (long)lda);
			e = sa_launch_status();
			if (e)
				return e;
			e = sa_gemm_launch(s, true, rows, rows, w, -1.0, Djj + w, lda, Djj + w, lda, 1.0,
			                   A + (j + w) + (long)(j + w) * lda, lda);
			if (e)
				return e;
		}
	}
	return 0;
}
