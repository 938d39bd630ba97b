// Kernels of the Cholesky solve phase (L y = b, L^T x = y) on gfx950, for
// a thin block of right-hand sides (a few to a few tens of columns).  Every
// L tile is read once per solve, so these kernels are bound by HBM and L2
// bandwidth, not by the matrix cores: they run on the VALU and keep the
// right-hand sides in registers and LDS.
//
// k_trsv_blocked<UPPER>: X <- alpha * T^{-1} X for one triangular tile T
//   (m <= 1024), 8 waves per group of 8 right-hand sides.  The tile is
//   walked in 64-row blocks (forward for lower T, backward for upper T);
//   for each block the waves first sum T(block rows, solved columns) *
//   Y(solved) over interleaved columns (coalesced 512-byte column reads),
//   then one wave per right-hand side finishes the block by substitution
//   inside the 64x64 diagonal block, lane = row.  Solved rows stay in LDS.
// k_transpose_lower: U = L^T for the backward solve, so that it reads U
//   column by column like the forward solve reads L.
// k_gemm_thin_nn / k_gemm_thin_tn: C(m x r) = beta*C + alpha*op(A)*B with
//   r small: NN has one lane per row of A and waves splitting k; TN one
//   wave per 4 columns of A with lanes along k and a cross-lane reduction.
#include "sa_common.h"
#include "starpu_amd_kernels.h"

namespace {

constexpr int RG = 8;             // right-hand sides per workgroup
constexpr int TRSV_MAX = 1024;    // rows of the triangular tile one launch solves

__device__ __forceinline__ double readlane_d(double v, int lane)
{
	const long long bits = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), lane);
	const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
	return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <bool UPPER>
__global__ __launch_bounds__(512)
void k_trsv_blocked(int m, int r, double alpha, const double *__restrict__ T, long ldt,
                    double *__restrict__ X, long ldx)
{
	__shared__ double Y[TRSV_MAX * RG];     // solved rows, Y[row * RG + c]
	__shared__ double P[8 * 64 * RG];       // per-wave partial sums
	__shared__ double D[64 * 65];           // diagonal block, ld 65
	__shared__ double Rd[64];               // reciprocals of its diagonal
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int c0 = blockIdx.x * RG;
	const int rg = (r - c0) < RG ? (r - c0) : RG;
	const int nblk = (m + 63) / 64;
	for (int s = 0; s < nblk; s++) {
		const int jb = UPPER ? nblk - 1 - s : s;
		const int j0 = 64 * jb;
		const int w = (m - j0) < 64 ? (m - j0) : 64;
		const int row = j0 + lane;
		const bool rin = lane < w;
		const int rowc = rin ? row : m - 1;
		for (int idx = tid; idx < 64 * 64; idx += 512) {
			const int i = idx & 63, j = idx >> 6;
			D[i + 65 * j] = (i < w && j < w) ? T[(j0 + i) + (long)(j0 + j) * ldt] : 0.0;
		}
		if (tid < 64)
			Rd[tid] = tid < w ? 1.0 / T[(j0 + tid) + (long)(j0 + tid) * ldt] : 0.0;
		// partial sums over the already solved rows: columns [0, j0) of a
		// lower T, [j0 + w, m) of an upper T; wave q takes columns q, q+8, ..
		double acc[RG];
#pragma unroll
		for (int c = 0; c < RG; c++)
			acc[c] = 0.0;
		const int kb = UPPER ? j0 + w : 0, ke = UPPER ? m : j0;
		int k = kb + wave;
		for (; k + 24 < ke; k += 32) {
			double a[4];
#pragma unroll
			for (int u = 0; u < 4; u++)
				a[u] = T[rowc + (long)(k + 8 * u) * ldt];
#pragma unroll
			for (int u = 0; u < 4; u++)
#pragma unroll
				for (int c = 0; c < RG; c++)
					acc[c] += a[u] * Y[(k + 8 * u) * RG + c];
		}
		for (; k < ke; k += 8) {
			const double a = T[rowc + (long)k * ldt];
#pragma unroll
			for (int c = 0; c < RG; c++)
				acc[c] += a * Y[k * RG + c];
		}
#pragma unroll
		for (int c = 0; c < RG; c++)
			P[(wave * 64 + lane) * RG + c] = acc[c];
		__syncthreads();
		if (wave < rg) {
			const int c = wave;
			double t = 0.0;
			if (rin) {
				t = alpha * X[row + (long)(c0 + c) * ldx];
#pragma unroll
				for (int q = 0; q < 8; q++)
					t -= P[(q * 64 + lane) * RG + c];
			}
			// substitution inside the diagonal block; lane i holds row i
			if (!UPPER) {
				for (int i = 0; i < w; i++) {
					const double yi = readlane_d(t, i) * Rd[i];
					const double l = D[lane + 65 * i];
					t = (lane == i) ? yi : ((lane > i) ? t - l * yi : t);
				}
			} else {
				for (int i = w - 1; i >= 0; i--) {
					const double yi = readlane_d(t, i) * Rd[i];
					const double u = D[lane + 65 * i];
					t = (lane == i) ? yi : ((lane < i) ? t - u * yi : t);
				}
			}
			if (rin) {
				Y[row * RG + c] = t;
				X[row + (long)(c0 + c) * ldx] = t;
			}
		} else if (wave < RG) {
			if (rin)
				Y[row * RG + wave] = 0.0;   // padding right-hand sides
		}
		__syncthreads();
	}
}

// U(j, i) = L(i, j) for i >= j (n x n, 64x64 blocks through LDS; blocks
// strictly above L's diagonal are skipped).
__global__ __launch_bounds__(256)
void k_transpose_lower(int n, const double *__restrict__ L, long ldl, double *__restrict__ U, long ldu)
{
	__shared__ double S[64 * 65];
	const int bi = blockIdx.x, bj = blockIdx.y;      // block of L: rows bi, cols bj
	if (bi < bj)
		return;
	const int tid = threadIdx.x;
	for (int idx = tid; idx < 64 * 64; idx += 256) {
		const int i = idx & 63, j = idx >> 6;
		const int gi = 64 * bi + i, gj = 64 * bj + j;
		S[i + 65 * j] = (gi < n && gj < n) ? L[gi + (long)gj * ldl] : 0.0;
	}
	__syncthreads();
	for (int idx = tid; idx < 64 * 64; idx += 256) {
		const int i = idx & 63, j = idx >> 6;       // U block (bj, bi): row i, col j
		const int gi = 64 * bj + i, gj = 64 * bi + j;
		if (gi < n && gj < n)
			U[gi + (long)gj * ldu] = S[j + 65 * i];
	}
}

// C(m x r) = beta*C + alpha * A(m x k) * B(k x r): lane = row of A, the 4
// waves split k, B is read with wave-uniform (scalar) loads.
__global__ __launch_bounds__(256)
void k_gemm_thin_nn(int m, int r, int kdim, double alpha, const double *__restrict__ A, long lda,
                    const double *__restrict__ B, long ldb, double beta, double *__restrict__ C, long ldc)
{
	__shared__ double P[4 * 64 * RG];
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int row = blockIdx.x * 64 + lane;
	const int c0 = blockIdx.y * RG;
	const int rg = (r - c0) < RG ? (r - c0) : RG;
	const bool rin = row < m;
	const int rowc = rin ? row : m - 1;
	double acc[RG];
#pragma unroll
	for (int c = 0; c < RG; c++)
		acc[c] = 0.0;
	const double *Bc[RG];
#pragma unroll
	for (int c = 0; c < RG; c++)
		Bc[c] = B + (long)(c0 + (c < rg ? c : rg - 1)) * ldb;
	int k = wave;
	for (; k + 12 < kdim; k += 16) {
		double a[4];
#pragma unroll
		for (int u = 0; u < 4; u++)
			a[u] = A[rowc + (long)(k + 4 * u) * lda];
#pragma unroll
		for (int u = 0; u < 4; u++)
#pragma unroll
			for (int c = 0; c < RG; c++)
				acc[c] += a[u] * Bc[c][k + 4 * u];
	}
	for (; k < kdim; k += 4) {
		const double a = A[rowc + (long)k * lda];
#pragma unroll
		for (int c = 0; c < RG; c++)
			acc[c] += a * Bc[c][k];
	}
#pragma unroll
	for (int c = 0; c < RG; c++)
		P[(wave * 64 + lane) * RG + c] = acc[c];
	__syncthreads();
	// wave q writes right-hand sides q and q + 4
	for (int c = wave; c < rg; c += 4) {
		if (!rin)
			continue;
		double s = 0.0;
#pragma unroll
		for (int q = 0; q < 4; q++)
			s += P[(q * 64 + lane) * RG + c];
		double *p = C + row + (long)(c0 + c) * ldc;
		*p = (beta == 0.0 ? 0.0 : beta * *p) + alpha * s;
	}
}

// C(m x r) = beta*C + alpha * A^T * B, A (k x m), B (k x r): wave = 4
// columns of A (4 rows of C), lanes along k, then a butterfly reduction.
__global__ __launch_bounds__(256)
void k_gemm_thin_tn(int m, int r, int kdim, double alpha, const double *__restrict__ A, long lda,
                    const double *__restrict__ B, long ldb, double beta, double *__restrict__ C, long ldc)
{
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int i0 = (blockIdx.x * 4 + wave) * 4;
	const int c0 = blockIdx.y * RG;
	const int rg = (r - c0) < RG ? (r - c0) : RG;
	if (i0 >= m)
		return;
	double acc[4][RG];
#pragma unroll
	for (int q = 0; q < 4; q++)
#pragma unroll
		for (int c = 0; c < RG; c++)
			acc[q][c] = 0.0;
	const double *Ac[4];
#pragma unroll
	for (int q = 0; q < 4; q++)
		Ac[q] = A + (long)(i0 + q < m ? i0 + q : m - 1) * lda;
	const double *Bc[RG];
#pragma unroll
	for (int c = 0; c < RG; c++)
		Bc[c] = B + (long)(c0 + (c < rg ? c : rg - 1)) * ldb;
	for (int k = lane; k < kdim; k += 64) {
		double a[4], b[RG];
#pragma unroll
		for (int q = 0; q < 4; q++)
			a[q] = Ac[q][k];
#pragma unroll
		for (int c = 0; c < RG; c++)
			b[c] = Bc[c][k];
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int c = 0; c < RG; c++)
				acc[q][c] += a[q] * b[c];
	}
#pragma unroll
	for (int q = 0; q < 4; q++)
#pragma unroll
		for (int c = 0; c < RG; c++) {
			double v = acc[q][c];
#pragma unroll
			for (int off = 32; off >= 1; off >>= 1)
				v += __shfl_xor(v, off);
			acc[q][c] = v;
		}
	// lane q*RG + c writes C(i0 + q, c0 + c)
	const int q = lane / RG, c = lane % RG;
	if (lane < 4 * RG && i0 + q < m && c < rg) {
		double v = 0.0;
#pragma unroll
		for (int qq = 0; qq < 4; qq++)
#pragma unroll
			for (int cc = 0; cc < RG; cc++)
				if (qq == q && cc == c)
					v = acc[qq][cc];
		double *p = C + (i0 + q) + (long)(c0 + c) * ldc;
		*p = (beta == 0.0 ? 0.0 : beta * *p) + alpha * v;
	}
}

int gemm_thin_nn(hipStream_t s, int m, int r, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc)
{
	if (m == 0 || r == 0)
		return 0;
	if (k == 0)
		return beta == 1.0 ? 0 : sa_scale_launch(s, m, r, C, ldc, beta, false);
	hipLaunchKernelGGL(k_gemm_thin_nn, dim3((unsigned)((m + 63) / 64), (unsigned)((r + RG - 1) / RG)),
	                   dim3(256), 0, s, m, r, k, alpha, A, (long)lda, B, (long)ldb, beta, C, (long)ldc);
	return sa_launch_status();
}

int gemm_thin_tn(hipStream_t s, int m, int r, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc)
{
	if (m == 0 || r == 0)
		return 0;
	if (k == 0)
		return beta == 1.0 ? 0 : sa_scale_launch(s, m, r, C, ldc, beta, false);
	hipLaunchKernelGGL(k_gemm_thin_tn, dim3((unsigned)((m + 15) / 16), (unsigned)((r + RG - 1) / RG)),
	                   dim3(256), 0, s, m, r, k, alpha, A, (long)lda, B, (long)ldb, beta, C, (long)ldc);
	return sa_launch_status();
}

bool ld_ok(int ld, int rows) { return ld >= (rows > 1 ? rows : 1); }

} // namespace

extern "C" int sa_dgemm_nn(hipStream_t s, int m, int n, int k, double alpha, const double *A, int lda,
                           const double *B, int ldb, double beta, double *C, int ldc)
{
	if (m < 0 || n < 0 || k < 0 || !ld_ok(lda, m) || !ld_ok(ldb, k) || !ld_ok(ldc, m))
		return -EINVAL;
	return gemm_thin_nn(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int sa_dgemm_tn(hipStream_t s, int m, int n, int k, double alpha, const double *A, int lda,
                           const double *B, int ldb, double beta, double *C, int ldc)
{
	if (m < 0 || n < 0 || k < 0 || !ld_ok(lda, k) || !ld_ok(ldb, k) || !ld_ok(ldc, m))
		return -EINVAL;
	return gemm_thin_tn(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int sa_dtrsm_llnn(hipStream_t s, int m, int n, double alpha, const double *L, int ldl,
                             double *B, int ldb)
{
	if (m < 0 || n < 0 || !ld_ok(ldl, m) || !ld_ok(ldb, m))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	const unsigned grid = (unsigned)((n + RG - 1) / RG);
	for (int j = 0; j < m; j += TRSV_MAX) {
		const int w = (m - j) < TRSV_MAX ? (m - j) : TRSV_MAX;
		const double a = (j == 0) ? alpha : 1.0;
		hipLaunchKernelGGL(k_trsv_blocked<false>, dim3(grid), dim3(512), 0, s, w, n, a,
		                   L + j + (long)j * ldl, (long)ldl, B + j, (long)ldb);
		int e = sa_launch_status();
		if (e)
			return e;
		if (j + w < m) {
			// B2 <- a*B2 - L21 * X1, with a = alpha only before the first block
			e = gemm_thin_nn(s, m - j - w, n, w, -1.0, L + (j + w) + (long)j * ldl, ldl, B + j, ldb,
			                 j == 0 ? alpha : 1.0, B + j + w, ldb);
			if (e)
				return e;
		}
	}
	return 0;
}

extern "C" size_t sa_dtrsm_lltn_worksize(int m)
{
	const size_t w = (size_t)(m < TRSV_MAX ? (m > 0 ? m : 0) : TRSV_MAX);
	return w * w * sizeof(double);
}

extern "C" int sa_dtrsm_lltn(hipStream_t s, int m, int n, double alpha, const double *L, int ldl,
                             double *B, int ldb, void *work, size_t work_bytes)
{
	if (m < 0 || n < 0 || !ld_ok(ldl, m) || !ld_ok(ldb, m))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	if (!work || work_bytes < sa_dtrsm_lltn_worksize(m))
		return -EINVAL;
	double *U = (double *)work;
	const unsigned grid = (unsigned)((n + RG - 1) / RG);
	// chunks of TRSV_MAX rows from the bottom up: the last chunk of X first
	const int nch = (m + TRSV_MAX - 1) / TRSV_MAX;
	for (int c = nch - 1; c >= 0; c--) {
		const int j = c * TRSV_MAX;
		const int w = (m - j) < TRSV_MAX ? (m - j) : TRSV_MAX;
		const double a = (c == nch - 1) ? alpha : 1.0;
		const unsigned nb64 = (unsigned)((w + 63) / 64);
		hipLaunchKernelGGL(k_transpose_lower, dim3(nb64, nb64), dim3(256), 0, s, w,
		                   L + j + (long)j * ldl, (long)ldl, U, (long)w);
		int e = sa_launch_status();
		if (e)
			return e;
		hipLaunchKernelGGL(k_trsv_blocked<true>, dim3(grid), dim3(512), 0, s, w, n, a, U, (long)w,
		                   B + j, (long)ldb);
		e = sa_launch_status();
		if (e)
			return e;
		if (j > 0) {
			// X0 <- a0*X0 - L10^T * X1 (L10 = L(j:j+w, 0:j))
			e = gemm_thin_tn(s, j, n, w, -1.0, L + j, ldl, B + j, ldb, c == nch - 1 ? alpha : 1.0, B, ldb);
			if (e)
				return e;
		}
	}
	return 0;
}
