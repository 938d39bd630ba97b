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
#include "sa_common.h"
#include "starpu_amd_kernels.h"

namespace {

constexpr int LD = 65;      // LDS leading dimension of a 64x64 block (bank spread)

// v of lane K of each 16-lane row, broadcast to the whole row with the DPP
// row_newbcast mode (a VALU move: no SGPR round trip as with v_readlane).
template <int K>
__device__ __forceinline__ double rb(double v)
{
	const long long bits = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), 0x150 + K, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), 0x150 + K, 0xf, 0xf, false);
	return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v of lane `lane` (< 16) broadcast to the whole wave through an SGPR pair.
__device__ __forceinline__ double bcast_rl(double v, int lane)
{
	const long long bits = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), lane);
	const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
	return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// k is a compile-time constant wherever this is called from fully unrolled
// loops, so the switch folds to one broadcast.
__device__ __forceinline__ double bcast(double v, int k)
{
	switch (k) {
	case 0: return rb<0>(v);
	case 1: return rb<1>(v);
	case 2: return rb<2>(v);
	case 3: return rb<3>(v);
	case 4: return rb<4>(v);
	case 5: return rb<5>(v);
	case 6: return rb<6>(v);
	case 7: return rb<7>(v);
	case 8: return rb<8>(v);
	case 9: return rb<9>(v);
	case 10: return rb<10>(v);
	case 11: return rb<11>(v);
	case 12: return rb<12>(v);
	case 13: return rb<13>(v);
	case 14: return rb<14>(v);
	default: return rb<15>(v);
	}
}

// 16x16 block products on LDS operands (block origins, leading dimension LD).
// The result is in MFMA output layout: register q of lane l holds element
// (4q + l/16, l%16).
__device__ __forceinline__ sa_d4 mm16_nt(const double *A, const double *B, sa_d4 acc, double sign)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
	for (int s = 0; s < 4; s++)
		acc = sa_mfma(sign * A[r + (4 * s + g) * LD], B[r + (4 * s + g) * LD], acc);
	return acc;
}

__device__ __forceinline__ sa_d4 mm16_nn(const double *A, const double *B, sa_d4 acc)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
	for (int s = 0; s < 4; s++)
		acc = sa_mfma(A[r + (4 * s + g) * LD], B[(4 * s + g) + r * LD], acc);
	return acc;
}

// acc += A * T with T a 16x16 block already in MFMA output layout: register
// s of lane l holds T(4s + l/16, l%16), which is exactly the B operand of
// k-step s.
__device__ __forceinline__ sa_d4 mm16_nreg(const double *A, sa_d4 T, sa_d4 acc)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
	for (int s = 0; s < 4; s++)
		acc = sa_mfma(A[r + (4 * s + g) * LD], T[s], acc);
	return acc;
}

__device__ __forceinline__ void st16(double *P, sa_d4 acc, double sign)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
	for (int q = 0; q < 4; q++)
		P[(4 * q + g) + r * LD] = sign * acc[q];
}

__device__ __forceinline__ sa_d4 ld16(const double *P)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
	sa_d4 acc;
#pragma unroll
	for (int q = 0; q < 4; q++)
		acc[q] = P[(4 * q + g) + r * LD];
	return acc;
}

// Lane t < 16 holds row t of a 16x16 lower block in x[0..15] (columns
// above the diagonal are ignored).  In-place Cholesky; returns the 1-based
// index of the first non-positive pivot, or 0.  The pivots and multipliers
// are broadcast with v_readlane: the row-broadcast DPP form gave a wrong
// factor here (a 17x17 input) although the same loop in isolation was
// correct, so this chain stays on the SGPR path.
__device__ __forceinline__ int potrf16_regs(double (&x)[16])
{
	const int t = threadIdx.x & 63;
	int bad = 0;
#pragma unroll
	for (int k = 0; k < 16; k++) {
		const double d = bcast_rl(x[k], k);
		if (!(d > 0.0) && bad == 0)
			bad = k + 1;
		const double sd = sqrt(d);
		x[k] = (t == k) ? sd : x[k] / sd;
#pragma unroll
		for (int j = k + 1; j < 16; j++)
			x[j] -= x[k] * bcast_rl(x[k], j);
	}
	return bad;
}

// Lane t < 16 holds row t of a factored 16x16 lower L; lane c gets column c
// of inv(L) in y: y_i = (delta_ic - sum_{q<i} L(i,q) y_q) / L(i,i).
__device__ __forceinline__ void inv16_regs(const double (&x)[16], double (&y)[16])
{
	const int c = threadIdx.x & 63;
#pragma unroll
	for (int i = 0; i < 16; i++) {
		double s = (c == i) ? 1.0 : 0.0;
#pragma unroll
		for (int q = 0; q < i; q++)
			s -= bcast(x[q], i) * y[q];
		y[i] = s / bcast(x[i], i);
	}
}

// Rows of the 16x16 diagonal block J of S into lanes 0..15 (lower part);
// other lanes get an identity row so that they run the same code.
__device__ __forceinline__ void load_rows16(const double *S, int J, double (&x)[16])
{
	const int t = threadIdx.x & 63;
#pragma unroll
	for (int c = 0; c < 16; c++)
		x[c] = (t < 16) ? ((c <= t) ? S[(16 * J + t) + (16 * J + c) * LD] : 0.0)
		                : ((c == (t & 15)) ? 1.0 : 0.0);
}

__device__ __forceinline__ void store_inv16(double *Li, int J, const double (&y)[16])
{
	const int c = threadIdx.x & 63;
	if (c < 16)
#pragma unroll
		for (int i = 0; i < 16; i++)
			Li[(16 * J + i) + (16 * J + c) * LD] = y[i];
}

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
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	// workgroup b handles the diagonal block b (rows/cols 64b .. 64b+63)
	const int j0 = 64 * blockIdx.x;
	const int w = (n - j0) < 64 ? (n - j0) : 64;
	double *D = Dbase + j0 + (long)j0 * ldd;
	double *Dinv = Dinvbase + 4096L * blockIdx.x;
	for (int idx = tid; idx < 64 * 64; idx += 256) {
		const int i = idx & 63, j = idx >> 6;
		double v;
		if (i < w && j < w)
			v = (i >= j) ? D[i + (long)j * ldd] : 0.0;
		else
			v = (i == j) ? 1.0 : 0.0;
		S[i + j * LD] = v;
		Li[i + j * LD] = 0.0;
	}
	__syncthreads();

	int bad = 0;
	if (FACTOR) {
		for (int J = 0; J < 4; J++) {
			if (wave == 0) {
				double x[16], y[16];
				load_rows16(S, J, x);
				const int b = potrf16_regs(x);
				if (b && !bad)
					bad = 16 * J + b;
				if (lane < 16)
#pragma unroll
					for (int c = 0; c < 16; c++)
						if (c <= lane)
							S[(16 * J + lane) + (16 * J + c) * LD] = x[c];
				inv16_regs(x, y);
				store_inv16(Li, J, y);
			}
			__syncthreads();
			// panel: L_IJ = A_IJ * inv(L_JJ)^T
			if (wave < 3 - J) {
				const int I = J + 1 + wave;
				const sa_d4 z = {0.0, 0.0, 0.0, 0.0};
				const sa_d4 acc = mm16_nt(&S[16 * I + 16 * J * LD], &Li[16 * J + 16 * J * LD], z, 1.0);
				st16(&S[16 * I + 16 * J * LD], acc, 1.0);
			}
			__syncthreads();
			// trailing: A_IK -= L_IJ * L_KJ^T for J < K <= I < 4
			int p = 0;
			for (int I = J + 1; I < 4; I++)
				for (int K = J + 1; K <= I; K++, p++) {
					if ((p & 3) != wave)
						continue;
					sa_d4 acc = ld16(&S[16 * I + 16 * K * LD]);
					acc = mm16_nt(&S[16 * I + 16 * J * LD], &S[16 * K + 16 * J * LD], acc, -1.0);
					st16(&S[16 * I + 16 * K * LD], acc, 1.0);
				}
			__syncthreads();
		}
	} else {
		double x[16], y[16];
		load_rows16(S, wave, x);
		inv16_regs(x, y);
		store_inv16(Li, wave, y);
		__syncthreads();
	}
	// off-diagonal blocks of inv(L): Li_IJ = -Li_II * sum_{K=J}^{I-1} L_IK Li_KJ
	for (int d = 1; d < 4; d++) {
		if (wave < 4 - d) {
			const int J = wave, I = J + d;
			sa_d4 T = {0.0, 0.0, 0.0, 0.0};
			for (int K = J; K < I; K++)
				T = mm16_nn(&S[16 * I + 16 * K * LD], &Li[16 * K + 16 * J * LD], T);
			const sa_d4 z = {0.0, 0.0, 0.0, 0.0};
			const sa_d4 X = mm16_nreg(&Li[16 * I + 16 * I * LD], T, z);
			st16(&Li[16 * I + 16 * J * LD], X, -1.0);
		}
		__syncthreads();
	}
	if (FACTOR)
		for (int idx = tid; idx < 64 * 64; idx += 256) {
			const int i = idx & 63, j = idx >> 6;
			if (i < w && j < w && i >= j)
				D[i + (long)j * ldd] = S[i + j * LD];
		}
	for (int idx = tid; idx < 64 * 64; idx += 256) {
		const int i = idx & 63, j = idx >> 6;
		Dinv[i + j * 64] = Li[i + j * LD];
	}
	if (FACTOR && tid == 0) {
		if (col_base == 0)
			*info = (bad && bad <= w) ? bad : 0;
		else if (bad && bad <= w)
			atomicCAS(info, 0, col_base + bad);
	}
}

// P (m x w, w <= 64) <- P * Dinv^T, Dinv 64x64 (ld 64, identity past w).
// One wave per 16 rows reads its rows completely before writing them, so
// the update is in place.
__global__ __launch_bounds__(256)
void k_apply_inv(int m, int w, const double *__restrict__ Dinv, double *__restrict__ P, long ldp)
{
	const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
	const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
	if (row0 >= m)
		return;
	const int row = row0 + r;
	const int rc = row < m ? row : m - 1;
	double a[16];
#pragma unroll
	for (int s = 0; s < 16; s++) {
		const int k = 4 * s + g;
		const double v = P[rc + (long)(k < w ? k : w - 1) * ldp];
		a[s] = k < w ? v : 0.0;
	}
	sa_d4 acc[4];
#pragma unroll
	for (int cb = 0; cb < 4; cb++) {
		acc[cb] = sa_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
		for (int s = 0; s < 16; s++)
			acc[cb] = sa_mfma(a[s], Dinv[(16 * cb + r) + (4 * s + g) * 64], acc[cb]);
	}
	// acc[cb][q] = X(row0 + 4q + g, 16cb + r)
#pragma unroll
	for (int cb = 0; cb < 4; cb++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int rr = row0 + 4 * q + g, cc = 16 * cb + r;
			if (rr < m && cc < w)
				P[rr + (long)cc * ldp] = acc[cb][q];
		}
}

// B (m x n, n <= 1024) <- alpha * B * L^{-T}; Dinv holds inv(L_jj) of the
// 64x64 diagonal blocks of L (block j at Dinv + 4096*j, ld 64).
// 8 waves: wave w computes output columns 16*(w%4) .. +15 of the current
// 64-column block over the solved 64-column chunks of parity w/4, so two
// waves share each SIMD and hide each other's L2 latency; the halves are
// summed through LDS.
template <int NMAX>
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
	for (int jb = 0; jb < nblk; jb++) {
		const int c0 = 64 * jb;
		const int col = c0 + 16 * ob + r;                // this lane's C column (as B operand row)
		const int colc = col < n ? col : n - 1;
		// acc = sum over solved 64-chunks k0 = 64*(2i + half) < c0 of
		// X(:, k0..k0+63) * L(col, k0..k0+63)^T; the next chunk's L values
		// are loaded (clamped past the end) while this one computes.
		sa_d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
		const double *Lrow = L + colc;
		double bc[16], bn[16];
		int k0 = 64 * half;
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
				const double bv = (col < n) ? B[rrc + (long)col * ldb] : 0.0;
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

extern "C" size_t sa_dpotrf_l_worksize(int n)
{
	(void)n;
	return 64 * 64 * sizeof(double);
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
