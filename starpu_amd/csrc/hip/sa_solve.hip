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
// k_gemm_thin<TRANS>: C(m x r) = beta*C + alpha*op(A)*B with r small, split
//   over k across workgroups so that many loads are in flight; partial sums
//   are reduced in a fixed order (k_thin_reduce).
// sa_dtrtri_lt / sa_dtrmm_lu: the inverse-based tile solve.  W = L^{-T} of a
//   diagonal tile is computed once (the right-side MFMA TRSM applied to an
//   identity), after which L^{-1} b = W^T b and L^{-T} b = W b are thin
//   GEMMs that run across the whole chip instead of a sequential
//   substitution.
#include <cstdlib>

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

constexpr int TR = 16;   // right-hand sides per workgroup of the thin GEMM

// k-columns staged per step: 64 for an FP64 A, 128 for an FP32 A, whose
// loads carry half the bytes, so that as many bytes are in flight.
template <typename TA> constexpr int thin_tk() { return sizeof(TA) == 4 ? 128 : 64; }

// Thin GEMM, split over k.  Workgroup (x, y, z) takes rows 64x.., right-hand
// sides 16z.. and k in [y*kc, (y+1)*kc).  NN (TRANS false): lane = row of A,
// wave w loads the 16 columns w, w+4, .. of each 64-column step at once
// (16 coalesced 512-byte loads in flight per wave).  TN (TRANS true, A is
// k x m): the 64 x 64 block of A is staged transposed through LDS with
// loads running down A's columns, then read with lane = output row.  B's
// 64 x 16 block is staged in LDS and read as broadcasts.  The four waves'
// sums are added in a fixed order; with one k slice the result goes to C,
// otherwise to P[(y*r + c)*m + row] for k_thin_reduce, so the result never
// depends on timing.
template <bool TRANS, typename TA>
__global__ __launch_bounds__(256)
void k_gemm_thin(int m, int r, int kdim, int kc, double alpha, const TA *__restrict__ A, long lda,
                 const double *__restrict__ B, long ldb, double beta, double *__restrict__ C, long ldc,
                 double *__restrict__ P)
{
	constexpr int TK = thin_tk<TA>();
	// with an FP32 A the final reduction reuses the staging block of A (free
	// after the last step), which keeps two workgroups per CU
	constexpr bool RED_IN_AS = TRANS && sizeof(TA) == 4;
	__shared__ __align__(16) TA As[TRANS ? TK * 65 : 1];
	__shared__ double Bs[TK * TR];
	__shared__ double Red_[RED_IN_AS ? 1 : 4 * 64 * TR];
	static_assert(!RED_IN_AS || sizeof(As) >= 4 * 64 * TR * sizeof(double), "reduction does not fit");
	double *Red = RED_IN_AS ? reinterpret_cast<double *>(As) : Red_;
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int row0 = blockIdx.x * 64;
	const int kb = blockIdx.y * kc;
	const int ke = (kdim - kb) < kc ? kdim : kb + kc;
	const int c0 = blockIdx.z * TR;
	const int rg = (r - c0) < TR ? (r - c0) : TR;
	const int row = row0 + lane;
	const int rowc = row < m ? row : m - 1;
	double acc[TR];
#pragma unroll
	for (int c = 0; c < TR; c++)
		acc[c] = 0.0;
	// The next step's B block (and, for TN, A block) is loaded into registers
	// while the current step computes, so a workgroup keeps its loads in
	// flight across the LDS round trip.
	constexpr int NB = TK * TR / 256, NA = TRANS ? TK * 64 / 256 : 1;
	double bn[NB];
	TA an[NA];
	auto fetch = [&](int k0) {
		const int kw = (ke - k0) < TK ? (ke - k0) : TK;
#pragma unroll
		for (int q = 0; q < NB; q++) {
			const int idx = tid + 256 * q, kk = idx % TK, c = idx / TK;
			const bool in = kk < kw && c < rg;
			const double v = B[(k0 + (kk < kw ? kk : kw - 1)) + (long)(c0 + (c < rg ? c : rg - 1)) * ldb];
			bn[q] = in ? v : 0.0;
		}
		if (TRANS) {
#pragma unroll
			for (int q = 0; q < NA; q++) {
				const int idx = tid + 256 * q, kk = idx % TK, i = idx / TK;
				const int gi = row0 + i;
				const bool in = kk < kw && gi < m;
				const TA v = A[(k0 + (kk < kw ? kk : kw - 1)) + (long)(gi < m ? gi : m - 1) * lda];
				an[q] = in ? v : TA(0);
			}
		}
	};
	if (kb < ke)
		fetch(kb);
	for (int k0 = kb; k0 < ke; k0 += TK) {
		const int kw = (ke - k0) < TK ? (ke - k0) : TK;
#pragma unroll
		for (int q = 0; q < NB; q++) {
			const int idx = tid + 256 * q;
			Bs[(idx % TK) * TR + idx / TK] = bn[q];
		}
		if (TRANS) {
#pragma unroll
			for (int q = 0; q < NA; q++) {
				const int idx = tid + 256 * q;
				As[(idx % TK) * 65 + idx / TK] = an[q];
			}
		}
		__syncthreads();
		if (k0 + TK < ke)
			fetch(k0 + TK);
		if (!TRANS) {
			double a[TK / 4];
#pragma unroll
			for (int u = 0; u < TK / 4; u++) {
				const int kk = wave + 4 * u;
				const double v = (double)A[rowc + (long)(k0 + (kk < kw ? kk : kw - 1)) * lda];
				a[u] = kk < kw ? v : 0.0;
			}
#pragma unroll
			for (int u = 0; u < TK / 4; u++)
#pragma unroll
				for (int c = 0; c < TR; c++)
					acc[c] += a[u] * Bs[(wave + 4 * u) * TR + c];
		} else {
#pragma unroll
			for (int u = 0; u < TK / 4; u++) {
				const int kk = wave + 4 * u;
				const double a = (double)As[kk * 65 + lane];
#pragma unroll
				for (int c = 0; c < TR; c++)
					acc[c] += a * Bs[kk * TR + c];
			}
		}
		__syncthreads();
	}
#pragma unroll
	for (int c = 0; c < TR; c++)
		Red[(wave * 64 + lane) * TR + c] = acc[c];
	__syncthreads();
	const int i = tid & 63, orow = row0 + i;
	for (int c = tid >> 6; c < rg; c += 4) {
		const double sum = ((Red[i * TR + c] + Red[(64 + i) * TR + c]) + Red[(128 + i) * TR + c]) +
		                   Red[(192 + i) * TR + c];
		if (orow >= m)
			continue;
		if (P) {
			P[((long)blockIdx.y * r + c0 + c) * m + orow] = sum;
		} else {
			double *p = C + orow + (long)(c0 + c) * ldc;
			const double cv = *p;
			*p = (beta == 0.0 ? 0.0 : beta * cv) + alpha * sum;
		}
	}
}

// C(i, c) = beta*C + alpha * sum over the ns k slices of P (in slice order).
__global__ __launch_bounds__(256)
void k_thin_reduce(int m, int r, int ns, double alpha, const double *__restrict__ P, double beta,
                   double *__restrict__ C, long ldc)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)m * r)
		return;
	const int i = (int)(idx % m), c = (int)(idx / m);
	double s = 0.0;
	for (int q = 0; q < ns; q++)
		s += P[((long)q * r + c) * m + i];
	double *p = C + i + (long)c * ldc;
	const double cv = *p;
	*p = (beta == 0.0 ? 0.0 : beta * cv) + alpha * s;
}

// MFMA form of the thin GEMM for tall operands (m >= 2048): FP64
// MFMAs (v_mfma_f64_16x16x4f64) instead of FMAs on the VALU, whose
// right-hand-side reads from LDS (16 doubles per k and lane) bounded
// k_gemm_thin at ~1.5 TB/s on such shapes.  Workgroup (x, y, z): four waves
// of 64 rows each (rows 256x + 64 wave ..), k slice y (kc columns of op(A)),
// right-hand sides 16z .. 16z+15.  A wave holds four 16x16 output blocks (row
// blocks j = 0..3) and walks its slice 32 k at a time: at MFMA step b = 0..7
// lane l (group g = l/16) supplies k = k0 + 8g + b -- a fixed partition of
// the 32 k, the same for both operands -- so that its A values come from one
// contiguous run:
//   TN: the 8 consecutive k of A's column row0 + 16j + l%16 (row block j
//       holds rows row0 + 16j + i, i = 0..15);
//   NN: per step b, the 4 consecutive rows row0 + 4(l%16) + e of column
//       k0 + 8g + b, one for each row block (row block j holds rows
//       row0 + 4i + j).
// Lane l loads B(k0 + 8g + b, 16z + l%16) itself (a slice of B is small and
// stays in L2).  Register q of block j's accumulator holds block row
// i = 4q + g.  Partial sums go to P and k_thin_reduce as in k_gemm_thin, so
// the result never depends on timing.  The host guarantees kdim % 8 == 0,
// kc % 32 == 0, lda % 4 == 0, A aligned to 4 elements, and m % 4 == 0 for
// NN, so that a lane's run of k (or rows) lies wholly inside or outside.
constexpr int MFMA_MIN_M = 2048;

template <bool TRANS, typename TA>
__global__ __launch_bounds__(256)
void k_gemm_thin_mfma(int m, int r, int kdim, int kc, double alpha, const TA *__restrict__ A, long lda,
                      const double *__restrict__ B, long ldb, double beta, double *__restrict__ C, long ldc,
                      double *__restrict__ P)
{
	typedef TA TA4 __attribute__((ext_vector_type(4)));
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int l16 = lane & 15, g = lane >> 4;
	const int row0 = (blockIdx.x * 4 + wave) * 64;
	const int kb = blockIdx.y * kc;
	const int ke = (kdim - kb) < kc ? kdim : kb + kc;
	const int c = blockIdx.z * 16 + l16;
	const bool cin = c < r;
	const double *Bc = B + (long)(cin ? c : r - 1) * ldb;
	sa_d4 acc[4];
#pragma unroll
	for (int j = 0; j < 4; j++)
		acc[j] = sa_d4{0.0, 0.0, 0.0, 0.0};
	for (int k0 = kb; k0 < ke; k0 += 32) {
		const int kk = k0 + 8 * g;
		const bool kin = kk < ke;
		const int kl = kin ? kk : kb;
		double bv[8], av[4][8];
#pragma unroll
		for (int b = 0; b < 8; b++) {
			const double v = Bc[kl + b];
			bv[b] = (kin && cin) ? v : 0.0;
		}
		if (TRANS) {
#pragma unroll
			for (int j = 0; j < 4; j++) {
				const int row = row0 + 16 * j + l16;
				const TA *p = A + kl + (long)(row < m ? row : m - 1) * lda;
				const TA4 x0 = *reinterpret_cast<const TA4 *>(p), x1 = *reinterpret_cast<const TA4 *>(p + 4);
#pragma unroll
				for (int e = 0; e < 4; e++) {
					av[j][e] = (double)x0[e];
					av[j][4 + e] = (double)x1[e];
				}
			}
		} else {
			const int rw = row0 + 4 * l16;
			const TA *p = A + (rw < m ? rw : 0);
#pragma unroll
			for (int b = 0; b < 8; b++) {
				const TA4 x = *reinterpret_cast<const TA4 *>(p + (long)(kl + b) * lda);
#pragma unroll
				for (int j = 0; j < 4; j++)
					av[j][b] = (double)x[j];
			}
		}
#pragma unroll
		for (int b = 0; b < 8; b++)
#pragma unroll
			for (int j = 0; j < 4; j++)
				acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[j][b], bv[b], acc[j], 0, 0, 0);
	}
	if (!cin)
		return;
#pragma unroll
	for (int j = 0; j < 4; j++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int i = 4 * q + g;
			const int row = TRANS ? row0 + 16 * j + i : row0 + 4 * i + j;
			if (row >= m)
				continue;
			if (P) {
				P[((long)blockIdx.y * r + c) * m + row] = acc[j][q];
			} else {
				double *pc = C + row + (long)c * ldc;
				*pc = (beta == 0.0 ? 0.0 : beta * *pc) + alpha * acc[j][q];
			}
		}
}

// Thin product for the solves with an FP32 factor, C(m x r) = beta*C +
// alpha*op(A)*B with A FP32 and B, C FP64, without partial sums in memory:
// a workgroup computes 16 RB rows x 16 right-hand sides (rows 16 RB x..,
// columns 16y..), its NW waves each take a contiguous share of k (whole
// 32-wide steps) and all RB 16-row blocks, with the lane layout of
// k_gemm_thin_mfma -- lane l supplies k = k0 + 8g + b at MFMA step b
// (g = l/16), B's element for column l%16 and A's for row l%16 of each
// block -- and the waves' partial sums meet in LDS, added in wave order, so
// the result never depends on timing.  A wave issues the loads of U steps
// before their MFMAs: the product is bound by load latency, and one round
// trip per U steps replaces one per step.  The RB blocks of a wave share its
// B values: with RB = 1 (short products, 16-row workgroups: a 1024-row
// product gets 64 of them) every workgroup reads all of its k-range of B
// per 16 rows, which on tall operands moves twice the bytes of A through
// the caches and ran a 31744 x 1024 panel at 2 TB/s; RB = 4 shares each B
// load among four blocks.  TN reads 8 consecutive k of A's column as two
// 4-float vectors (host: k % 8 == 0, lda % 4 == 0, A 16-byte aligned); NN
// reads A(row, k) per k, for RB = 4 as one 4-row vector per k with block j
// holding rows 16 RB x + 4i + j (host: also m % 4 == 0).
template <bool TRANS, int RB, int NW>
__global__ __launch_bounds__(64 * NW)
void k_thin_rows(int m, int r, int kdim, double alpha, const float *__restrict__ A, long lda,
                 const double *__restrict__ B, long ldb, double beta, double *__restrict__ C, long ldc)
{
	typedef float TA4 __attribute__((ext_vector_type(4)));
	static_assert(RB == 1 || RB == 4, "row blocks per wave: 1 or 4");
	constexpr int U = RB == 1 ? 4 : 2;
	__shared__ sa_d4 part[NW - 1][RB][64];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int l16 = lane & 15, g = lane >> 4;
	const int row0 = blockIdx.x * 16 * RB;
	const int c = blockIdx.y * 16 + l16;
	const bool cin = c < r;
	const double *Bc = B + (long)(cin ? c : r - 1) * ldb;
	const int steps = (kdim + 31) / 32, per = (steps + NW - 1) / NW;
	const int sb = wave * per;
	const int se = (steps - sb) < per ? steps : sb + per;
	sa_d4 acc[RB];
#pragma unroll
	for (int j = 0; j < RB; j++)
		acc[j] = sa_d4{0.0, 0.0, 0.0, 0.0};
	for (int s0 = sb; s0 < se; s0 += U) {
		float av[U][RB][8];
		double bv[U][8];
		bool kin[U];
#pragma unroll
		for (int u = 0; u < U; u++) {
			const int kk = 32 * (s0 + u) + 8 * g;
			kin[u] = s0 + u < se && kk < kdim;
			const int kl = kin[u] ? kk : 32 * sb;
#pragma unroll
			for (int b = 0; b < 8; b++)
				bv[u][b] = Bc[kl + b];
			if (TRANS) {
#pragma unroll
				for (int j = 0; j < RB; j++) {
					const int row = row0 + 16 * j + l16;
					const float *pa = A + kl + (long)(row < m ? row : m - 1) * lda;
					const TA4 x0 = *reinterpret_cast<const TA4 *>(pa), x1 = *reinterpret_cast<const TA4 *>(pa + 4);
#pragma unroll
					for (int e = 0; e < 4; e++) {
						av[u][j][e] = x0[e];
						av[u][j][4 + e] = x1[e];
					}
				}
			} else if (RB == 1) {
				const int row = row0 + l16;
				const int rowc = row < m ? row : m - 1;
#pragma unroll
				for (int b = 0; b < 8; b++)
					av[u][0][b] = A[rowc + (long)(kl + b) * lda];
			} else {
				const int rw = row0 + 4 * l16;
				const float *pa = A + (rw < m ? rw : 0);
#pragma unroll
				for (int b = 0; b < 8; b++) {
					const TA4 x = *reinterpret_cast<const TA4 *>(pa + (long)(kl + b) * lda);
#pragma unroll
					for (int j = 0; j < RB; j++)
						av[u][j][b] = x[j];
				}
			}
		}
#pragma unroll
		for (int u = 0; u < U; u++)
#pragma unroll
			for (int b = 0; b < 8; b++)
#pragma unroll
				for (int j = 0; j < RB; j++)
					acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(kin[u] ? (double)av[u][j][b] : 0.0,
					                                               (kin[u] && cin) ? bv[u][b] : 0.0, acc[j], 0, 0, 0);
	}
	if (wave > 0)
#pragma unroll
		for (int j = 0; j < RB; j++)
			part[wave - 1][j][lane] = acc[j];
	__syncthreads();
	if (wave > 0 || !cin)
		return;
	// register q of lane l: row 4q + g of block j, column l%16
#pragma unroll
	for (int j = 0; j < RB; j++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int i = 4 * q + g;
			const int rr = TRANS ? row0 + 16 * j + i : row0 + RB * i + j;
			if (rr >= m)
				continue;
			double sum = acc[j][q];
#pragma unroll
			for (int w = 0; w < NW - 1; w++)
				sum += part[w][j][lane][q];
			double *pc = C + rr + (long)c * ldc;
			*pc = (beta == 0.0 ? 0.0 : beta * *pc) + alpha * sum;
		}
}

// Shape of k_thin_rows for an m-row product: 16-row workgroups of eight
// waves below thin_tall_min_m() rows, 64-row workgroups (four blocks per
// wave) of thin_tall_waves() waves from there.  SA_THIN_TALL_MIN_M and
// SA_THIN_TALL_WAVES (4 or 8) override, for tuning.
int thin_tall_min_m()
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_THIN_TALL_MIN_M");
		const int x = e ? atoi(e) : 0;
		v = x > 0 ? x : 4096;
	}
	return v;
}

int thin_tall_waves()
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_THIN_TALL_WAVES");
		v = (e && atoi(e) == 4) ? 4 : 8;
	}
	return v;
}

// k_thin_rows when its layout conditions hold (SA_THIN_ROWS16=0 turns it
// off); 1 if the caller should take the split-k forms.
template <bool TRANS>
int thin_rows16(hipStream_t s, int m, int r, int k, double alpha, const float *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc, int *status)
{
	static int on = -1;
	if (on < 0) {
		const char *e = getenv("SA_THIN_ROWS16");
		on = (e && atoi(e) == 0) ? 0 : 1;
	}
	if (!on || k % 8 != 0 || (TRANS && (lda % 4 != 0 || ((uintptr_t)A % 16) != 0)))
		return 1;
	const bool tall = m >= thin_tall_min_m() &&
	                  (TRANS || (m % 4 == 0 && lda % 4 == 0 && ((uintptr_t)A % 16) == 0));
	if (tall) {
		const dim3 grid((unsigned)((m + 63) / 64), (unsigned)((r + 15) / 16));
		if (thin_tall_waves() == 4)
			hipLaunchKernelGGL((k_thin_rows<TRANS, 4, 4>), grid, dim3(256), 0, s, m, r, k, alpha, A, (long)lda, B,
			                   (long)ldb, beta, C, (long)ldc);
		else
			hipLaunchKernelGGL((k_thin_rows<TRANS, 4, 8>), grid, dim3(512), 0, s, m, r, k, alpha, A, (long)lda, B,
			                   (long)ldb, beta, C, (long)ldc);
	} else {
		const dim3 grid((unsigned)((m + 15) / 16), (unsigned)((r + 15) / 16));
		hipLaunchKernelGGL((k_thin_rows<TRANS, 1, 8>), grid, dim3(512), 0, s, m, r, k, alpha, A, (long)lda, B,
		                   (long)ldb, beta, C, (long)ldc);
	}
	*status = sa_launch_status();
	return 0;
}

// Target workgroup count of the MFMA thin GEMM (SA_THIN_MFMA_WORKGROUPS
// overrides; SA_THIN_MFMA=0 turns the MFMA form off).
int thin_mfma_target()
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_THIN_MFMA_WORKGROUPS");
		const int x = e ? atoi(e) : 0;
		v = (x >= 16 && x <= 8192) ? x : 512;
	}
	return v;
}

// Smallest m that takes the MFMA form (SA_THIN_MFMA_MIN_M overrides, >= 256).
int thin_mfma_min_m()
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_THIN_MFMA_MIN_M");
		const int x = e ? atoi(e) : 0;
		v = x >= 256 ? x : MFMA_MIN_M;
	}
	return v;
}

bool thin_mfma_enabled()
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_THIN_MFMA");
		v = (e && atoi(e) == 0) ? 0 : 1;
	}
	return v == 1;
}

// k slices of the MFMA form: about thin_mfma_target() workgroups, whole
// 32-wide k steps per slice.
void thin_mfma_split(int m, int r, int k, int *ns, int *kc)
{
	const int wx = (m + 255) / 256, nz = (r + 15) / 16;
	const int steps = (k + 31) / 32;
	int want = (thin_mfma_target() + wx * nz - 1) / (wx * nz);
	if (want > steps)
		want = steps;
	if (want < 1)
		want = 1;
	const int per = (steps + want - 1) / want;
	*kc = per * 32;
	*ns = (k + *kc - 1) / *kc;
}

size_t thin_mfma_worksize(int m, int r, int k)
{
	if (m < thin_mfma_min_m() || r <= 0 || k <= 0)
		return 0;
	int ns, kc;
	thin_mfma_split(m, r, k, &ns, &kc);
	return ns > 1 ? (size_t)ns * (size_t)r * (size_t)m * sizeof(double) : 0;
}

// Target workgroup count of the thin GEMM (SA_THIN_WORKGROUPS overrides).
int thin_target()
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_THIN_WORKGROUPS");
		const int x = e ? atoi(e) : 0;
		v = (x >= 16 && x <= 4096) ? x : 512;
	}
	return v;
}

// k slices of the thin GEMM: enough workgroups (about 512) to keep many
// loads in flight, at least one 64-column step per slice.
void thin_split(int m, int k, int tk, int *ns, int *kc)
{
	const int rb = (m + 63) / 64;
	const int steps = (k + tk - 1) / tk;
	int want = (thin_target() + rb - 1) / rb;
	if (want > steps)
		want = steps;
	if (want < 1)
		want = 1;
	const int per = (steps + want - 1) / want;     // tk-column steps per slice
	*kc = per * tk;
	*ns = (k + *kc - 1) / *kc;
}

size_t thin_worksize(int m, int r, int k, int tk = 64)
{
	if (m <= 0 || r <= 0 || k <= 0)
		return 0;
	int ns, kc;
	thin_split(m, k, tk, &ns, &kc);
	return ns > 1 ? (size_t)ns * (size_t)r * (size_t)m * sizeof(double) : 0;
}

// The MFMA form on a tall operand when its runs of k (TN) or rows (NN) can
// be loaded as aligned 4-element vectors; 0 if it ran (or failed to launch),
// 1 if the caller should take the VALU form.
template <bool TRANS, typename TA>
int gemm_thin_mfma(hipStream_t s, int m, int r, int k, double alpha, const TA *A, int lda, const double *B,
                   int ldb, double beta, double *C, int ldc, void *work, size_t work_bytes, int *status)
{
	if (!thin_mfma_enabled() || m < thin_mfma_min_m() || k % 8 != 0 || lda % 4 != 0 ||
	    ((uintptr_t)A % (4 * sizeof(TA))) != 0 || (!TRANS && m % 4 != 0))
		return 1;
	int ns, kc;
	thin_mfma_split(m, r, k, &ns, &kc);
	const size_t per_slice = (size_t)r * (size_t)m * sizeof(double);
	if (ns > 1 && (!work || work_bytes < (size_t)ns * per_slice)) {
		// fewer slices, as many as the workspace holds
		const size_t fit = work ? work_bytes / per_slice : 0;
		const int nfit = fit > (size_t)ns ? ns : (int)fit;
		if (nfit <= 1) {
			ns = 1;
			kc = (k + 31) / 32 * 32;
		} else {
			kc = ((k + nfit - 1) / nfit + 31) / 32 * 32;
			ns = (k + kc - 1) / kc;
		}
	}
	double *P = ns > 1 ? (double *)work : nullptr;
	const dim3 grid((unsigned)((m + 255) / 256), (unsigned)ns, (unsigned)((r + 15) / 16));
	hipLaunchKernelGGL((k_gemm_thin_mfma<TRANS, TA>), grid, dim3(256), 0, s, m, r, k, kc, alpha, A, (long)lda, B,
	                   (long)ldb, beta, C, (long)ldc, P);
	int e = sa_launch_status();
	if (!e && P) {
		const long cnt = (long)m * r;
		hipLaunchKernelGGL(k_thin_reduce, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m, r, ns, alpha,
		                   (const double *)P, beta, C, (long)ldc);
		e = sa_launch_status();
	}
	*status = e;
	return 0;
}

template <bool TRANS, typename TA = double>
int gemm_thin(hipStream_t s, int m, int r, int k, double alpha, const TA *A, int lda,
              const double *B, int ldb, double beta, double *C, int ldc, void *work, size_t work_bytes)
{
	if (m == 0 || r == 0)
		return 0;
	if (k == 0)
		return beta == 1.0 ? 0 : sa_scale_launch(s, m, r, C, ldc, beta, false);
	int status = 0;
	if (gemm_thin_mfma<TRANS, TA>(s, m, r, k, alpha, A, lda, B, ldb, beta, C, ldc, work, work_bytes, &status) == 0)
		return status;
	constexpr int tk = thin_tk<TA>();
	int ns, kc;
	thin_split(m, k, tk, &ns, &kc);
	double *P = nullptr;
	if (ns > 1) {
		if (work && work_bytes >= thin_worksize(m, r, k, tk)) {
			P = (double *)work;
		} else {            // no room for partial sums: one k slice
			ns = 1;
			kc = k;
		}
	}
	const dim3 grid((unsigned)((m + 63) / 64), (unsigned)ns, (unsigned)((r + TR - 1) / TR));
	hipLaunchKernelGGL((k_gemm_thin<TRANS, TA>), grid, dim3(256), 0, s, m, r, k, kc, alpha, A, (long)lda, B,
	                   (long)ldb, beta, C, (long)ldc, P);
	int e = sa_launch_status();
	if (e || !P)
		return e;
	const long cnt = (long)m * r;
	hipLaunchKernelGGL(k_thin_reduce, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m, r, ns, alpha,
	                   (const double *)P, beta, C, (long)ldc);
	return sa_launch_status();
}

bool ld_ok(int ld, int rows) { return ld >= (rows > 1 ? rows : 1); }

} // namespace

extern "C" size_t sa_dgemm_thin_worksize(int m, int n, int k)
{
	const size_t a = thin_worksize(m, n, k), b = thin_mfma_worksize(m, n, k);
	return a > b ? a : b;
}

extern "C" int sa_dgemm_nn(hipStream_t s, int m, int n, int k, double alpha, const double *A, int lda,
                           const double *B, int ldb, double beta, double *C, int ldc, void *work,
                           size_t work_bytes)
{
	if (m < 0 || n < 0 || k < 0 || !ld_ok(lda, m) || !ld_ok(ldb, k) || !ld_ok(ldc, m))
		return -EINVAL;
	return gemm_thin<false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work, work_bytes);
}

extern "C" int sa_dgemm_tn(hipStream_t s, int m, int n, int k, double alpha, const double *A, int lda,
                           const double *B, int ldb, double beta, double *C, int ldc, void *work,
                           size_t work_bytes)
{
	if (m < 0 || n < 0 || k < 0 || !ld_ok(lda, k) || !ld_ok(ldb, k) || !ld_ok(ldc, m))
		return -EINVAL;
	return gemm_thin<true>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work, work_bytes);
}

extern "C" size_t sa_dtrtri_lt_worksize(int n) { return sa_dtrsm_rltn_worksize(n); }

extern "C" int sa_dtrtri_lt(hipStream_t s, int n, const double *L, int ldl, double *W, int ldw, void *work,
                            size_t work_bytes)
{
	if (n < 0 || !ld_ok(ldl, n) || !ld_ok(ldw, n))
		return -EINVAL;
	if (n == 0)
		return 0;
	if (n <= 1024)
		return sa_trtri_lt_rows16(s, n, L, ldl, W, ldw, work, work_bytes);
	int e = sa_dlaset(s, n, n, 0.0, 1.0, W, ldw);
	if (e)
		return e;
	return sa_dtrsm_rltn(s, n, n, 1.0, L, ldl, W, ldw, work, work_bytes);
}

extern "C" size_t sa_dtrmm_lu_worksize(int m, int n)
{
	if (m <= 0 || n <= 0)
		return 0;
	const size_t copy = ((size_t)m * (size_t)n * sizeof(double) + 255) & ~(size_t)255;
	return copy + thin_worksize(m, n, m);
}

extern "C" int sa_dtrmm_lu(hipStream_t s, int trans, int m, int n, double alpha, const double *U, int ldu,
                           double *B, int ldb, void *work, size_t work_bytes)
{
	if (m < 0 || n < 0 || !ld_ok(ldu, m) || !ld_ok(ldb, m) || (trans != 0 && trans != 1))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	const size_t copy = ((size_t)m * (size_t)n * sizeof(double) + 255) & ~(size_t)255;
	if (!work || work_bytes < copy)
		return -EINVAL;
	double *T = (double *)work;
	int e = sa_dlacpy(s, m, n, B, ldb, T, m);
	if (e)
		return e;
	void *split = (char *)work + copy;
	const size_t split_bytes = work_bytes - copy;
	return trans ? gemm_thin<true>(s, m, n, m, alpha, U, ldu, T, m, 0.0, B, ldb, split, split_bytes)
	             : gemm_thin<false>(s, m, n, m, alpha, U, ldu, T, m, 0.0, B, ldb, split, split_bytes);
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
			e = gemm_thin<false>(s, m - j - w, n, w, -1.0, L + (j + w) + (long)j * ldl, ldl, B + j, ldb,
			                      j == 0 ? alpha : 1.0, B + j + w, ldb, nullptr, 0);
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
			e = gemm_thin<true>(s, j, n, w, -1.0, L + j, ldl, B + j, ldb, c == nch - 1 ? alpha : 1.0, B, ldb,
			                    nullptr, 0);
			if (e)
				return e;
		}
	}
	return 0;
}

// ---- FP32 factor, FP64 right-hand sides (the refinement solves) ----------

extern "C" int sa_gemm_nn_sd(hipStream_t s, int m, int n, int k, double alpha, const float *A, int lda,
                             const double *B, int ldb, double beta, double *C, int ldc, void *work,
                             size_t work_bytes)
{
	if (m < 0 || n < 0 || k < 0 || !ld_ok(lda, m) || !ld_ok(ldb, k) || !ld_ok(ldc, m))
		return -EINVAL;
	int status = 0;
	if (m > 0 && n > 0 && k > 0 && thin_rows16<false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, &status) == 0)
		return status;
	return gemm_thin<false, float>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work, work_bytes);
}

extern "C" int sa_gemm_tn_sd(hipStream_t s, int m, int n, int k, double alpha, const float *A, int lda,
                             const double *B, int ldb, double beta, double *C, int ldc, void *work,
                             size_t work_bytes)
{
	if (m < 0 || n < 0 || k < 0 || !ld_ok(lda, k) || !ld_ok(ldb, k) || !ld_ok(ldc, m))
		return -EINVAL;
	int status = 0;
	if (m > 0 && n > 0 && k > 0 && thin_rows16<true>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, &status) == 0)
		return status;
	return gemm_thin<true, float>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work, work_bytes);
}

extern "C" int sa_trmm_lu_sd(hipStream_t s, int trans, int m, int n, double alpha, const float *U, int ldu,
                             double *B, int ldb, void *work, size_t work_bytes)
{
	if (m < 0 || n < 0 || !ld_ok(ldu, m) || !ld_ok(ldb, m) || (trans != 0 && trans != 1))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	const size_t copy = ((size_t)m * (size_t)n * sizeof(double) + 255) & ~(size_t)255;
	if (!work || work_bytes < copy)
		return -EINVAL;
	double *T = (double *)work;
	int e = sa_dlacpy(s, m, n, B, ldb, T, m);
	if (e)
		return e;
	void *split = (char *)work + copy;
	const size_t split_bytes = work_bytes - copy;
	return trans ? gemm_thin<true, float>(s, m, n, m, alpha, U, ldu, T, m, 0.0, B, ldb, split, split_bytes)
	             : gemm_thin<false, float>(s, m, n, m, alpha, U, ldu, T, m, 0.0, B, ldb, split, split_bytes);
}
