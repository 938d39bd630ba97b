// Hand-written gfx950 FP64 tile kernels for the tiled Cholesky path:
// GEMM NT, SYRK lower, TRSM right/lower/trans and POTRF lower.
//
// Layout: column-major tiles, leading dimension in elements, as the StarPU
// matrix interface hands them to a codelet.
//
// GEMM/SYRK run on the CDNA matrix cores with v_mfma_f64_16x16x4f64.  Each
// wave owns a 64x64 block of C as 4x4 MFMA blocks; the row index inside
// the wave is interleaved (MFMA block b, MFMA row r  ->  C row 4*r + b) so
// that every lane's A and B fragments for one k are 4 contiguous doubles
// (two 16-byte loads) and every lane's C outputs come in 4-row runs.
// Operands stream from L2 straight into registers with a two-stage
// register pipeline (16 k-values per stage); FP64 MFMA is long enough
// (64 cycles per 16x16x4 on gfx950) that L1/L2 bandwidth is not the limit.
//
// TRSM and POTRF are blocked by 64 columns: a one-wave-per-64-rows panel
// kernel does the triangular part in registers against a 64x64 diagonal
// block held in LDS, and the trailing update goes to the MFMA GEMM/SYRK.
#include <hip/hip_runtime.h>
#include <errno.h>
#include <stdint.h>

#include "starpu_amd_kernels.h"

typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int WB = 4;           // MFMA blocks per wave edge
constexpr int WT = 16 * WB;     // wave tile edge (64)
constexpr int WV = 2;           // waves per WG edge
constexpr int BT = WT * WV;     // WG tile edge (128)
constexpr int KU = 4;           // k4 steps per pipeline stage
constexpr int KS = 4 * KU;      // k values per stage (16)

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c)
{
	return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Load one lane's 4 contiguous doubles X[row .. row+3, k].  Out-of-range
// k (and rows, for EDGE) read a clamped in-range address and are replaced
// by zero with a select, so the load sequence has no branches (branches
// around the loads make the compiler shuttle the MFMA accumulators between
// AGPRs and VGPRs every iteration).  Requires K >= 1 and nrows >= 1.
template <bool EDGE>
__device__ __forceinline__ void load4(double (&v)[WB], const double *__restrict__ X,
                                      long ld, int row, int nrows, int k, int K)
{
	const bool kin = k < K;
	const long kc = kin ? k : K - 1;
	if (!EDGE) {
		const double *p = X + row + kc * ld;
		const d2 lo = *reinterpret_cast<const d2 *>(p);
		const d2 hi = *reinterpret_cast<const d2 *>(p + 2);
		v[0] = kin ? lo.x : 0.0;
		v[1] = kin ? lo.y : 0.0;
		v[2] = kin ? hi.x : 0.0;
		v[3] = kin ? hi.y : 0.0;
	} else {
#pragma unroll
		for (int b = 0; b < WB; b++) {
			const int rr = row + b;
			const bool in = kin && rr < nrows;
			const double val = X[(in ? rr : nrows - 1) + kc * ld];
			v[b] = in ? val : 0.0;
		}
	}
}

// C = beta*C + alpha*A*B^T.  LOWER: only the lower triangle of the square C
// is touched and the grid enumerates lower WG tiles only (SYRK).
template <bool EDGE, bool LOWER, bool BETA0>
__global__ __launch_bounds__(64 * WV * WV)
void k_dgemm_nt(int M, int N, int K, double alpha,
                const double *__restrict__ A, long lda,
                const double *__restrict__ B, long ldb,
                double beta, double *__restrict__ C, long ldc, int tiles_m)
{
	int tm, tn;
	if (LOWER) {
		const int t = blockIdx.x;
		int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
		while ((i + 1) * (i + 2) / 2 <= t) i++;
		while (i * (i + 1) / 2 > t) i--;
		tm = i;
		tn = t - i * (i + 1) / 2;
	} else {
		tm = blockIdx.x % tiles_m;
		tn = blockIdx.x / tiles_m;
	}
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const int row0 = tm * BT + (wave % WV) * WT;
	const int col0 = tn * BT + (wave / WV) * WT;
	if (LOWER && row0 + WT <= col0)
		return;         // wave block strictly above the diagonal
	if (EDGE && (row0 >= M || col0 >= N))
		return;
	const int r = lane & 15;
	const int g = lane >> 4;
	const int arow = row0 + WB * r;
	const int bcol = col0 + WB * r;

	d4 acc[WB][WB];

	// Two register stages of KS k-values each; stages past K load zeros,
	// so the loop body is uniform (no early exit).  The first two stages are
	// peeled and the first MFMA takes a zero accumulator, so the
	// accumulators are born in AGPRs and stay there across the loop.
	double fa0[KU][WB], fb0[KU][WB], fa1[KU][WB], fb1[KU][WB];

	auto load_stage = [&](double (&fa)[KU][WB], double (&fb)[KU][WB], int kb) {
#pragma unroll
		for (int s = 0; s < KU; s++) {
			const int k = kb + 4 * s + g;
			load4<EDGE>(fa[s], A, lda, arow, M, k, K);
			load4<EDGE>(fb[s], B, ldb, bcol, N, k, K);
		}
	};
	auto compute_stage = [&](const double (&fa)[KU][WB], const double (&fb)[KU][WB]) {
#pragma unroll
		for (int s = 0; s < KU; s++)
#pragma unroll
			for (int i = 0; i < WB; i++)
#pragma unroll
				for (int j = 0; j < WB; j++)
					acc[i][j] = mfma_f64(fa[s][i], fb[s][j], acc[i][j]);
	};

	load_stage(fa0, fb0, 0);
	load_stage(fa1, fb1, KS);
#pragma unroll
	for (int s = 0; s < KU; s++)
#pragma unroll
		for (int i = 0; i < WB; i++)
#pragma unroll
			for (int j = 0; j < WB; j++)
				acc[i][j] = mfma_f64(fa0[s][i], fb0[s][j], s == 0 ? d4{0.0, 0.0, 0.0, 0.0} : acc[i][j]);
	load_stage(fa0, fb0, 2 * KS);
	compute_stage(fa1, fb1);
	for (int kb = 2 * KS; kb < K; kb += 2 * KS) {
		load_stage(fa1, fb1, kb + KS);
		compute_stage(fa0, fb0);
		load_stage(fa0, fb0, kb + 2 * KS);
		compute_stage(fa1, fb1);
	}

	// v_mfma_f64_16x16x4f64 leaves D(4*q + lane/16, lane%16) in register q
	// of each lane, so acc[i][j][q] is C(row0 + WB*(4q+g) + i, col0 + WB*r + j).
#pragma unroll
	for (int j = 0; j < WB; j++) {
		const int col = col0 + WB * r + j;
		if (EDGE && col >= N)
			continue;
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int row = row0 + WB * (4 * q + g);
			double *p = C + row + (long)col * ldc;
			const bool fast = !EDGE && !(LOWER && row < col);
			if (fast) {
				d2 lo, hi;
				if (BETA0) {
					lo = d2{0.0, 0.0};
					hi = d2{0.0, 0.0};
				} else {
					lo = *reinterpret_cast<d2 *>(p) * beta;
					hi = *reinterpret_cast<d2 *>(p + 2) * beta;
				}
				lo.x += alpha * acc[0][j][q];
				lo.y += alpha * acc[1][j][q];
				hi.x += alpha * acc[2][j][q];
				hi.y += alpha * acc[3][j][q];
				*reinterpret_cast<d2 *>(p) = lo;
				*reinterpret_cast<d2 *>(p + 2) = hi;
			} else {
#pragma unroll
				for (int i = 0; i < WB; i++) {
					const int rr = row + i;
					if ((EDGE && rr >= M) || (LOWER && rr < col))
						continue;
					double c0 = BETA0 ? 0.0 : beta * p[i];
					p[i] = c0 + alpha * acc[i][j][q];
				}
			}
		}
	}
}

__device__ __forceinline__ double bcast(double v, int lane)
{
	// v of lane `lane`, broadcast to the whole wave through SGPRs.
	const long long bits = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), lane);
	const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
	return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Cholesky of the w x w (w <= 64) lower block D by one wave, in registers:
// lane t holds row t; step k broadcasts L(k,k) and the column L(j,k) from
// the owning lanes with v_readlane.  Rows/cols past w are identity.  The
// strict upper part of each lane's row accumulates garbage that is never
// read (only lanes j > k are ever broadcast for column k) nor stored.
// info: col_base == 0 stores 0 or the failing column; later calls only
// record the first failure.
__global__ __launch_bounds__(64)
void k_potrf64(int w, double *__restrict__ D, long ldd, int *__restrict__ info, int col_base)
{
	const int t = threadIdx.x;
	double x[64];
#pragma unroll
	for (int c = 0; c < 64; c++) {
		double v = (t == c) ? 1.0 : 0.0;
		if (t < w && c < w && c <= t)
			v = D[t + (long)c * ldd];
		x[c] = v;
	}
	int bad = 0;
#pragma unroll
	for (int k = 0; k < 64; k++) {
		const double d = bcast(x[k], k);
		if (!(d > 0.0) && bad == 0)
			bad = k + 1;
		const double sd = sqrt(d);
		x[k] = (t == k) ? sd : x[k] / sd;
#pragma unroll
		for (int j = k + 1; j < 64; j++)
			x[j] -= x[k] * bcast(x[k], j);
	}
	if (t < w)
		for (int c = 0; c < 64; c++)   // not unrolled: see k_trsm_panel64
			if (c <= t && c < w) {
				double v = 0.0;
#pragma unroll
				for (int q = 0; q < 64; q++)
					v = (q == c) ? x[q] : v;
				D[t + (long)c * ldd] = v;
			}
	if (t == 0) {
		if (col_base == 0)
			*info = (bad && bad <= w) ? bad : 0;
		else if (bad && bad <= w)
			atomicCAS(info, 0, col_base + bad);
	}
}

// P(m x w) <- alpha * P * L^{-T} for the w x w (w <= 64) lower L, one row
// of P per lane: x_c = (x_c - sum_{q<c} x_q L(c,q)) / L(c,c), applied
// column by column.  L is read with wave-uniform addresses, so the
// compiler fetches it through the scalar cache into SGPRs; the FMAs then
// take L as a scalar operand and no LDS round trip sits on the
// dependency chain.
template <bool FULL>
__global__ __launch_bounds__(64)
void k_trsm_panel64(int m, int w, const double *__restrict__ L, long ldl,
                    double *__restrict__ P, long ldp, double alpha)
{
	// FULL (w == 64): L is read with wave-uniform addresses straight from
	// global memory (scalar loads into SGPRs).  Partial blocks (w < 64, only
	// the last block of a size that is not a multiple of 64) first copy L
	// into LDS padded with the identity and read it from there.
	__shared__ double stage[64][65];
	const int t = threadIdx.x;
	const int i = blockIdx.x * 64 + t;
	const int ic = i < m ? i : m - 1;
	if (!FULL) {
		for (int c = 0; c < 64; c++)
			stage[c][t] = (t < w && c < w) ? ((t >= c) ? L[t + (long)c * ldl] : 0.0)
			                               : ((t == c) ? 1.0 : 0.0);
		__syncthreads();
	}
	double x[64];
#pragma unroll
	for (int c = 0; c < 64; c++)
		x[c] = (FULL || c < w) ? alpha * P[ic + (long)c * ldp] : 0.0;
#pragma unroll
	for (int c = 0; c < 64; c++) {
		x[c] = x[c] / (FULL ? L[c + (long)c * ldl] : stage[c][c]);
#pragma unroll
		for (int q = c + 1; q < 64; q++)
			x[q] -= x[c] * (FULL ? L[q + (long)c * ldl] : stage[c][q]);
	}
	// Rows past m computed on a clamped copy and are not stored; the
	// result goes through LDS so that w only guards a plain store loop
	// (guarding 64 unrolled register stores makes the compiler spill).
	__syncthreads();
#pragma unroll
	for (int c = 0; c < 64; c++)
		stage[c][t] = x[c];
	__syncthreads();
	if (i < m)
		for (int c = 0; c < w; c++)
			P[i + (long)c * ldp] = stage[c][t];
}

inline void launch_trsm_panel(hipStream_t s, int m, int w, const double *L, long ldl,
                              double *P, long ldp, double alpha)
{
	const dim3 grid((unsigned)((m + 63) / 64));
	if (w == 64)
		hipLaunchKernelGGL(k_trsm_panel64<true>, grid, dim3(64), 0, s, m, w, L, ldl, P, ldp, alpha);
	else
		hipLaunchKernelGGL(k_trsm_panel64<false>, grid, dim3(64), 0, s, m, w, L, ldl, P, ldp, alpha);
}

// X <- a*X over an m x n region (lower: only i >= j).  a == 0 stores zeros,
// as BLAS does for beta == 0, so NaNs in X do not survive.
__global__ void k_scale(int m, int n, double *X, long ld, double a, int lower)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)m * n)
		return;
	const int i = (int)(idx % m), j = (int)(idx / m);
	if (lower && i < j)
		return;
	double *p = X + i + (long)j * ld;
	*p = (a == 0.0) ? 0.0 : a * *p;
}

// Counter-based generator: splitmix64 finaliser of (seed, i, j).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t h)   // in (0, 1)
{
	return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__global__ void k_fill_random(int m, int n, double *X, long ld, uint64_t seed, int dist)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)m * n)
		return;
	const int i = (int)(idx % m), j = (int)(idx / m);
	const uint64_t key = mix64(seed ^ mix64(((uint64_t)(uint32_t)j << 32) | (uint32_t)i));
	double v;
	if (dist == 0) {
		v = 2.0 * u01(mix64(key)) - 1.0;
	} else {
		const double u1 = u01(mix64(key));
		const double u2 = u01(mix64(key ^ 0xD1B54A32D192ED03ull));
		v = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
	}
	X[i + (long)j * ld] = v;
}

__global__ void k_add_diag(int n, double *X, long ld, double shift)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		X[i + (long)i * ld] += shift;
}

__global__ void k_symmetrize(int n, double *X, long ld)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)n * n)
		return;
	const int i = (int)(idx % n), j = (int)(idx / n);
	if (i < j)
		X[i + (long)j * ld] = X[j + (long)i * ld];
}

inline int launch_status()
{
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? 0 : (int)e;
}

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

template <bool LOWER>
int gemm_dispatch(hipStream_t s, int m, int n, int k, double alpha,
                  const double *A, int lda, const double *B, int ldb,
                  double beta, double *C, int ldc)
{
	if (k == 0) {
		if (beta == 1.0)
			return 0;
		const long cnt = (long)m * n;
		hipLaunchKernelGGL(k_scale, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
		                   m, n, C, (long)ldc, beta, LOWER ? 1 : 0);
		return launch_status();
	}
	const int tiles_m = (m + BT - 1) / BT;
	const int tiles_n = (n + BT - 1) / BT;
	const unsigned grid = LOWER ? (unsigned)(tiles_m * (tiles_m + 1) / 2)
	                            : (unsigned)(tiles_m * tiles_n);
	const bool vec = aligned16(A) && aligned16(B) && aligned16(C) &&
	                 (lda % 2 == 0) && (ldb % 2 == 0) && (ldc % 2 == 0);
	const bool edge = !vec || (m % BT) || (n % BT);
	const bool b0 = (beta == 0.0);
	dim3 blk(64 * WV * WV);
#define SA_GEMM_LAUNCH(E, Z) \
	hipLaunchKernelGGL((k_dgemm_nt<E, LOWER, Z>), dim3(grid), blk, 0, s, m, n, k, alpha, \
	                   A, (long)lda, B, (long)ldb, beta, C, (long)ldc, tiles_m)
	if (edge) {
		if (b0) SA_GEMM_LAUNCH(true, true); else SA_GEMM_LAUNCH(true, false);
	} else {
		if (b0) SA_GEMM_LAUNCH(false, true); else SA_GEMM_LAUNCH(false, false);
	}
#undef SA_GEMM_LAUNCH
	return launch_status();
}

} // namespace

extern "C" int sa_dgemm_nt(hipStream_t s, int m, int n, int k, double alpha,
                           const double *A, int lda, const double *B, int ldb,
                           double beta, double *C, int ldc)
{
	if (m < 0 || n < 0 || k < 0 || ldc < (m > 1 ? m : 1) ||
	    (k > 0 && (lda < (m > 1 ? m : 1) || ldb < (n > 1 ? n : 1))))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	return gemm_dispatch<false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int sa_dsyrk_ln(hipStream_t s, int n, int k, double alpha,
                           const double *A, int lda, double beta, double *C, int ldc)
{
	if (n < 0 || k < 0 || ldc < (n > 1 ? n : 1) || (k > 0 && lda < (n > 1 ? n : 1)))
		return -EINVAL;
	if (n == 0)
		return 0;
	return gemm_dispatch<true>(s, n, n, k, alpha, A, lda, A, lda, beta, C, ldc);
}

extern "C" int sa_dtrsm_rltn(hipStream_t s, int m, int n, double alpha,
                             const double *L, int ldl, double *B, int ldb)
{
	if (m < 0 || n < 0 || ldl < (n > 1 ? n : 1) || ldb < (m > 1 ? m : 1))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	for (int j = 0; j < n; j += 64) {
		const int w = (n - j) < 64 ? (n - j) : 64;
		launch_trsm_panel(s, m, w, L + j + (long)j * ldl, (long)ldl, B + (long)j * ldb,
		                  (long)ldb, j == 0 ? alpha : 1.0);
		int e = launch_status();
		if (e)
			return e;
		if (j + w < n) {
			// The remaining columns still carry alpha*B; fold alpha into
			// them before they are updated with already-scaled X columns.
			if (j == 0 && alpha != 1.0) {
				const long cnt = (long)m * (n - w);
				hipLaunchKernelGGL(k_scale, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
				                   m, n - w, B + (long)w * ldb, (long)ldb, alpha, 0);
				e = launch_status();
				if (e)
					return e;
			}
			e = gemm_dispatch<false>(s, m, n - j - w, w, -1.0, B + (long)j * ldb, ldb,
			                         L + (j + w) + (long)j * ldl, ldl, 1.0,
			                         B + (long)(j + w) * ldb, ldb);
			if (e)
				return e;
		}
	}
	return 0;
}

extern "C" int sa_dpotrf_l(hipStream_t s, int n, double *A, int lda, int *dev_info)
{
	if (n < 0 || lda < (n > 1 ? n : 1) || dev_info == nullptr)
		return -EINVAL;
	if (n == 0) {
		hipLaunchKernelGGL(k_potrf64, dim3(1), dim3(64), 0, s, 0, A, (long)lda, dev_info, 0);
		return launch_status();
	}
	for (int j = 0; j < n; j += 64) {
		const int w = (n - j) < 64 ? (n - j) : 64;
		const int rows = n - j - w;
		double *Djj = A + j + (long)j * lda;
		hipLaunchKernelGGL(k_potrf64, dim3(1), dim3(64), 0, s, w, Djj, (long)lda, dev_info, j);
		int e = launch_status();
		if (e)
			return e;
		if (rows > 0) {
			launch_trsm_panel(s, rows, w, Djj, (long)lda, Djj + w, (long)lda, 1.0);
			e = launch_status();
			if (e)
				return e;
			e = gemm_dispatch<true>(s, rows, rows, w, -1.0, Djj + w, lda, Djj + w, lda,
			                        1.0, A + (j + w) + (long)(j + w) * lda, lda);
			if (e)
				return e;
		}
	}
	return 0;
}

// ------------------------------------------------------------ helpers ----

extern "C" int sa_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess)
		return 0;
	return n;
}

extern "C" int sa_set_device(int dev) { return (int)hipSetDevice(dev); }

extern "C" void *sa_dev_malloc(size_t bytes)
{
	void *p = nullptr;
	if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess)
		return nullptr;
	return p;
}

extern "C" int sa_dev_free(void *p) { return (int)hipFree(p); }
extern "C" int sa_memcpy_h2d(void *d, const void *s, size_t b) { return (int)hipMemcpy(d, s, b, hipMemcpyHostToDevice); }
extern "C" int sa_memcpy_d2h(void *d, const void *s, size_t b) { return (int)hipMemcpy(d, s, b, hipMemcpyDeviceToHost); }
extern "C" int sa_memcpy_d2d(void *d, const void *s, size_t b) { return (int)hipMemcpy(d, s, b, hipMemcpyDeviceToDevice); }

extern "C" int sa_memcpy2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                                 size_t width, size_t height, int kind, hipStream_t s)
{
	return (int)hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, (hipMemcpyKind)kind, s);
}

extern "C" int sa_stream_create(hipStream_t *s) { return (int)hipStreamCreateWithFlags(s, hipStreamNonBlocking); }
extern "C" int sa_stream_destroy(hipStream_t s) { return (int)hipStreamDestroy(s); }
extern "C" int sa_stream_sync(hipStream_t s) { return (int)hipStreamSynchronize(s); }
extern "C" int sa_device_sync(void) { return (int)hipDeviceSynchronize(); }
extern "C" const char *sa_error_string(int e) { return hipGetErrorString((hipError_t)e); }

struct sa_timer {
	hipEvent_t a, b;
	hipStream_t s;
};

extern "C" int sa_timer_start(hipStream_t s, void **timer)
{
	sa_timer *t = new sa_timer;
	t->s = s;
	hipError_t e = hipEventCreate(&t->a);
	if (e == hipSuccess)
		e = hipEventCreate(&t->b);
	if (e == hipSuccess)
		e = hipEventRecord(t->a, s);
	if (e != hipSuccess) {
		delete t;
		return (int)e;
	}
	*timer = t;
	return 0;
}

extern "C" int sa_timer_stop(void *timer, float *ms)
{
	sa_timer *t = (sa_timer *)timer;
	hipError_t e = hipEventRecord(t->b, t->s);
	if (e == hipSuccess)
		e = hipEventSynchronize(t->b);
	if (e == hipSuccess)
		e = hipEventElapsedTime(ms, t->a, t->b);
	hipEventDestroy(t->a);
	hipEventDestroy(t->b);
	delete t;
	return (int)e;
}

extern "C" int sa_fill_random(hipStream_t s, int m, int n, double *X, int ld,
                              uint64_t seed, int dist)
{
	if (m < 0 || n < 0 || ld < (m > 1 ? m : 1) || (dist != 0 && dist != 1))
		return -EINVAL;
	const long cnt = (long)m * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_fill_random, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
	                   m, n, X, (long)ld, seed, dist);
	return launch_status();
}

extern "C" int sa_add_diagonal(hipStream_t s, int n, double *X, int ld, double shift)
{
	if (n < 0 || ld < (n > 1 ? n : 1))
		return -EINVAL;
	if (n == 0)
		return 0;
	hipLaunchKernelGGL(k_add_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, X, (long)ld, shift);
	return launch_status();
}

extern "C" int sa_symmetrize_lower(hipStream_t s, int n, double *X, int ld)
{
	if (n < 0 || ld < (n > 1 ? n : 1))
		return -EINVAL;
	const long cnt = (long)n * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, n, X, (long)ld);
	return launch_status();
}
