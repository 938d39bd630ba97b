// FP64 GEMM NT and SYRK (lower) on the CDNA matrix cores (gfx950).
//
// C = beta*C + alpha*A*B^T with A (m x k), B (n x k), C (m x n), column-major.
// Each wave owns a (16*WB) x (16*WB) block of C as WB x WB tiles of
// v_mfma_f64_16x16x4f64; a workgroup has WVM x WVN waves.  Rows inside a
// wave are interleaved (MFMA tile b, MFMA row r -> C row WB*r + b) so each
// lane's A and B fragments for one k are WB contiguous doubles (16-byte
// loads) and its C values come in WB-row runs.  Operands stream from L2
// straight into registers through a two-stage register pipeline; stages
// past k load clamped addresses and are zeroed with selects, so the loop
// body has no branches and the accumulators stay in AGPRs.
//
// Two shapes are built: 64x64 waves in 128x128 workgroups (large problems,
// two workgroups per compute unit) and 32x32 waves in 64x64 workgroups,
// which gives a 1024x1024 tile 256 workgroups -- one per compute unit -- so
// a single tile GEMM fills the GPU.
#include <atomic>

#include "sa_common.h"
#include "starpu_amd_kernels.h"

namespace {

template <bool EDGE, int WB>
__device__ __forceinline__ void loadw(double (&v)[WB], const double *__restrict__ X, long ld,
                                      int row, int nrows, int k, int K)
{
	const bool kin = k < K;
	const long kc = kin ? k : K - 1;
	if (!EDGE) {
		const double *p = X + row + kc * ld;
#pragma unroll
		for (int h = 0; h < WB / 2; h++) {
			const sa_d2 t = *reinterpret_cast<const sa_d2 *>(p + 2 * h);
			v[2 * h] = kin ? t.x : 0.0;
			v[2 * h + 1] = kin ? t.y : 0.0;
		}
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

template <int WB, int WVM, int WVN, int KU, bool EDGE, bool LOWER, bool BETA0, int OCC = 1>
__global__ __launch_bounds__(64 * WVM * WVN, OCC)
void k_dgemm_nt(int M, int N, int K, double alpha, const double *__restrict__ A, long lda,
                const double *__restrict__ B, long ldb, double beta, double *__restrict__ C,
                long ldc, int tiles_m, int ktri)
{
	constexpr int WT = 16 * WB;
	constexpr int KS = 4 * KU;
	static_assert(!LOWER || WVM == WVN, "SYRK needs square workgroup tiles");
	int tm, tn;
	if (LOWER) {
		const int t = blockIdx.x;
		int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
		while ((i + 1) * (i + 2) / 2 <= t) i++;
		while (i * (i + 1) / 2 > t) i--;
		tm = i;
		tn = t - i * (i + 1) / 2;
	} else {
		// Workgroups are dealt to the 8 XCDs round robin (XCD = id % 8), and
		// each XCD has its own L2.  Give each XCD a contiguous run of
		// logical ids and walk the tiles in groups of G tile rows (G a power
		// of two dividing tiles_m, G*G <= workgroups per XCD), so one XCD
		// covers a near-square block of C tiles: for a 1024x1024 product in
		// 64x64 workgroups a 4x8 block (4 A and 8 B operand panels instead
		// of 2 and 16), in 128x128 workgroups a 2x4 block (instead of 1x8).
		int t = blockIdx.x;
		const int T = gridDim.x;
		if ((T & 7) == 0)
			t = (t & 7) * (T >> 3) + (t >> 3);
		const int tiles_n = T / tiles_m;
		int G = 1;
		while (2 * G <= tiles_m && tiles_m % (2 * G) == 0 && 4 * G * G <= (T >> 3))
			G *= 2;
		tm = (t / (G * tiles_n)) * G + t % G;
		tn = (t / G) % tiles_n;
	}
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const int row0 = tm * WT * WVM + (wave % WVM) * WT;
	const int col0 = tn * WT * WVN + (wave / WVM) * WT;
	if (LOWER && row0 + WT <= col0)
		return;         // wave block strictly above the diagonal
	if (EDGE && (row0 >= M || col0 >= N))
		return;
	// ktri: B (N x K) is lower triangular with zeros stored above its
	// diagonal, so the wave's columns col0 .. col0+WT-1 take only
	// k < col0 + WT (the terms left out are exact zeros; no barrier ties the
	// waves, so each stops at its own bound)
	if (ktri && col0 + WT < K)
		K = col0 + WT;
	const int r = lane & 15;
	const int g = lane >> 4;
	const int arow = row0 + WB * r;
	const int bcol = col0 + WB * r;

	sa_d4 acc[WB][WB];
	double fa0[KU][WB], fb0[KU][WB], fa1[KU][WB], fb1[KU][WB];

	auto load_stage = [&](double (&fa)[KU][WB], double (&fb)[KU][WB], int kb) {
#pragma unroll
		for (int s = 0; s < KU; s++) {
			const int k = kb + 4 * s + g;
			loadw<EDGE, WB>(fa[s], A, lda, arow, M, k, K);
			loadw<EDGE, WB>(fb[s], B, ldb, bcol, N, k, K);
		}
	};
	auto compute_stage = [&](const double (&fa)[KU][WB], const double (&fb)[KU][WB]) {
#pragma unroll
		for (int s = 0; s < KU; s++)
#pragma unroll
			for (int i = 0; i < WB; i++)
#pragma unroll
				for (int j = 0; j < WB; j++)
					acc[i][j] = sa_mfma(fa[s][i], fb[s][j], acc[i][j]);
	};

	// The first two stages are peeled and the first MFMA takes a zero
	// accumulator, so the accumulators are born in AGPRs and stay there.
	load_stage(fa0, fb0, 0);
	load_stage(fa1, fb1, KS);
#pragma unroll
	for (int s = 0; s < KU; s++)
#pragma unroll
		for (int i = 0; i < WB; i++)
#pragma unroll
			for (int j = 0; j < WB; j++)
				acc[i][j] = sa_mfma(fa0[s][i], fb0[s][j], s == 0 ? sa_d4{0.0, 0.0, 0.0, 0.0} : acc[i][j]);
	load_stage(fa0, fb0, 2 * KS);
	compute_stage(fa1, fb1);
	for (int kb = 2 * KS; kb < K; kb += 2 * KS) {
		load_stage(fa1, fb1, kb + KS);
		compute_stage(fa0, fb0);
		load_stage(fa0, fb0, kb + 2 * KS);
		compute_stage(fa1, fb1);
	}

	// acc[i][j][q] is C(row0 + WB*(4q+g) + i, col0 + WB*r + j).
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
#pragma unroll
				for (int h = 0; h < WB / 2; h++) {
					sa_d2 v = BETA0 ? sa_d2{0.0, 0.0} : *reinterpret_cast<sa_d2 *>(p + 2 * h) * beta;
					v.x += alpha * acc[2 * h][j][q];
					v.y += alpha * acc[2 * h + 1][j][q];
					*reinterpret_cast<sa_d2 *>(p + 2 * h) = v;
				}
			} else {
#pragma unroll
				for (int i = 0; i < WB; i++) {
					const int rr = row + i;
					if ((EDGE && rr >= M) || (LOWER && rr < col))
						continue;
					const double c0 = BETA0 ? 0.0 : beta * p[i];
					p[i] = c0 + alpha * acc[i][j][q];
				}
			}
		}
	}
}

// X <- a*X over an m x n region (lower: only i >= j).
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

template <int WB, int WV, int KU, bool LOWER, int OCC = 1>
int gemm_shape(hipStream_t s, int m, int n, int k, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc, int ktri = 0)
{
	constexpr int BT = 16 * WB * WV;
	const int tiles_m = (m + BT - 1) / BT;
	const int tiles_n = (n + BT - 1) / BT;
	const unsigned grid = LOWER ? (unsigned)(tiles_m * (tiles_m + 1) / 2) : (unsigned)(tiles_m * tiles_n);
	const bool vec = sa_aligned16(A) && sa_aligned16(B) && sa_aligned16(C) && (lda % 2 == 0) &&
	                 (ldb % 2 == 0) && (ldc % 2 == 0);
	const bool edge = !vec || (m % BT) || (n % BT);
	const bool b0 = (beta == 0.0);
	const dim3 blk(64 * WV * WV);
#define SA_GEMM_LAUNCH(E, Z)                                                                      \
	hipLaunchKernelGGL((k_dgemm_nt<WB, WV, WV, KU, E, LOWER, Z, OCC>), dim3(grid), blk, 0, s, m, n, k,  \
	                   alpha, A, (long)lda, B, (long)ldb, beta, C, (long)ldc, tiles_m, ktri)
	if (edge) {
		if (b0) SA_GEMM_LAUNCH(true, true); else SA_GEMM_LAUNCH(true, false);
	} else {
		if (b0) SA_GEMM_LAUNCH(false, true); else SA_GEMM_LAUNCH(false, false);
	}
#undef SA_GEMM_LAUNCH
	return sa_launch_status();
}

// Launches the caller keeps in flight on one device (sa_set_launch_concurrency).
std::atomic<int> g_concurrency{1};

// Forced workgroup tile (64 or 128), or 0 for the choice by launch
// concurrency.  Read from SA_GEMM_SHAPE on first use; sa_set_gemm_shape
// changes it at run time.
std::atomic<int> g_shape_override{-1};

int tile_shape_override()
{
	int v = g_shape_override.load(std::memory_order_relaxed);
	if (v < 0) {
		const char *e = getenv("SA_GEMM_SHAPE");
		const int w = e ? atoi(e) : 0;
		g_shape_override.compare_exchange_strong(v, (w == 64 || w == 128) ? w : 0);
		v = g_shape_override.load(std::memory_order_relaxed);
	}
	return v;
}

// Workgroup tile edge (64 or 128) that a product of this shape runs with.
// 128x128 workgroups (half the operand traffic per flop) while the
// launches in flight together give every compute unit about 1.5 of them;
// otherwise 64x64 workgroups (4x as many).  Alone, a 1024x1024 tile product
// thus runs as 256 workgroups of 64x64 (40.6 TFLOP/s); with 12 streams of
// such products in flight, 128x128 workgroups ran the N=32768 tile-1024
// factorisation at 49.0 TFLOP/s against 44.9 with 64x64 (same machine, 16
// hardware queues).
int pick_shape(bool lower, int m, int n)
{
	const int forced = tile_shape_override();
	if (forced == 64 || forced == 128)
		return forced;
	const long big_tiles = lower ? (long)((m + 127) / 128) * ((m + 127) / 128 + 1) / 2
	                             : (long)((m + 127) / 128) * ((n + 127) / 128);
	return big_tiles * g_concurrency.load(std::memory_order_relaxed) >= 384 ? 128 : 64;
}

} // namespace

int sa_scale_launch(hipStream_t s, int m, int n, double *X, int ld, double a, bool lower)
{
	const long cnt = (long)m * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_scale, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m, n, X,
	                   (long)ld, a, lower ? 1 : 0);
	return sa_launch_status();
}

namespace {

int gemm_launch(hipStream_t s, bool lower, int m, int n, int k, double alpha, const double *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc, int ktri)
{
	if (m == 0 || n == 0)
		return 0;
	if (k == 0)
		return beta == 1.0 ? 0 : sa_scale_launch(s, m, n, C, ldc, beta, lower);
	const int shape = pick_shape(lower, m, n);
	// 128x128: two k-steps per pipeline stage, so that two workgroups fit on
	// a CU (216 VGPRs); with four k-steps per stage (316 registers, one
	// workgroup per CU) the tile-1024 factorisation ran at 51.2 against 54.0
	// TFLOP/s and an 8192^3 product at 63 against 66 TFLOP/s; one k-step per
	// stage at three workgroups per CU (164 VGPRs) ran it at 47.2.
	if (shape == 128)
		return lower ? gemm_shape<4, 2, 2, true, 2>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ktri)
		             : gemm_shape<4, 2, 2, false, 2>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ktri);
	return lower ? gemm_shape<2, 2, 8, true>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ktri)
	             : gemm_shape<2, 2, 8, false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ktri);
}

} // namespace

int sa_gemm_launch(hipStream_t s, bool lower, int m, int n, int k, double alpha, const double *A,
                   int lda, const double *B, int ldb, double beta, double *C, int ldc)
{
	return gemm_launch(s, lower, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

extern "C" int sa_set_launch_concurrency(int n)
{
	return g_concurrency.exchange(n > 0 ? n : 1);
}

extern "C" int sa_set_gemm_shape(int edge)
{
	if (edge != 0 && edge != 64 && edge != 128)
		return -EINVAL;
	tile_shape_override();   // settle the environment's value first
	return g_shape_override.exchange(edge);
}

extern "C" int sa_gemm_tile_shape(int lower, int m, int n)
{
	if (m <= 0 || n <= 0)
		return -EINVAL;
	return pick_shape(lower != 0, m, lower ? m : n);
}

extern "C" int sa_dgemm_nt(hipStream_t s, int m, int n, int k, double alpha, const double *A,
                           int lda, const double *B, int ldb, double beta, double *C, int ldc)
{
	if (m < 0 || n < 0 || k < 0 || ldc < (m > 1 ? m : 1) ||
	    (k > 0 && (lda < (m > 1 ? m : 1) || ldb < (n > 1 ? n : 1))))
		return -EINVAL;
	return sa_gemm_launch(s, false, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int sa_dgemm_nt_ltri(hipStream_t s, int m, int n, double alpha, const double *A, int lda,
                                const double *L, int ldl, double beta, double *C, int ldc)
{
	if (m < 0 || n < 0 || ldc < (m > 1 ? m : 1) || (n > 0 && (lda < (m > 1 ? m : 1) || ldl < (n > 1 ? n : 1))))
		return -EINVAL;
	return gemm_launch(s, false, m, n, n, alpha, A, lda, L, ldl, beta, C, ldc, 1);
}

extern "C" int sa_dsyrk_ln(hipStream_t s, int n, int k, double alpha, const double *A, int lda,
                           double beta, double *C, int ldc)
{
	if (n < 0 || k < 0 || ldc < (n > 1 ? n : 1) || (k > 0 && lda < (n > 1 ? n : 1)))
		return -EINVAL;
	return sa_gemm_launch(s, true, n, n, k, alpha, A, lda, A, lda, beta, C, ldc);
}
