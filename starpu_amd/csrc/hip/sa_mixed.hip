// Kernels of the mixed-precision Cholesky on gfx950.
//
// The FP32 working copy of A is factored as A = U^T U on its UPPER triangle,
// stored column-major.  Every trailing update then reads both operands
// along k with unit stride -- U_km^T U_kn for C_mn -- which is the operand
// layout the MFMA instructions want (8 consecutive k per lane for the BF16
// 16x16x32 form), so operands stream from L2 straight into registers with
// 16-byte loads and no LDS transpose.
//
// k_sgemm_tn<MODE>: C(m x n) = beta*C + alpha * A^T B, A (k x m) and
//   B (k x n) FP32, FP32 accumulation:
//   MODE_F32   v_mfma_f32_16x16x4_f32;
//   MODE_BF16  v_mfma_f32_16x16x32_bf16 on bf16(A), bf16(B);
//   MODE_BF16X3 the same instruction on a split x = hi + lo (hi = bf16(x),
//              lo = bf16(x - hi)): hi*hi + hi*lo + lo*hi, about 16 bits of
//              each product, three times the BF16 work.
//   The MFMA computes the transposed block D = B^T A (= C^T), so that a
//   lane's output values are consecutive rows of column-major C.
// k_tri_convert: out^T <- (Tout) in over one triangle (LDS-staged 64x64
//   blocks): the FP64 lower triangle of A to the FP32 upper working copy,
//   and the factor's diagonal tiles between FP32 upper and FP64 lower.
#include <stdlib.h>

#include <mutex>

#include "sa_common.h"
#include "starpu_amd_kernels.h"

namespace {

typedef __bf16 sa_bf8 __attribute__((ext_vector_type(8)));

enum { MODE_F32 = 0, MODE_BF16 = 1, MODE_BF16X3 = 2 };

template <int MODE>
struct Frag;

template <>
struct Frag<MODE_F32> {
	float v[8];
};
template <>
struct Frag<MODE_BF16> {
	sa_bf8 h;
};
template <>
struct Frag<MODE_BF16X3> {
	sa_bf8 h, l;
};

template <int MODE>
__device__ __forceinline__ Frag<MODE> make_frag(const float (&x)[8])
{
	Frag<MODE> f;
	if constexpr (MODE == MODE_F32) {
#pragma unroll
		for (int e = 0; e < 8; e++)
			f.v[e] = x[e];
	} else if constexpr (MODE == MODE_BF16) {
#pragma unroll
		for (int e = 0; e < 8; e++)
			f.h[e] = (__bf16)x[e];
	} else {
#pragma unroll
		for (int e = 0; e < 8; e++) {
			const __bf16 h = (__bf16)x[e];
			f.h[e] = h;
			f.l[e] = (__bf16)(x[e] - (float)h);
		}
	}
	return f;
}

// Lane group g holds k = 8g + e of a 32-wide k chunk in element e of both
// operand fragments; any fixed assignment of k to (g, e) gives the same
// sum, as long as the two operands agree.

// 8 consecutive k of column `col` starting at k0 + 8g; past K (or past the
// last column) the values are zero.  The fast path is two 16-byte loads.
template <bool EDGE>
__device__ __forceinline__ void load8(float (&v)[8], const float *__restrict__ X, long ld, int col, int ncols,
                                      int k, int K)
{
	if (!EDGE) {
		const sa_f4 *p = reinterpret_cast<const sa_f4 *>(X + k + (long)col * ld);
		const sa_f4 a = p[0], b = p[1];
		v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
		v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
	} else {
		const bool cin = col < ncols;
		const long cc = (long)(cin ? col : ncols - 1) * ld;
#pragma unroll
		for (int e = 0; e < 8; e++) {
			const bool in = cin && (k + e) < K;
			const float t = X[(in ? k + e : 0) + cc];
			v[e] = in ? t : 0.0f;
		}
	}
}

template <int MODE, int WB, bool UPPER, bool EDGE>
__global__ __launch_bounds__(256)
void k_sgemm_tn(int M, int N, int K, float alpha, const float *__restrict__ A, long lda,
                const float *__restrict__ B, long ldb, float beta, float *__restrict__ C, long ldc, int tiles_m)
{
	constexpr int WT = 16 * WB;      // wave tile
	constexpr int BT = 2 * WT;       // workgroup tile: 2 x 2 waves
	int tm, tn;
	if (UPPER) {
		// workgroup tiles on and above the diagonal, column by column
		const int t = blockIdx.x;
		int j = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
		while ((j + 1) * (j + 2) / 2 <= t) j++;
		while (j * (j + 1) / 2 > t) j--;
		tn = j;
		tm = t - j * (j + 1) / 2;
	} else {
		// XCD-grouped tile order, as the FP64 tile GEMM
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
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int i0 = tm * BT + (wave & 1) * WT;    // rows of C = columns of A
	const int j0 = tn * BT + (wave >> 1) * WT;   // columns of C = columns of B
	if (UPPER && i0 >= j0 + WT)
		return;                                   // wave tile strictly below the diagonal
	if (EDGE && (i0 >= M || j0 >= N))
		return;
	const int r = lane & 15, g = lane >> 4;

	sa_f4 acc[WB][WB];                           // acc[bj][bi]: block (C rows bi, C cols bj)
	float xa[WB][8], xb[WB][8], na[WB][8], nb[WB][8];
	auto load_chunk = [&](float (&fa)[WB][8], float (&fb)[WB][8], int k0) {
		const int k = k0 + 8 * g;
		const bool kin = k < K;
		const int kc = kin ? k : 0;
#pragma unroll
		for (int b = 0; b < WB; b++) {
			load8<EDGE>(fa[b], A, lda, i0 + 16 * b + r, M, kc, EDGE ? (kin ? K : 0) : K);
			load8<EDGE>(fb[b], B, ldb, j0 + 16 * b + r, N, kc, EDGE ? (kin ? K : 0) : K);
		}
	};
	auto compute_chunk = [&](const float (&fa)[WB][8], const float (&fb)[WB][8], bool first) {
		Frag<MODE> ya[WB], yb[WB];
#pragma unroll
		for (int b = 0; b < WB; b++) {
			ya[b] = make_frag<MODE>(fa[b]);
			yb[b] = make_frag<MODE>(fb[b]);
		}
		// one term at a time across all WB x WB independent accumulators, so
		// consecutive MFMAs never wait on each other's result
		if constexpr (MODE == MODE_F32) {
#pragma unroll
			for (int e = 0; e < 8; e++)
#pragma unroll
				for (int bj = 0; bj < WB; bj++)
#pragma unroll
					for (int bi = 0; bi < WB; bi++)
						acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x4f32(
						    yb[bj].v[e], ya[bi].v[e],
						    (first && e == 0) ? sa_f4{0.0f, 0.0f, 0.0f, 0.0f} : acc[bj][bi], 0, 0, 0);
		} else if constexpr (MODE == MODE_BF16) {
#pragma unroll
			for (int bj = 0; bj < WB; bj++)
#pragma unroll
				for (int bi = 0; bi < WB; bi++)
					acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
					    yb[bj].h, ya[bi].h, first ? sa_f4{0.0f, 0.0f, 0.0f, 0.0f} : acc[bj][bi], 0, 0, 0);
		} else {
#pragma unroll
			for (int bj = 0; bj < WB; bj++)
#pragma unroll
				for (int bi = 0; bi < WB; bi++)
					acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
					    yb[bj].l, ya[bi].h, first ? sa_f4{0.0f, 0.0f, 0.0f, 0.0f} : acc[bj][bi], 0, 0, 0);
#pragma unroll
			for (int bj = 0; bj < WB; bj++)
#pragma unroll
				for (int bi = 0; bi < WB; bi++)
					acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(yb[bj].h, ya[bi].l, acc[bj][bi], 0, 0, 0);
#pragma unroll
			for (int bj = 0; bj < WB; bj++)
#pragma unroll
				for (int bi = 0; bi < WB; bi++)
					acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(yb[bj].h, ya[bi].h, acc[bj][bi], 0, 0, 0);
		}
	};

	// two register stages: the next chunk loads while this one computes;
	// the first chunk is peeled so the accumulators start from zero
	load_chunk(xa, xb, 0);
	if (32 < K)
		load_chunk(na, nb, 32);
	compute_chunk(xa, xb, true);
	int k0 = 32;
	while (k0 < K) {
		if (k0 + 32 < K)
			load_chunk(xa, xb, k0 + 32);
		compute_chunk(na, nb, false);
		k0 += 32;
		if (k0 >= K)
			break;
		if (k0 + 32 < K)
			load_chunk(na, nb, k0 + 32);
		compute_chunk(xa, xb, false);
		k0 += 32;
	}

	// lane holds C(i0 + 16 bi + r, j0 + 16 bj + 4g + q) in acc[bj][bi][q].
	// All C values are loaded first (clamped addresses, no branches), so the
	// loads are in flight together; beta == 0 selects zero without reading
	// anything into the result.
	float cv[WB][4][WB];
#pragma unroll
	for (int bj = 0; bj < WB; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int bi = 0; bi < WB; bi++) {
				int jc = j0 + 16 * bj + 4 * g + q, ic = i0 + 16 * bi + r;
				if (EDGE) {
					jc = jc < N ? jc : N - 1;
					ic = ic < M ? ic : M - 1;
				}
				cv[bj][q][bi] = C[ic + (long)jc * ldc];
			}
#pragma unroll
	for (int bj = 0; bj < WB; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int jc = j0 + 16 * bj + 4 * g + q;
#pragma unroll
			for (int bi = 0; bi < WB; bi++) {
				const int ic = i0 + 16 * bi + r;
				const float c0 = beta == 0.0f ? 0.0f : beta * cv[bj][q][bi];
				const bool keep = !(EDGE && (jc >= N || ic >= M)) && !(UPPER && ic > jc);
				if (keep)
					C[ic + (long)jc * ldc] = c0 + alpha * acc[bj][bi][q];
			}
		}
}

// LDS-staged form for the BF16 modes on full 128x128 workgroup tiles: per
// 32-wide k chunk the workgroup loads its A and B panels (128 columns x 32
// k each) once with 16-byte loads, splits them into bf16 planes (hi, and
// lo for BF16x3) and stores them column by column in LDS (80-byte column
// stride, so the 16-byte fragment reads of a 16-lane group hit distinct
// banks); the four waves (2x2, 64x64 each) then read their MFMA fragments
// from LDS.  Each operand is thus fetched from L2 and converted once per
// workgroup instead of once per wave.  Two LDS buffers: the next chunk is
// loaded into registers while the MFMAs run on the current one, and
// stored into the other buffer; one barrier per chunk.
constexpr int LBT = 128;    // workgroup tile
constexpr int LCOL = 40;    // bf16 per staged column: 32 k + 8 padding

typedef __bf16 sa_bf4 __attribute__((ext_vector_type(4)));

template <int MODE, bool UPPER, int NBUF, int PF = 1>
__global__ __launch_bounds__(256)
void k_sgemm_tn_lds(int M, int N, int K, float alpha, const float *__restrict__ A, long lda,
                    const float *__restrict__ B, long ldb, float beta, float *__restrict__ C, long ldc, int tiles_m,
                    int ktri)
{
	constexpr int NPL = (MODE == MODE_BF16X3) ? 2 : 1;
	__shared__ __attribute__((aligned(16))) __bf16 Ls[NBUF][2][NPL][LBT * LCOL];
	int tm, tn;
	if (UPPER) {
		const int t = blockIdx.x;
		int j = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
		while ((j + 1) * (j + 2) / 2 <= t) j++;
		while (j * (j + 1) / 2 > t) j--;
		tn = j;
		tm = t - j * (j + 1) / 2;
	} else {
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
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int r = lane & 15, g = lane >> 4;
	const int I0 = tm * LBT, J0 = tn * LBT;
	const int i0w = (wave & 1) * 64, j0w = (wave >> 1) * 64;
	// on a diagonal tile of the upper form, the wave block strictly below
	// the diagonal takes part in loads and barriers but computes nothing.
	// (A run-time value in both forms: with compute() unconditional the
	// compiler interleaves it with the staging and runs out of registers.)
	const bool active = !(UPPER && I0 + i0w >= J0 + j0w + 64) && M > 0;
	// ktri: A (K x M) is upper triangular with zeros stored below its
	// diagonal, so output rows I0 .. I0+127 take only k < I0 + 128 (the
	// terms left out are exact zeros)
	if (ktri && I0 + LBT < K)
		K = I0 + LBT;

	// thread t loads columns (t >> 3) + 32u, k offset 4 (t & 7), u = 0..3
	const int lcol = tid >> 3, lk = (tid & 7) * 4;
	sa_f4 ra[4], rb[4], ra2[4], rb2[4];
	auto load_into = [&](int k0, sa_f4 (&xa)[4], sa_f4 (&xb)[4]) {
#pragma unroll
		for (int u = 0; u < 4; u++) {
			xa[u] = *reinterpret_cast<const sa_f4 *>(A + (k0 + lk) + (long)(I0 + lcol + 32 * u) * lda);
			xb[u] = *reinterpret_cast<const sa_f4 *>(B + (k0 + lk) + (long)(J0 + lcol + 32 * u) * ldb);
		}
	};
	auto load_stage = [&](int k0) { load_into(k0, ra, rb); };
	auto store_from = [&](int buf, const sa_f4 (&ra)[4], const sa_f4 (&rb)[4]) {
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int off = (lcol + 32 * u) * LCOL + lk;
			sa_bf4 ha, hb, la, lb;
#pragma unroll
			for (int e = 0; e < 4; e++) {
				const __bf16 xa = (__bf16)ra[u][e], xb = (__bf16)rb[u][e];
				ha[e] = xa;
				hb[e] = xb;
				if (NPL == 2) {
					la[e] = (__bf16)(ra[u][e] - (float)xa);
					lb[e] = (__bf16)(rb[u][e] - (float)xb);
				}
			}
			*reinterpret_cast<sa_bf4 *>(&Ls[buf][0][0][off]) = ha;
			*reinterpret_cast<sa_bf4 *>(&Ls[buf][1][0][off]) = hb;
			if (NPL == 2) {
				*reinterpret_cast<sa_bf4 *>(&Ls[buf][0][NPL - 1][off]) = la;
				*reinterpret_cast<sa_bf4 *>(&Ls[buf][1][NPL - 1][off]) = lb;
			}
		}
	};
	auto store_stage = [&](int buf) { store_from(buf, ra, rb); };
	sa_f4 acc[4][4];
#pragma unroll
	for (int a = 0; a < 4; a++)
#pragma unroll
		for (int b = 0; b < 4; b++)
			acc[a][b] = sa_f4{0.0f, 0.0f, 0.0f, 0.0f};
	auto compute = [&](int buf) {
		// the wave issuing MFMAs goes first while the other waves of the CU
		// convert and stage the next chunk (+4-8 % on 1024-wide products)
		__builtin_amdgcn_s_setprio(1);
		sa_bf8 ah[4], bh[4], al[4], bl[4];
#pragma unroll
		for (int b = 0; b < 4; b++) {
			const int oa = (i0w + 16 * b + r) * LCOL + 8 * g, ob = (j0w + 16 * b + r) * LCOL + 8 * g;
			ah[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][0][0][oa]);
			bh[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][1][0][ob]);
			if (NPL == 2) {
				al[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][0][NPL - 1][oa]);
				bl[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][1][NPL - 1][ob]);
			}
		}
		if (NPL == 2) {
#pragma unroll
			for (int bj = 0; bj < 4; bj++)
#pragma unroll
				for (int bi = 0; bi < 4; bi++)
					acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[bj], ah[bi], acc[bj][bi], 0, 0, 0);
#pragma unroll
			for (int bj = 0; bj < 4; bj++)
#pragma unroll
				for (int bi = 0; bi < 4; bi++)
					acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[bj], al[bi], acc[bj][bi], 0, 0, 0);
		}
#pragma unroll
		for (int bj = 0; bj < 4; bj++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[bj], ah[bi], acc[bj][bi], 0, 0, 0);
		__builtin_amdgcn_s_setprio(0);
	};

	if constexpr (NBUF == 2 && PF == 2) {
		// global loads two chunks ahead: chunk c + 2 is loaded into one
		// register stage while chunk c is computed and chunk c + 1 (loaded
		// a chunk earlier) is converted into the other LDS buffer, so a
		// round trip to L2 or HBM has two chunks' MFMAs to hide behind.
		// Two chunks per trip name the register stages statically.
		const int nch = K / 32;
		load_into(0, ra, rb);
		if (nch > 1)
			load_into(32, ra2, rb2);
		store_from(0, ra, rb);
		__syncthreads();
		for (int c = 0; c < nch; c += 2) {
			if (c + 2 < nch)
				load_into(32 * (c + 2), ra, rb);
			if (active)
				compute(0);
			if (c + 1 < nch)
				store_from(1, ra2, rb2);
			__syncthreads();
			if (c + 1 >= nch)
				break;
			if (c + 3 < nch)
				load_into(32 * (c + 3), ra2, rb2);
			if (active)
				compute(1);
			if (c + 2 < nch)
				store_from(0, ra, rb);
			__syncthreads();
		}
	} else if constexpr (NBUF == 2) {
		load_stage(0);
		store_stage(0);
		__syncthreads();
		for (int k0 = 0; k0 < K; k0 += 32) {
			const int cur = (k0 >> 5) & 1;
			const bool more = k0 + 32 < K;
			if (more)
				load_stage(k0 + 32);
			if (active)
				compute(cur);
			if (more)
				store_stage(cur ^ 1);
			__syncthreads();
		}
	} else {
		// one LDS buffer (half the LDS, twice the resident workgroups): the
		// next chunk's loads are in flight during the MFMAs, two barriers
		// per chunk
		load_stage(0);
		for (int k0 = 0; k0 < K; k0 += 32) {
			store_stage(0);
			__syncthreads();
			if (k0 + 32 < K)
				load_stage(k0 + 32);
			if (active)
				compute(0);
			__syncthreads();
		}
	}
	if (!active)
		return;
	// lane holds C(I0 + i0w + 16 bi + r, J0 + j0w + 16 bj + 4g + q) in acc[bj][bi][q]
	float cv[4][4][4];
#pragma unroll
	for (int bj = 0; bj < 4; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				cv[bj][q][bi] = C[(I0 + i0w + 16 * bi + r) + (long)(J0 + j0w + 16 * bj + 4 * g + q) * ldc];
#pragma unroll
	for (int bj = 0; bj < 4; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int jc = J0 + j0w + 16 * bj + 4 * g + q;
#pragma unroll
			for (int bi = 0; bi < 4; bi++) {
				const int ic = I0 + i0w + 16 * bi + r;
				const float c0 = beta == 0.0f ? 0.0f : beta * cv[bj][q][bi];
				if (!(UPPER && ic > jc))
					C[ic + (long)jc * ldc] = c0 + alpha * acc[bj][bi][q];
			}
		}
}

// LDS-staged form for FP32 (v_mfma_f32_16x16x4f32) on full 128x128
// workgroup tiles, the same scheme as the BF16 form: per 32-wide k chunk the
// workgroup loads its A and B panels (128 columns x 32 k) once with 16-byte
// loads and stores them column by column in LDS as FP32 (34-float column
// stride: the 16 lanes of a group read 16 columns at the same k, two floats
// apart in the banks); the four waves (2x2, 64x64 each) take their MFMA
// fragments from LDS, 8 k-steps of 4 per chunk.  Two LDS buffers (68 KB):
// the next chunk is loaded into registers while the MFMAs run on the current
// one; one barrier per chunk.
constexpr int LCOLF = 34;   // floats per staged column: 32 k + 2 padding

template <bool UPPER>
__global__ __launch_bounds__(256)
void k_sgemm_tn_lds_f32(int M, int N, int K, float alpha, const float *__restrict__ A, long lda,
                        const float *__restrict__ B, long ldb, float beta, float *__restrict__ C, long ldc,
                        int tiles_m, int ktri)
{
	__shared__ __attribute__((aligned(16))) float Lf[2][2][LBT * LCOLF];
	int tm, tn;
	if (UPPER) {
		const int t = blockIdx.x;
		int j = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
		while ((j + 1) * (j + 2) / 2 <= t) j++;
		while (j * (j + 1) / 2 > t) j--;
		tn = j;
		tm = t - j * (j + 1) / 2;
	} else {
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
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int r = lane & 15, g = lane >> 4;
	const int I0 = tm * LBT, J0 = tn * LBT;
	const int i0w = (wave & 1) * 64, j0w = (wave >> 1) * 64;
	const bool active = !(UPPER && I0 + i0w >= J0 + j0w + 64);
	// ktri: A (K x M) is upper triangular with zeros stored below its
	// diagonal, so output rows I0 .. I0+127 take only k < I0 + 128 (the
	// terms left out are exact zeros)
	if (ktri && I0 + LBT < K)
		K = I0 + LBT;

	// thread t loads columns (t >> 3) + 32u, k offset 4 (t & 7), u = 0..3
	const int lcol = tid >> 3, lk = (tid & 7) * 4;
	sa_f4 ra[4], rb[4];
	auto load_stage = [&](int k0) {
#pragma unroll
		for (int u = 0; u < 4; u++) {
			ra[u] = *reinterpret_cast<const sa_f4 *>(A + (k0 + lk) + (long)(I0 + lcol + 32 * u) * lda);
			rb[u] = *reinterpret_cast<const sa_f4 *>(B + (k0 + lk) + (long)(J0 + lcol + 32 * u) * ldb);
		}
	};
	auto store_stage = [&](int buf) {
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int off = (lcol + 32 * u) * LCOLF + lk;     // 8-byte aligned: two float2 stores
			float *pa = &Lf[buf][0][off], *pb = &Lf[buf][1][off];
			*reinterpret_cast<float2 *>(pa) = make_float2(ra[u][0], ra[u][1]);
			*reinterpret_cast<float2 *>(pa + 2) = make_float2(ra[u][2], ra[u][3]);
			*reinterpret_cast<float2 *>(pb) = make_float2(rb[u][0], rb[u][1]);
			*reinterpret_cast<float2 *>(pb + 2) = make_float2(rb[u][2], rb[u][3]);
		}
	};
	sa_f4 acc[4][4];
#pragma unroll
	for (int a = 0; a < 4; a++)
#pragma unroll
		for (int b = 0; b < 4; b++)
			acc[a][b] = sa_f4{0.0f, 0.0f, 0.0f, 0.0f};
	auto compute = [&](int buf) {
		__builtin_amdgcn_s_setprio(1);
#pragma unroll
		for (int s = 0; s < 8; s++) {
			float af[4], bf[4];
#pragma unroll
			for (int b = 0; b < 4; b++) {
				af[b] = Lf[buf][0][(i0w + 16 * b + r) * LCOLF + 4 * s + g];
				bf[b] = Lf[buf][1][(j0w + 16 * b + r) * LCOLF + 4 * s + g];
			}
#pragma unroll
			for (int bj = 0; bj < 4; bj++)
#pragma unroll
				for (int bi = 0; bi < 4; bi++)
					acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[bj], af[bi], acc[bj][bi], 0, 0, 0);
		}
		__builtin_amdgcn_s_setprio(0);
	};

	load_stage(0);
	store_stage(0);
	__syncthreads();
	for (int k0 = 0; k0 < K; k0 += 32) {
		const int cur = (k0 >> 5) & 1;
		const bool more = k0 + 32 < K;
		if (more)
			load_stage(k0 + 32);
		if (active)
			compute(cur);
		if (more)
			store_stage(cur ^ 1);
		__syncthreads();
	}
	if (!active)
		return;
	// lane holds C(I0 + i0w + 16 bi + r, J0 + j0w + 16 bj + 4g + q) in acc[bj][bi][q]
	float cv[4][4][4];
#pragma unroll
	for (int bj = 0; bj < 4; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				cv[bj][q][bi] = C[(I0 + i0w + 16 * bi + r) + (long)(J0 + j0w + 16 * bj + 4 * g + q) * ldc];
#pragma unroll
	for (int bj = 0; bj < 4; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int jc = J0 + j0w + 16 * bj + 4 * g + q;
#pragma unroll
			for (int bi = 0; bi < 4; bi++) {
				const int ic = I0 + i0w + 16 * bi + r;
				const float c0 = beta == 0.0f ? 0.0f : beta * cv[bj][q][bi];
				if (!(UPPER && ic > jc))
					C[ic + (long)jc * ldc] = c0 + alpha * acc[bj][bi][q];
			}
		}
}

// BF16x3 form on operands split ahead of time: A and B arrive as bf16
// planes hi = bf16(x), lo = bf16(x - hi), each laid out like the FP32
// matrix (k-contiguous columns, same leading dimension), as k_split_bf16x2
// writes them once per factor panel.  Same tiles, LDS layout, MFMA order and
// epilogue as k_sgemm_tn_lds<MODE_BF16X3>, so the result is bit-identical;
// the workgroup copies 16-byte runs of 8 k per plane into LDS instead of
// loading FP32 and splitting it, which removes the conversion from every
// product that reads the panel.
template <bool UPPER>
__global__ __launch_bounds__(256)
void k_sgemm_tn_lds_split(int M, int N, int K, float alpha, const __bf16 *__restrict__ Ah,
                          const __bf16 *__restrict__ Al, long lda, const __bf16 *__restrict__ Bh,
                          const __bf16 *__restrict__ Bl, long ldb, float beta, float *__restrict__ C, long ldc,
                          int tiles_m)
{
	__shared__ __attribute__((aligned(16))) __bf16 Ls[2][2][2][LBT * LCOL];
	int tm, tn;
	if (UPPER) {
		const int t = blockIdx.x;
		int j = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
		while ((j + 1) * (j + 2) / 2 <= t) j++;
		while (j * (j + 1) / 2 > t) j--;
		tn = j;
		tm = t - j * (j + 1) / 2;
	} else {
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
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int r = lane & 15, g = lane >> 4;
	const int I0 = tm * LBT, J0 = tn * LBT;
	const int i0w = (wave & 1) * 64, j0w = (wave >> 1) * 64;
	const bool active = !(UPPER && I0 + i0w >= J0 + j0w + 64);

	// thread t copies columns (t >> 2) + 64u, k offset 8 (t & 3), u = 0..1
	const int lcol = tid >> 2, lk = (tid & 3) * 8;
	sa_bf8 rah[2], ral[2], rbh[2], rbl[2];
	auto load_stage = [&](int k0) {
#pragma unroll
		for (int u = 0; u < 2; u++) {
			const long oa = (k0 + lk) + (long)(I0 + lcol + 64 * u) * lda;
			const long ob = (k0 + lk) + (long)(J0 + lcol + 64 * u) * ldb;
			rah[u] = *reinterpret_cast<const sa_bf8 *>(Ah + oa);
			ral[u] = *reinterpret_cast<const sa_bf8 *>(Al + oa);
			rbh[u] = *reinterpret_cast<const sa_bf8 *>(Bh + ob);
			rbl[u] = *reinterpret_cast<const sa_bf8 *>(Bl + ob);
		}
	};
	auto store_stage = [&](int buf) {
#pragma unroll
		for (int u = 0; u < 2; u++) {
			const int off = (lcol + 64 * u) * LCOL + lk;
			*reinterpret_cast<sa_bf8 *>(&Ls[buf][0][0][off]) = rah[u];
			*reinterpret_cast<sa_bf8 *>(&Ls[buf][0][1][off]) = ral[u];
			*reinterpret_cast<sa_bf8 *>(&Ls[buf][1][0][off]) = rbh[u];
			*reinterpret_cast<sa_bf8 *>(&Ls[buf][1][1][off]) = rbl[u];
		}
	};
	sa_f4 acc[4][4];
#pragma unroll
	for (int a = 0; a < 4; a++)
#pragma unroll
		for (int b = 0; b < 4; b++)
			acc[a][b] = sa_f4{0.0f, 0.0f, 0.0f, 0.0f};
	auto compute = [&](int buf) {
		// the wave issuing MFMAs goes first while the other waves of the CU
		// convert and stage the next chunk (+4-8 % on 1024-wide products)
		__builtin_amdgcn_s_setprio(1);
		sa_bf8 ah[4], bh[4], al[4], bl[4];
#pragma unroll
		for (int b = 0; b < 4; b++) {
			const int oa = (i0w + 16 * b + r) * LCOL + 8 * g, ob = (j0w + 16 * b + r) * LCOL + 8 * g;
			ah[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][0][0][oa]);
			bh[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][1][0][ob]);
			al[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][0][1][oa]);
			bl[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][1][1][ob]);
		}
#pragma unroll
		for (int bj = 0; bj < 4; bj++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[bj], ah[bi], acc[bj][bi], 0, 0, 0);
#pragma unroll
		for (int bj = 0; bj < 4; bj++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[bj], al[bi], acc[bj][bi], 0, 0, 0);
#pragma unroll
		for (int bj = 0; bj < 4; bj++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[bj], ah[bi], acc[bj][bi], 0, 0, 0);
		__builtin_amdgcn_s_setprio(0);
	};
	load_stage(0);
	store_stage(0);
	__syncthreads();
	for (int k0 = 0; k0 < K; k0 += 32) {
		const int cur = (k0 >> 5) & 1;
		const bool more = k0 + 32 < K;
		if (more)
			load_stage(k0 + 32);
		if (active)
			compute(cur);
		if (more)
			store_stage(cur ^ 1);
		__syncthreads();
	}
	if (!active)
		return;
	// lane holds C(I0 + i0w + 16 bi + r, J0 + j0w + 16 bj + 4g + q) in acc[bj][bi][q]
	float cv[4][4][4];
#pragma unroll
	for (int bj = 0; bj < 4; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				cv[bj][q][bi] = C[(I0 + i0w + 16 * bi + r) + (long)(J0 + j0w + 16 * bj + 4 * g + q) * ldc];
#pragma unroll
	for (int bj = 0; bj < 4; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int jc = J0 + j0w + 16 * bj + 4 * g + q;
#pragma unroll
			for (int bi = 0; bi < 4; bi++) {
				const int ic = I0 + i0w + 16 * bi + r;
				const float c0 = beta == 0.0f ? 0.0f : beta * cv[bj][q][bi];
				if (!(UPPER && ic > jc))
					C[ic + (long)jc * ldc] = c0 + alpha * acc[bj][bi][q];
			}
		}
}

// The split-plane product with the planes copied from memory straight into
// LDS (global_load_lds_dwordx4: no registers, no VALU on the way), so the
// staging of a chunk is 8 load instructions per wave that run on their own
// while the MFMAs of the previous chunk issue.  One such instruction writes
// 64 consecutive 16-byte LDS slots (1 KB); slot s of a plane's 128 x 32
// chunk holds unit u = (s % 4) ^ ((c >> 2) & 3) (8 k, 16 bytes) of column
// c = s / 4, so each lane loads the global run that belongs in its slot.
// The swizzle spreads the 16 columns one fragment read touches over the
// banks without padding (padding would break the 1 KB runs).  NST LDS
// stages of 32 KB (2: two workgroups per CU; 3 and 4: one): chunk c + NST - 1
// loads while chunk c is read; a wait on the wave's own loads and one
// barrier per chunk.  MFMA order and epilogue as k_sgemm_tn_lds_split:
// bit-identical results.  With 12 streams of 1024^3 products: 13.6-14.9 us
// at 2 stages, 13.9-14.1 us at 4, against 14.1 us through registers.
__device__ __forceinline__ int split_slot(int col, int u) { return col * 4 + (u ^ ((col >> 2) & 3)); }

template <bool UPPER, int NST = 2>
__global__ __launch_bounds__(256)
void k_sgemm_tn_lds_direct(int M, int N, int K, float alpha, const __bf16 *__restrict__ Ah,
                           const __bf16 *__restrict__ Al, long lda, const __bf16 *__restrict__ Bh,
                           const __bf16 *__restrict__ Bl, long ldb, float beta, float *__restrict__ C, long ldc,
                           int tiles_m)
{
	// [stage][operand A, B][plane hi, lo][128 columns x 32 k], NST x 32 KB of
	// dynamic LDS (more than the 64 KB a static array may take)
	extern __shared__ __attribute__((aligned(1024))) __bf16 lds_dyn[];
	typedef __bf16 stage_t[2][2][LBT * 32];
	stage_t *Ls = reinterpret_cast<stage_t *>(lds_dyn);
	int tm, tn;
	if (UPPER) {
		const int t = blockIdx.x;
		int j = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
		while ((j + 1) * (j + 2) / 2 <= t) j++;
		while (j * (j + 1) / 2 > t) j--;
		tn = j;
		tm = t - j * (j + 1) / 2;
	} else {
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
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int r = lane & 15, g = lane >> 4;
	const int I0 = tm * LBT, J0 = tn * LBT;
	const int i0w = (wave & 1) * 64, j0w = (wave >> 1) * 64;
	const bool active = !(UPPER && I0 + i0w >= J0 + j0w + 64);

	// wave w stages operand w >> 1, plane w & 1: eight 1 KB blocks q, slot
	// 64 q + lane = column 16 q + lane / 4, unit (lane % 4) ^ swizzle
	const int op = wave >> 1, pl = wave & 1;
	const __bf16 *src = op == 0 ? (pl == 0 ? Ah : Al) : (pl == 0 ? Bh : Bl);
	const long ld = op == 0 ? lda : ldb;
	const int base_col = op == 0 ? I0 : J0;
	const __bf16 *gsrc[8];
#pragma unroll
	for (int q = 0; q < 8; q++) {
		const int col = 16 * q + (lane >> 2);
		const int u = (lane & 3) ^ ((col >> 2) & 3);
		gsrc[q] = src + 8 * u + (long)(base_col + col) * ld;
	}
	auto stage = [&](int buf, int k0) {
#pragma unroll
		for (int q = 0; q < 8; q++)
			__builtin_amdgcn_global_load_lds(gsrc[q] + k0, &Ls[buf][op][pl][512 * q], 16, 0, 0);
	};
	sa_f4 acc[4][4];
#pragma unroll
	for (int a = 0; a < 4; a++)
#pragma unroll
		for (int b = 0; b < 4; b++)
			acc[a][b] = sa_f4{0.0f, 0.0f, 0.0f, 0.0f};
	auto compute = [&](int buf) {
		__builtin_amdgcn_s_setprio(1);
		sa_bf8 ah[4], bh[4], al[4], bl[4];
#pragma unroll
		for (int b = 0; b < 4; b++) {
			const int oa = split_slot(i0w + 16 * b + r, g) * 8, ob = split_slot(j0w + 16 * b + r, g) * 8;
			ah[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][0][0][oa]);
			bh[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][1][0][ob]);
			al[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][0][1][oa]);
			bl[b] = *reinterpret_cast<const sa_bf8 *>(&Ls[buf][1][1][ob]);
		}
#pragma unroll
		for (int bj = 0; bj < 4; bj++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[bj], ah[bi], acc[bj][bi], 0, 0, 0);
#pragma unroll
		for (int bj = 0; bj < 4; bj++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[bj], al[bi], acc[bj][bi], 0, 0, 0);
#pragma unroll
		for (int bj = 0; bj < 4; bj++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[bj], ah[bi], acc[bj][bi], 0, 0, 0);
		__builtin_amdgcn_s_setprio(0);
	};
	// chunk c goes to stage c % NST, NST - 1 chunks ahead of its use.  The LDS
	// written by a wave's loads is visible to the other waves once that wave
	// has waited for them (vmcnt counts its own outstanding loads, 8 per
	// chunk, oldest first) and passed the barrier; the barrier also tells
	// that every wave is done with the stage the next loads overwrite.
	const int nch = K / 32;
#pragma unroll
	for (int c = 0; c < NST - 1; c++)
		if (c < nch)
			stage(c, 32 * c);
	for (int c = 0; c < nch; c++) {
		if (c + NST - 2 < nch) {
			if constexpr (NST == 4)
				asm volatile("s_waitcnt vmcnt(16)" ::: "memory");
			else if constexpr (NST == 3)
				asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
			else
				asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
		} else {
			asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
		}
		__syncthreads();
		if (c + NST - 1 < nch)
			stage((c + NST - 1) % NST, 32 * (c + NST - 1));
		if (active)
			compute(c % NST);
	}
	if (!active)
		return;
	// lane holds C(I0 + i0w + 16 bi + r, J0 + j0w + 16 bj + 4g + q) in acc[bj][bi][q]
	float cv[4][4][4];
#pragma unroll
	for (int bj = 0; bj < 4; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int bi = 0; bi < 4; bi++)
				cv[bj][q][bi] = C[(I0 + i0w + 16 * bi + r) + (long)(J0 + j0w + 16 * bj + 4 * g + q) * ldc];
#pragma unroll
	for (int bj = 0; bj < 4; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int jc = J0 + j0w + 16 * bj + 4 * g + q;
#pragma unroll
			for (int bi = 0; bi < 4; bi++) {
				const int ic = I0 + i0w + 16 * bi + r;
				const float c0 = beta == 0.0f ? 0.0f : beta * cv[bj][q][bi];
				if (!(UPPER && ic > jc))
					C[ic + (long)jc * ldc] = c0 + alpha * acc[bj][bi][q];
			}
		}
}

// H(i, j) = bf16(A(i, j)), L(i, j) = bf16(A(i, j) - H(i, j)) over m x n, the
// split the BF16x3 kernels apply to their operands.
__global__ __launch_bounds__(256)
void k_split_bf16x2(int m, int n, const float *__restrict__ A, long lda, __bf16 *__restrict__ H,
                    __bf16 *__restrict__ L, long ldh)
{
	const long idx = (long)blockIdx.x * 256 + threadIdx.x;
	if (idx >= (long)m * n)
		return;
	const int i = (int)(idx % m), j = (int)(idx / m);
	const float x = A[i + (long)j * lda];
	const __bf16 h = (__bf16)x;
	H[i + (long)j * ldh] = h;
	L[i + (long)j * ldh] = (__bf16)(x - (float)h);
}

template <int MODE, int WB, bool UPPER>
int sgemm_launch(hipStream_t s, int m, int n, int k, float alpha, const float *A, int lda, const float *B,
                 int ldb, float beta, float *C, int ldc)
{
	constexpr int BT = 32 * WB;
	const int tiles_m = (m + BT - 1) / BT, tiles_n = (n + BT - 1) / BT;
	const unsigned grid = UPPER ? (unsigned)(tiles_n * (tiles_n + 1) / 2) : (unsigned)(tiles_m * tiles_n);
	const bool vec = sa_aligned16(A) && sa_aligned16(B) && (lda % 4 == 0) && (ldb % 4 == 0) && (k % 32 == 0);
	if (vec && m % BT == 0 && n % BT == 0)
		hipLaunchKernelGGL((k_sgemm_tn<MODE, WB, UPPER, false>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, A,
		                   (long)lda, B, (long)ldb, beta, C, (long)ldc, tiles_m);
	else
		hipLaunchKernelGGL((k_sgemm_tn<MODE, WB, UPPER, true>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, A,
		                   (long)lda, B, (long)ldb, beta, C, (long)ldc, tiles_m);
	return sa_launch_status();
}

// LDS buffers of the staged kernel: 2 (default: the next chunk is stored
// while this one computes, one barrier per chunk) or 1 (half the LDS, two
// barriers per chunk; measured slower on 1024-wide tiles, 96 against 109
// TFLOP/s); SA_SGEMM_LDS_BUFFERS=1 selects it, for tuning.
int lds_buffers()
{
	static int v = 0;
	if (v == 0) {
		const char *e = getenv("SA_SGEMM_LDS_BUFFERS");
		v = (e && atoi(e) == 1) ? 1 : 2;
	}
	return v;
}

// Chunks of k between a global load of the staged kernel and its use: 2
// (default) or 1 (SA_SGEMM_PREFETCH=1, for tuning).  Both give the same bits.
// (Three chunks, with a third register stage, measured the same speed.)
int prefetch_distance()
{
	static int v = 0;
	if (v == 0) {
		const char *e = getenv("SA_SGEMM_PREFETCH");
		v = (e && atoi(e) == 1) ? 1 : 2;
	}
	return v;
}

template <int MODE, bool UPPER>
int sgemm_shape(hipStream_t s, int m, int n, int k, float alpha, const float *A, int lda, const float *B,
                int ldb, float beta, float *C, int ldc, int ktri = 0)
{
	if (m % LBT == 0 && n % LBT == 0 && k % 32 == 0 && sa_aligned16(A) && sa_aligned16(B) && lda % 4 == 0 &&
	    ldb % 4 == 0) {
		const int tiles_m = m / LBT, tiles_n = n / LBT;
		const unsigned grid = UPPER ? (unsigned)(tiles_n * (tiles_n + 1) / 2) : (unsigned)(tiles_m * tiles_n);
		if constexpr (MODE == MODE_F32) {
			hipLaunchKernelGGL((k_sgemm_tn_lds_f32<UPPER>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, A, (long)lda,
			                   B, (long)ldb, beta, C, (long)ldc, tiles_m, ktri);
			return sa_launch_status();
		} else {
			if (lds_buffers() == 1)
				hipLaunchKernelGGL((k_sgemm_tn_lds<MODE, UPPER, 1>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, A,
				                   (long)lda, B, (long)ldb, beta, C, (long)ldc, tiles_m, ktri);
			else if (prefetch_distance() == 2)
				hipLaunchKernelGGL((k_sgemm_tn_lds<MODE, UPPER, 2, 2>), dim3(grid), dim3(256), 0, s, m, n, k, alpha,
				                   A, (long)lda, B, (long)ldb, beta, C, (long)ldc, tiles_m, ktri);
			else
				hipLaunchKernelGGL((k_sgemm_tn_lds<MODE, UPPER, 2>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, A,
				                   (long)lda, B, (long)ldb, beta, C, (long)ldc, tiles_m, ktri);
			return sa_launch_status();
		}
	}
	// 128x128 workgroups (64x64 per wave) once they give every CU a
	// workgroup, 64x64 below that
	const long big = (long)((m + 127) / 128) * ((n + 127) / 128);
	if (big >= 64)
		return sgemm_launch<MODE, 4, UPPER>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
	return sgemm_launch<MODE, 2, UPPER>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Block (r, c), r >= c, that is the t-th of the lower triangle of a grid of
// blocks taken row by row (row r starts at t = r (r + 1) / 2).
__device__ __forceinline__ void tri_block(unsigned t, int &r, int &c)
{
	int q = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
	while ((long)q * (q + 1) / 2 > (long)t)
		q--;
	while ((long)(q + 1) * (q + 2) / 2 <= (long)t)
		q++;
	r = q;
	c = (int)((long)t - (long)q * (q + 1) / 2);
}

// out(j, i) = (Tout) in(i, j) for the elements (i, j) of in's lower (LOWER)
// or upper triangle, n x n, through 64x64 LDS blocks: one workgroup per
// block of that triangle (nb (nb + 1) / 2 of them, nb = ceil(n / 64)).
// Whole blocks move as 16-byte vectors when the leading dimensions and base
// addresses allow it, edge blocks element by element.  The LDS copy holds
// the converted values.  With NORM (the FP64 lower triangle of a symmetric
// matrix) each block also leaves in part[bi + nb bj] its share of
// ||in||_F^2: its diagonal elements once, the elements below the diagonal
// twice (they stand for their mirror images), summed in a fixed order.
template <typename Tin, typename Tout, bool LOWER, bool NORM>
__device__ __forceinline__ void tri_convert_block(unsigned t, int n, int nb, const Tin *__restrict__ in, long ldi,
                                                  Tout *__restrict__ out, long ldo, double *__restrict__ part)
{
	constexpr int VI = 16 / sizeof(Tin), VO = 16 / sizeof(Tout);
	typedef Tin TinV __attribute__((ext_vector_type(VI)));
	typedef Tout ToutV __attribute__((ext_vector_type(VO)));
	__shared__ Tout S[64 * 65];
	__shared__ double red[4];
	int r, c;
	tri_block(t, r, c);
	const int bi = LOWER ? r : c, bj = LOWER ? c : r;     // block (bi, bj) of `in`
	const int r0 = 64 * bi, c0 = 64 * bj;
	const int tid = threadIdx.x;
	const Tin *src = in + r0 + (long)c0 * ldi;
	Tout *dst = out + c0 + (long)r0 * ldo;                // out block (bj, bi)
	const bool full = r0 + 64 <= n && c0 + 64 <= n;
	double ss = 0.0;
	if (full && ldi % VI == 0 && ((uintptr_t)src % 16) == 0) {
		TinV x[16 / VI];
#pragma unroll
		for (int q = 0; q < 16 / VI; q++) {
			const int v = tid + 256 * q, i = (v % (64 / VI)) * VI, j = v / (64 / VI);
			x[q] = *reinterpret_cast<const TinV *>(src + i + (long)j * ldi);
		}
#pragma unroll
		for (int q = 0; q < 16 / VI; q++) {
			const int v = tid + 256 * q, i = (v % (64 / VI)) * VI, j = v / (64 / VI);
#pragma unroll
			for (int e = 0; e < VI; e++) {
				S[i + e + 65 * j] = (Tout)x[q][e];
				if (NORM && r0 + i + e >= c0 + j)
					ss += (r0 + i + e == c0 + j ? 1.0 : 2.0) * (double)x[q][e] * (double)x[q][e];
			}
		}
	} else {
		for (int idx = tid; idx < 64 * 64; idx += 256) {
			const int i = idx & 63, j = idx >> 6;
			const int gi = r0 + i, gj = c0 + j;
			const Tin x = (gi < n && gj < n) ? src[i + (long)j * ldi] : (Tin)0;
			S[i + 65 * j] = (Tout)x;
			if (NORM && gi >= gj)
				ss += (gi == gj ? 1.0 : 2.0) * (double)x * (double)x;
		}
	}
	if constexpr (NORM) {
		for (int off = 32; off >= 1; off >>= 1)
			ss += __shfl_xor(ss, off);
		if ((tid & 63) == 0)
			red[tid >> 6] = ss;
	}
	__syncthreads();
	if (NORM && tid == 0)
		part[bi + (long)nb * bj] = (red[0] + red[1]) + (red[2] + red[3]);
	// out(c0 + i, r0 + j) = in(r0 + j, c0 + i) = S[j + 65 i], where the
	// source element lies in the converted triangle
	if (full && ldo % VO == 0 && ((uintptr_t)dst % 16) == 0) {
#pragma unroll
		for (int q = 0; q < 16 / VO; q++) {
			const int v = tid + 256 * q, i = (v % (64 / VO)) * VO, j = v / (64 / VO);
			ToutV y;
#pragma unroll
			for (int e = 0; e < VO; e++)
				y[e] = S[j + 65 * (i + e)];
			// off the diagonal every element is inside; on it (r0 = c0) the
			// vector is whole when its last (LOWER) or first element is
			if (bi != bj || (LOWER ? j >= i + VO - 1 : j <= i)) {
				*reinterpret_cast<ToutV *>(dst + i + (long)j * ldo) = y;
			} else {
#pragma unroll
				for (int e = 0; e < VO; e++)
					if (LOWER ? j >= i + e : j <= i + e)
						dst[i + e + (long)j * ldo] = y[e];
			}
		}
	} else {
		for (int idx = tid; idx < 64 * 64; idx += 256) {
			const int i = idx & 63, j = idx >> 6;        // out block row i, column j
			const int gi = c0 + i, gj = r0 + j;          // = in(gj, gi)
			const bool tri = LOWER ? gj >= gi : gj <= gi;
			if (gi < n && gj < n && tri)
				dst[i + (long)j * ldo] = S[j + 65 * i];
		}
	}
}

template <typename Tin, typename Tout, bool LOWER, bool NORM = false>
__global__ __launch_bounds__(256)
void k_tri_convert(int n, int nb, const Tin *__restrict__ in, long ldi, Tout *__restrict__ out, long ldo,
                   double *__restrict__ part = nullptr)
{
	tri_convert_block<Tin, Tout, LOWER, NORM>(blockIdx.x, n, nb, in, ldi, out, ldo, part);
}

// The FP32 outputs of a diagonal tile factored in FP64 in one launch: the
// first nb (nb + 1) / 2 workgroups write L's lower triangle transposed into
// U's upper one (as k_tri_convert), the others convert W (n x n) to W32,
// 256 elements each.
__global__ __launch_bounds__(256)
void k_tile_outputs_f32(int n, int nb, const double *__restrict__ L, long ldl, float *__restrict__ U, long ldu,
                        const double *__restrict__ W, long ldw, float *__restrict__ W32, long ldw32)
{
	const unsigned nbl = (unsigned)nb * (unsigned)(nb + 1) / 2;
	if (blockIdx.x < nbl) {
		tri_convert_block<double, float, true, false>(blockIdx.x, n, nb, L, ldl, U, ldu, nullptr);
		return;
	}
	const long idx = (long)(blockIdx.x - nbl) * 256 + threadIdx.x;
	if (idx >= (long)n * n)
		return;
	const int i = (int)(idx % n), j = (int)(idx / n);
	W32[i + (long)j * ldw32] = (float)W[i + (long)j * ldw];
}

// out[0] = the sum of part[bi + nb bj] over the blocks bi >= bj, in a fixed
// order (one workgroup).
__global__ __launch_bounds__(256)
void k_sum_lower_blocks(int nb, const double *__restrict__ part, double *__restrict__ out)
{
	__shared__ double red[256];
	const int tid = threadIdx.x;
	double s = 0.0;
	for (long idx = tid; idx < (long)nb * nb; idx += 256) {
		const int bi = (int)(idx % nb), bj = (int)(idx / nb);
		if (bi >= bj)
			s += part[idx];
	}
	red[tid] = s;
	__syncthreads();
	for (int h = 128; h > 0; h >>= 1) {
		if (tid < h)
			red[tid] += red[tid + h];
		__syncthreads();
	}
	if (tid == 0)
		out[0] = red[0];
}

__global__ void k_convert_d2s(int m, int n, const double *__restrict__ A, long lda, float *__restrict__ B, long ldb)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)m * n)
		return;
	const int i = (int)(idx % m), j = (int)(idx / m);
	B[i + (long)j * ldb] = (float)A[i + (long)j * lda];
}

// X(m x n) += alpha * D, FP64
__global__ void k_daxpy2d(int m, int n, double alpha, const double *__restrict__ D, long ldd, double *__restrict__ X,
                          long ldx)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)m * n)
		return;
	const int i = (int)(idx % m), j = (int)(idx / m);
	X[i + (long)j * ldx] += alpha * D[i + (long)j * ldd];
}

bool ldok(int ld, int rows) { return ld >= (rows > 1 ? rows : 1); }

} // namespace

extern "C" int sa_sgemm_tn(hipStream_t s, int mode, int upper, int m, int n, int k, float alpha, const float *A,
                           int lda, const float *B, int ldb, float beta, float *C, int ldc)
{
	if (m < 0 || n < 0 || k < 0 || !ldok(lda, k) || !ldok(ldb, k) || !ldok(ldc, m) || mode < 0 || mode > 2 ||
	    (upper && m != n))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	if (k == 0) {
		if (beta == 1.0f)
			return 0;
		return -EINVAL;     // no FP32 scaling kernel: callers always pass k > 0
	}
#define SA_SG(MODE)                                                                                   \
	return upper ? sgemm_shape<MODE, true>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)           \
	             : sgemm_shape<MODE, false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
	switch (mode) {
	case MODE_F32: SA_SG(MODE_F32);
	case MODE_BF16: SA_SG(MODE_BF16);
	default: SA_SG(MODE_BF16X3);
	}
#undef SA_SG
}

extern "C" int sa_sgemm_tn_utri(hipStream_t s, int mode, int m, int n, float alpha, const float *U, int ldu,
                                const float *B, int ldb, float beta, float *C, int ldc)
{
	if (m < 0 || n < 0 || !ldok(ldu, m) || !ldok(ldb, m) || !ldok(ldc, m) || mode < 0 || mode > 2)
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	switch (mode) {
	case MODE_F32: return sgemm_shape<MODE_F32, false>(s, m, n, m, alpha, U, ldu, B, ldb, beta, C, ldc, 1);
	case MODE_BF16: return sgemm_shape<MODE_BF16, false>(s, m, n, m, alpha, U, ldu, B, ldb, beta, C, ldc, 1);
	default: return sgemm_shape<MODE_BF16X3, false>(s, m, n, m, alpha, U, ldu, B, ldb, beta, C, ldc, 1);
	}
}

static int g_split_direct = 1;

template <bool U, int S>
void launch_direct(hipStream_t s, unsigned grid, int m, int n, int k, float alpha, const __bf16 *ah, const __bf16 *al,
                   int lda, const __bf16 *bh, const __bf16 *bl, int ldb, float beta, float *C, int ldc, int tiles_m)
{
	constexpr unsigned bytes = S * 2 * 2 * LBT * 32 * sizeof(__bf16);
	static std::once_flag once;
	std::call_once(once, [] {
		(void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sgemm_tn_lds_direct<U, S>),
		                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
	});
	hipLaunchKernelGGL((k_sgemm_tn_lds_direct<U, S>), dim3(grid), dim3(256), bytes, s, m, n, k, alpha, ah, al,
	                   (long)lda, bh, bl, (long)ldb, beta, C, (long)ldc, tiles_m);
}

extern "C" int sa_sgemm_tn_split_ok(int m, int n, int k, const void *Ah, const void *Al, int lda, const void *Bh,
                                    const void *Bl, int ldb)
{
	return m > 0 && n > 0 && k > 0 && m % LBT == 0 && n % LBT == 0 && k % 32 == 0 && lda % 8 == 0 &&
	       ldb % 8 == 0 && ldok(lda, k) && ldok(ldb, k) && sa_aligned16(Ah) && sa_aligned16(Al) &&
	       sa_aligned16(Bh) && sa_aligned16(Bl);
}

extern "C" int sa_sgemm_tn_split(hipStream_t s, int upper, int m, int n, int k, float alpha, const void *Ah,
                                 const void *Al, int lda, const void *Bh, const void *Bl, int ldb, float beta,
                                 float *C, int ldc)
{
	if (!sa_sgemm_tn_split_ok(m, n, k, Ah, Al, lda, Bh, Bl, ldb) || !ldok(ldc, m) || (upper && m != n))
		return -EINVAL;
	const int tiles_m = m / LBT, tiles_n = n / LBT;
	const unsigned grid = upper ? (unsigned)(tiles_n * (tiles_n + 1) / 2) : (unsigned)(tiles_m * tiles_n);
	const __bf16 *ah = (const __bf16 *)Ah, *al = (const __bf16 *)Al, *bh = (const __bf16 *)Bh,
	             *bl = (const __bf16 *)Bl;
	if (g_split_direct) {
		static int nst = -1;
		if (nst < 0) {
			const char *e = getenv("SA_SGEMM_DIRECT_STAGES");
			const int x = e ? atoi(e) : 2;
			nst = (x == 3 || x == 4) ? x : 2;
		}
		if (upper) {
			if (nst == 4)
				launch_direct<true, 4>(s, grid, m, n, k, alpha, ah, al, lda, bh, bl, ldb, beta, C, ldc, tiles_m);
			else if (nst == 3)
				launch_direct<true, 3>(s, grid, m, n, k, alpha, ah, al, lda, bh, bl, ldb, beta, C, ldc, tiles_m);
			else
				launch_direct<true, 2>(s, grid, m, n, k, alpha, ah, al, lda, bh, bl, ldb, beta, C, ldc, tiles_m);
		} else {
			if (nst == 4)
				launch_direct<false, 4>(s, grid, m, n, k, alpha, ah, al, lda, bh, bl, ldb, beta, C, ldc, tiles_m);
			else if (nst == 3)
				launch_direct<false, 3>(s, grid, m, n, k, alpha, ah, al, lda, bh, bl, ldb, beta, C, ldc, tiles_m);
			else
				launch_direct<false, 2>(s, grid, m, n, k, alpha, ah, al, lda, bh, bl, ldb, beta, C, ldc, tiles_m);
		}
	} else if (upper) {
		hipLaunchKernelGGL((k_sgemm_tn_lds_split<true>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, ah, al,
		                   (long)lda, bh, bl, (long)ldb, beta, C, (long)ldc, tiles_m);
	} else {
		hipLaunchKernelGGL((k_sgemm_tn_lds_split<false>), dim3(grid), dim3(256), 0, s, m, n, k, alpha, ah, al,
		                   (long)lda, bh, bl, (long)ldb, beta, C, (long)ldc, tiles_m);
	}
	return sa_launch_status();
}

// Staging of the split-plane product: straight to LDS (1, the default) or
// through registers (0).  Both give the same bits.  Returns the previous
// setting, -EINVAL for another value.
extern "C" int sa_set_sgemm_direct(int on)
{
	if (on != 0 && on != 1)
		return -EINVAL;
	const int prev = g_split_direct;
	g_split_direct = on;
	return prev;
}

extern "C" int sa_split_bf16x2(hipStream_t s, int m, int n, const float *A, int lda, void *H, void *L, int ldh)
{
	if (m < 0 || n < 0 || !ldok(lda, m) || !ldok(ldh, m))
		return -EINVAL;
	const long cnt = (long)m * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_split_bf16x2, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m, n, A, (long)lda,
	                   (__bf16 *)H, (__bf16 *)L, (long)ldh);
	return sa_launch_status();
}

extern "C" int sa_dlag2s_lower_to_upper(hipStream_t s, int n, const double *A, int lda, float *U, int ldu)
{
	if (n < 0 || !ldok(lda, n) || !ldok(ldu, n))
		return -EINVAL;
	if (n == 0)
		return 0;
	const unsigned nb = (unsigned)((n + 63) / 64);
	hipLaunchKernelGGL((k_tri_convert<double, float, true>), dim3(nb * (nb + 1) / 2), dim3(256), 0, s, n, (int)nb, A,
	                   (long)lda, U, (long)ldu);
	return sa_launch_status();
}

extern "C" size_t sa_dlag2s_lower_to_upper_fro_worksize(int n)
{
	const size_t nb = (size_t)(n > 0 ? (n + 63) / 64 : 1);
	return (nb * nb + 1) * sizeof(double);
}

extern "C" int sa_dlag2s_lower_to_upper_fro(hipStream_t s, int n, const double *A, int lda, float *U, int ldu,
                                            double *fro2, void *work, size_t work_bytes)
{
	if (n < 0 || !ldok(lda, n) || !ldok(ldu, n) || !fro2 || !work || work_bytes < sa_dlag2s_lower_to_upper_fro_worksize(n))
		return -EINVAL;
	const unsigned nb = (unsigned)((n + 63) / 64);
	double *part = (double *)work;
	if (n > 0) {
		hipLaunchKernelGGL((k_tri_convert<double, float, true, true>), dim3(nb * (nb + 1) / 2), dim3(256), 0, s, n,
		                   (int)nb, A, (long)lda, U, (long)ldu, part);
		int e = sa_launch_status();
		if (e)
			return e;
	}
	hipLaunchKernelGGL(k_sum_lower_blocks, dim3(1), dim3(256), 0, s, n > 0 ? (int)nb : 0, part, fro2);
	return sa_launch_status();
}

extern "C" int sa_slag2d_upper_to_lower(hipStream_t s, int n, const float *U, int ldu, double *L, int ldl)
{
	if (n < 0 || !ldok(ldu, n) || !ldok(ldl, n))
		return -EINVAL;
	if (n == 0)
		return 0;
	const unsigned nb = (unsigned)((n + 63) / 64);
	hipLaunchKernelGGL((k_tri_convert<float, double, false>), dim3(nb * (nb + 1) / 2), dim3(256), 0, s, n, (int)nb, U,
	                   (long)ldu, L, (long)ldl);
	return sa_launch_status();
}

extern "C" int sa_tile_outputs_f32(hipStream_t s, int n, const double *L, int ldl, float *U, int ldu, const double *W,
                                  int ldw, float *W32, int ldw32)
{
	if (n < 0 || !ldok(ldl, n) || !ldok(ldu, n) || !ldok(ldw, n) || !ldok(ldw32, n) || (n > 0 && (!L || !U || !W || !W32)))
		return -EINVAL;
	if (n == 0)
		return 0;
	const unsigned nb = (unsigned)((n + 63) / 64);
	const unsigned grid = nb * (nb + 1) / 2 + (unsigned)(((long)n * n + 255) / 256);
	hipLaunchKernelGGL(k_tile_outputs_f32, dim3(grid), dim3(256), 0, s, n, (int)nb, L, (long)ldl, U, (long)ldu, W,
	                   (long)ldw, W32, (long)ldw32);
	return sa_launch_status();
}

extern "C" int sa_dlag2s(hipStream_t s, int m, int n, const double *A, int lda, float *B, int ldb)
{
	if (m < 0 || n < 0 || !ldok(lda, m) || !ldok(ldb, m))
		return -EINVAL;
	const long cnt = (long)m * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_convert_d2s, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m, n, A, (long)lda, B,
	                   (long)ldb);
	return sa_launch_status();
}

extern "C" int sa_daxpy2d(hipStream_t s, int m, int n, double alpha, const double *D, int ldd, double *X, int ldx)
{
	if (m < 0 || n < 0 || !ldok(ldd, m) || !ldok(ldx, m))
		return -EINVAL;
	const long cnt = (long)m * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_daxpy2d, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m, n, alpha, D, (long)ldd,
	                   X, (long)ldx);
	return sa_launch_status();
}
