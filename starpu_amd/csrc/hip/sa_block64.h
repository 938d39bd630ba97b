// Device building blocks on 64x64 diagonal blocks of a lower Cholesky
// factor (gfx950): 16x16 register Cholesky and inverse, 16x16x16 MFMA
// block products on LDS operands, the factor-and-invert of one 64x64 block
// by a 256-thread workgroup, and the application of a block inverse to a
// panel.  Shared by the tile kernels (sa_chol.hip, sa_potrf.hip) and the
// batched persistent kernels (sa_batched.hip).
//
// Everything is a template on the element type T (double or float).  The
// FP32 and FP64 16x16x4 MFMAs read their operands alike (lane l supplies
// A(l%16, k) and B(k, l%16) for k = l/16 of the step) but return the 16x16
// result in different orders: register q of lane l holds row 4q + l/16
// (FP64) or row 4(l/16) + q (FP32) of column l%16 -- drow<T> below.  The
// double-named entry points at the end keep the FP64 interface of the
// callers.
#ifndef SA_BLOCK64_H
#define SA_BLOCK64_H

#include "sa_common.h"

namespace sa64 {

constexpr int LD = 65;      // LDS leading dimension of a 64x64 block (bank spread)

template <typename T> struct V4;
template <> struct V4<double> { typedef sa_d4 type; };
template <> struct V4<float> { typedef sa_f4 type; };

// v_mfma_f32_16x16x4f32: the lane layout of sa_mfma's FP64 form.
using ::sa_mfma;
__device__ __forceinline__ sa_f4 sa_mfma(float a, float b, sa_f4 c)
{
	return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of the 16x16 MFMA result held in register q of the lanes of group g
// (g = lane / 16; the column is lane % 16).
template <typename T> __device__ __forceinline__ constexpr int drow(int q, int g);
template <> __device__ __forceinline__ constexpr int drow<double>(int q, int g) { return 4 * q + g; }
template <> __device__ __forceinline__ constexpr int drow<float>(int q, int g) { return 4 * g + q; }

template <typename T>
__device__ __forceinline__ typename V4<T>::type zero4()
{
	return typename V4<T>::type{T(0), T(0), T(0), T(0)};
}

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

template <int K>
__device__ __forceinline__ float rb(float v)
{
	return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + K, 0xf, 0xf, false));
}

// v of lane `lane` (< 16) broadcast to the whole wave through SGPRs.
__device__ __forceinline__ double bcast_rl(double v, int lane)
{
	const long long bits = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), lane);
	const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
	return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float bcast_rl(float v, int lane)
{
	return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// k is a compile-time constant wherever this is called from fully unrolled
// loops, so the switch folds to one broadcast.
template <typename T>
__device__ __forceinline__ T bcast(T v, int k)
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
// (drow<T>(q, l/16), l%16).
template <typename T>
__device__ __forceinline__ typename V4<T>::type mm16_nt(const T *A, const T *B, typename V4<T>::type acc, T sign)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
	for (int s = 0; s < 4; s++)
		acc = sa_mfma(sign * A[r + (4 * s + g) * LD], B[r + (4 * s + g) * LD], acc);
	return acc;
}

template <typename T>
__device__ __forceinline__ typename V4<T>::type mm16_nn(const T *A, const T *B, typename V4<T>::type acc)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
	for (int s = 0; s < 4; s++)
		acc = sa_mfma(A[r + (4 * s + g) * LD], B[(4 * s + g) + r * LD], acc);
	return acc;
}

// acc += A * T with T a 16x16 block already in MFMA output layout: register
// s of lane l holds T(k, l%16) with k = drow<T>(s, l/16), so k-step s takes
// Tb[s] as its B operand and A(l%16, k) as its A operand (the 16 k values
// are covered once over the four steps, in a permuted order for FP32).
template <typename T>
__device__ __forceinline__ typename V4<T>::type mm16_nreg(const T *A, typename V4<T>::type Tb,
                                                          typename V4<T>::type acc)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
	for (int s = 0; s < 4; s++)
		acc = sa_mfma(A[r + drow<T>(s, g) * LD], Tb[s], acc);
	return acc;
}

template <typename T>
__device__ __forceinline__ void st16(T *P, typename V4<T>::type acc, T sign)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
	for (int q = 0; q < 4; q++)
		P[drow<T>(q, g) + r * LD] = sign * acc[q];
}

template <typename T>
__device__ __forceinline__ typename V4<T>::type ld16(const T *P)
{
	const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
	typename V4<T>::type acc;
#pragma unroll
	for (int q = 0; q < 4; q++)
		acc[q] = P[drow<T>(q, g) + r * LD];
	return acc;
}

// 1/sqrt(d) to full precision: the hardware estimate refined by Newton
// steps r <- r + r (1/2 - (d/2) r^2), each doubling the correct bits (two
// for double, one for float).
__device__ __forceinline__ double rsqrt_nr(double d)
{
	double r = __builtin_amdgcn_rsq(d);
	const double h = 0.5 * d;
	r = fma(r, fma(-h * r, r, 0.5), r);
	r = fma(r, fma(-h * r, r, 0.5), r);
	return r;
}

__device__ __forceinline__ float rsqrt_nr(float d)
{
	const float r = __builtin_amdgcn_rsqf(d);
	const float h = 0.5f * d;
	return fmaf(r, fmaf(-h * r, r, 0.5f), r);
}

// Lane t < 16 holds row t of a 16x16 lower block in x[0..15] (columns
// above the diagonal are ignored).  In-place Cholesky; returns the 1-based
// index of the first non-positive pivot, or 0.  Each column costs one
// reciprocal square root (L(k,k) = d * d^{-1/2}, multipliers scaled by
// d^{-1/2}) instead of a square root and a division.  The pivots and
// multipliers are broadcast with v_readlane: the row-broadcast DPP form gave
// a wrong factor here (a 17x17 input) although the same loop in isolation
// was correct, so this chain stays on the SGPR path.
template <typename T>
__device__ __forceinline__ int potrf16_regs(T (&x)[16])
{
	const int t = threadIdx.x & 63;
	int bad = 0;
#pragma unroll
	for (int k = 0; k < 16; k++) {
		const T d = bcast_rl(x[k], k);
		if (!(d > T(0)) && bad == 0)
			bad = k + 1;
		const T r = rsqrt_nr(d);
		x[k] = (t == k) ? d * r : x[k] * r;
#pragma unroll
		for (int j = k + 1; j < 16; j++)
			x[j] -= x[k] * bcast_rl(x[k], j);
	}
	return bad;
}

// 1 / L(t,t) for the row t = lane % 16 held in x: one division per lane, all
// lanes at once (off the dependency chain of the factorisation), so the
// inverse is the same whether the block was just factored or loaded.
template <typename T>
__device__ __forceinline__ T diag_recip(const T (&x)[16])
{
	T dg = T(1);
#pragma unroll
	for (int c = 0; c < 16; c++)
		dg = (c == (threadIdx.x & 15)) ? x[c] : dg;
	return T(1) / dg;
}

// Lane t < 16 holds row t of a factored 16x16 lower L and in rdiag the
// reciprocal of L(t,t); lane c gets column c of inv(L) in y:
// y_i = (delta_ic - sum_{q<i} L(i,q) y_q) / L(i,i).
template <typename T>
__device__ __forceinline__ void inv16_regs(const T (&x)[16], T rdiag, T (&y)[16])
{
	const int c = threadIdx.x & 63;
#pragma unroll
	for (int i = 0; i < 16; i++) {
		T s = (c == i) ? T(1) : T(0);
#pragma unroll
		for (int q = 0; q < i; q++)
			s -= bcast(x[q], i) * y[q];
		y[i] = s * bcast(rdiag, i);
	}
}

// Rows of the 16x16 diagonal block J of S into lanes 0..15 (lower part);
// other lanes get an identity row so that they run the same code.
template <typename T>
__device__ __forceinline__ void load_rows16(const T *S, int J, T (&x)[16])
{
	const int t = threadIdx.x & 63;
#pragma unroll
	for (int c = 0; c < 16; c++)
		x[c] = (t < 16) ? ((c <= t) ? S[(16 * J + t) + (16 * J + c) * LD] : T(0))
		                : ((c == (t & 15)) ? T(1) : T(0));
}

template <typename T>
__device__ __forceinline__ void store_inv16(T *Li, int J, const T (&y)[16])
{
	const int c = threadIdx.x & 63;
	if (c < 16)
#pragma unroll
		for (int i = 0; i < 16; i++)
			Li[(16 * J + i) + (16 * J + c) * LD] = y[i];
}

// FACTOR: Cholesky of the w x w (w <= 64) lower block D (ld ldd, in place)
// and its inverse into Dinv (64 x 64, ld 64, identity past w).
// !FACTOR: D is already factored; only the inverse is formed.
// One 256-thread workgroup; S and Li are two LDS blocks of 64 * LD elements.
// Returns, in every thread of wave 0, the 1-based column of the first
// non-positive pivot within the block (or 0); other threads return 0.
// Ends with a workgroup barrier, so the caller may reuse S and Li.
template <typename T, bool FACTOR>
__device__ int block_potrf_inv64_t(int w, T *__restrict__ D, long ldd, T *__restrict__ Dinv, T *__restrict__ S,
                                   T *__restrict__ Li)
{
	typedef typename V4<T>::type V;
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	// all 16 loads of a thread in flight before the first LDS store (the
	// block sits in L2; one round trip instead of sixteen)
	T v[16];
#pragma unroll
	for (int q = 0; q < 16; q++) {
		const int i = (tid + 256 * q) & 63, j = (tid + 256 * q) >> 6;
		const bool in = i < w && j < w && i >= j;
		const T x = D[(in ? i : 0) + (long)(in ? j : 0) * ldd];
		v[q] = in ? x : ((i == j && (i >= w || j >= w)) ? T(1) : T(0));
	}
#pragma unroll
	for (int q = 0; q < 16; q++) {
		const int i = (tid + 256 * q) & 63, j = (tid + 256 * q) >> 6;
		S[i + j * LD] = v[q];
		Li[i + j * LD] = T(0);
	}
	__syncthreads();

	int bad = 0;
	if (FACTOR) {
		for (int J = 0; J < 4; J++) {
			if (wave == 0) {
				T x[16], y[16];
				load_rows16(S, J, x);
				const int b = potrf16_regs(x);
				if (b && !bad)
					bad = 16 * J + b;
				if (lane < 16)
#pragma unroll
					for (int c = 0; c < 16; c++)
						if (c <= lane)
							S[(16 * J + lane) + (16 * J + c) * LD] = x[c];
				inv16_regs(x, diag_recip(x), y);
				store_inv16(Li, J, y);
			}
			__syncthreads();
			// panel: L_IJ = A_IJ * inv(L_JJ)^T
			if (wave < 3 - J) {
				const int I = J + 1 + wave;
				const V acc = mm16_nt(&S[16 * I + 16 * J * LD], &Li[16 * J + 16 * J * LD], zero4<T>(), T(1));
				st16(&S[16 * I + 16 * J * LD], acc, T(1));
			}
			__syncthreads();
			// trailing: A_IK -= L_IJ * L_KJ^T for J < K <= I < 4
			int p = 0;
			for (int I = J + 1; I < 4; I++)
				for (int K = J + 1; K <= I; K++, p++) {
					if ((p & 3) != wave)
						continue;
					V acc = ld16(&S[16 * I + 16 * K * LD]);
					acc = mm16_nt(&S[16 * I + 16 * J * LD], &S[16 * K + 16 * J * LD], acc, T(-1));
					st16(&S[16 * I + 16 * K * LD], acc, T(1));
				}
			__syncthreads();
		}
	} else {
		T x[16], y[16];
		load_rows16(S, wave, x);
		inv16_regs(x, diag_recip(x), y);
		store_inv16(Li, wave, y);
		__syncthreads();
	}
	// off-diagonal blocks of inv(L): Li_IJ = -Li_II * sum_{K=J}^{I-1} L_IK Li_KJ
	for (int d = 1; d < 4; d++) {
		if (wave < 4 - d) {
			const int J = wave, I = J + d;
			V Tb = zero4<T>();
			for (int K = J; K < I; K++)
				Tb = mm16_nn(&S[16 * I + 16 * K * LD], &Li[16 * K + 16 * J * LD], Tb);
			const V X = mm16_nreg(&Li[16 * I + 16 * I * LD], Tb, zero4<T>());
			st16(&Li[16 * I + 16 * J * LD], X, T(-1));
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
	__syncthreads();
	return bad;
}

// P (m x w, w <= 64) <- P * Dinv^T for the 16 rows row0 .. row0+15 of P,
// one wave (Dinv 64x64, ld 64, identity past w).  The wave reads its rows
// completely before writing them, so the update is in place.  With PT the
// result is also written transposed there (PT(c, r) = P(r, c), ld ldpt).
template <typename T>
__device__ __forceinline__ void rows_apply_inv_t(int row0, int m, int w, const T *__restrict__ Dinv,
                                                 T *__restrict__ P, long ldp, T *__restrict__ PT = nullptr,
                                                 long ldpt = 0)
{
	typedef typename V4<T>::type V;
	const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
	if (row0 >= m)
		return;
	const int row = row0 + r;
	const int rc = row < m ? row : m - 1;
	// operand loads in two batches (the panel rows with the first half of
	// Dinv, then the second half) rather than one round trip per MFMA
	T a[16], dv[2][16];
#pragma unroll
	for (int s = 0; s < 16; s++) {
		const int k = 4 * s + g;
		const T v = P[rc + (long)(k < w ? k : w - 1) * ldp];
		a[s] = k < w ? v : T(0);
	}
	V acc[4];
#pragma unroll
	for (int h = 0; h < 2; h++) {
#pragma unroll
		for (int c = 0; c < 2; c++)
#pragma unroll
			for (int s = 0; s < 16; s++)
				dv[c][s] = Dinv[(16 * (2 * h + c) + r) + (4 * s + g) * 64];
#pragma unroll
		for (int c = 0; c < 2; c++) {
			acc[2 * h + c] = zero4<T>();
#pragma unroll
			for (int s = 0; s < 16; s++)
				acc[2 * h + c] = sa_mfma(a[s], dv[c][s], acc[2 * h + c]);
		}
	}
	// acc[cb][q] = X(row0 + drow(q, g), 16cb + r)
#pragma unroll
	for (int cb = 0; cb < 4; cb++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int rr = row0 + drow<T>(q, g), cc = 16 * cb + r;
			if (rr < m && cc < w) {
				P[rr + (long)cc * ldp] = acc[cb][q];
				if (PT)
					PT[cc + (long)rr * ldpt] = acc[cb][q];
			}
		}
}

// FP64 entry points of the callers.
template <bool FACTOR>
__device__ __forceinline__ int block_potrf_inv64(int w, double *__restrict__ D, long ldd,
                                                 double *__restrict__ Dinv, double *__restrict__ S,
                                                 double *__restrict__ Li)
{
	return block_potrf_inv64_t<double, FACTOR>(w, D, ldd, Dinv, S, Li);
}

__device__ __forceinline__ void rows_apply_inv(int row0, int m, int w, const double *__restrict__ Dinv,
                                               double *__restrict__ P, long ldp)
{
	rows_apply_inv_t<double>(row0, m, w, Dinv, P, ldp);
}

} // namespace sa64

#endif
