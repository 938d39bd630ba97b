// Device building blocks on 64x64 diagonal blocks of a lower Cholesky
// factor (gfx950): 16x16 register Cholesky and inverse, 16x16x16 MFMA
// block products on LDS operands, the factor-and-invert of one 64x64 block
// by a 256-thread workgroup, and the application of a block inverse to a
// panel.  Shared by the tile kernels (sa_chol.hip) and the batched
// persistent kernel (sa_batched.hip).
#ifndef SA_BLOCK64_H
#define SA_BLOCK64_H

#include "sa_common.h"

namespace sa64 {


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

// 1/sqrt(d) to full double precision: the hardware estimate refined by two
// Newton steps r <- r + r (1/2 - (d/2) r^2) (each doubles the correct bits).
__device__ __forceinline__ double rsqrt_nr(double d)
{
	double r = __builtin_amdgcn_rsq(d);
	const double h = 0.5 * d;
	r = fma(r, fma(-h * r, r, 0.5), r);
	r = fma(r, fma(-h * r, r, 0.5), r);
	return r;
}

// Lane t < 16 holds row t of a 16x16 lower block in x[0..15] (columns
// above the diagonal are ignored).  In-place Cholesky; returns the 1-based
// index of the first non-positive pivot, or 0.  Each column costs one
// reciprocal square root (L(k,k) = d * d^{-1/2}, multipliers scaled by
// d^{-1/2}) instead of a square root and a division.  The pivots and
// multipliers are broadcast with v_readlane: the row-broadcast DPP form gave
// a wrong factor here (a 17x17 input) although the same loop in isolation
// was correct, so this chain stays on the SGPR path.
__device__ __forceinline__ int potrf16_regs(double (&x)[16])
{
	const int t = threadIdx.x & 63;
	int bad = 0;
#pragma unroll
	for (int k = 0; k < 16; k++) {
		const double d = bcast_rl(x[k], k);
		if (!(d > 0.0) && bad == 0)
			bad = k + 1;
		const double r = rsqrt_nr(d);
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
__device__ __forceinline__ double diag_recip(const double (&x)[16])
{
	double dg = 1.0;
#pragma unroll
	for (int c = 0; c < 16; c++)
		dg = (c == (threadIdx.x & 15)) ? x[c] : dg;
	return 1.0 / dg;
}

// Lane t < 16 holds row t of a factored 16x16 lower L and in rdiag the
// reciprocal of L(t,t); lane c gets column c of inv(L) in y:
// y_i = (delta_ic - sum_{q<i} L(i,q) y_q) / L(i,i).
__device__ __forceinline__ void inv16_regs(const double (&x)[16], double rdiag, double (&y)[16])
{
	const int c = threadIdx.x & 63;
#pragma unroll
	for (int i = 0; i < 16; i++) {
		double s = (c == i) ? 1.0 : 0.0;
#pragma unroll
		for (int q = 0; q < i; q++)
			s -= bcast(x[q], i) * y[q];
		y[i] = s * bcast(rdiag, i);
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

// FACTOR: Cholesky of the w x w (w <= 64) lower block D (ld ldd, in place)
// and its inverse into Dinv (64 x 64, ld 64, identity past w).
// !FACTOR: D is already factored; only the inverse is formed.
// One 256-thread workgroup; S and Li are two LDS blocks of 64 * LD doubles.
// Returns, in every thread of wave 0, the 1-based column of the first
// non-positive pivot within the block (or 0); other threads return 0.
// Ends with a workgroup barrier, so the caller may reuse S and Li.
template <bool FACTOR>
__device__ int block_potrf_inv64(int w, double *__restrict__ D, long ldd, double *__restrict__ Dinv,
                                 double *__restrict__ S, double *__restrict__ Li)
{
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	// all 16 loads of a thread in flight before the first LDS store (the
	// block sits in L2; one round trip instead of sixteen)
	double v[16];
#pragma unroll
	for (int q = 0; q < 16; q++) {
		const int i = (tid + 256 * q) & 63, j = (tid + 256 * q) >> 6;
		const bool in = i < w && j < w && i >= j;
		const double x = D[(in ? i : 0) + (long)(in ? j : 0) * ldd];
		v[q] = in ? x : ((i == j && (i >= w || j >= w)) ? 1.0 : 0.0);
	}
#pragma unroll
	for (int q = 0; q < 16; q++) {
		const int i = (tid + 256 * q) & 63, j = (tid + 256 * q) >> 6;
		S[i + j * LD] = v[q];
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
				inv16_regs(x, diag_recip(x), y);
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
		inv16_regs(x, diag_recip(x), y);
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
	__syncthreads();
	return bad;
}

// P (m x w, w <= 64) <- P * Dinv^T for the 16 rows row0 .. row0+15 of P,
// one wave (Dinv 64x64, ld 64, identity past w).  The wave reads its rows
// completely before writing them, so the update is in place.
__device__ __forceinline__ void rows_apply_inv(int row0, int m, int w, const double *__restrict__ Dinv,
                                               double *__restrict__ P, long ldp)
{
	const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
	if (row0 >= m)
		return;
	const int row = row0 + r;
	const int rc = row < m ? row : m - 1;
	// operand loads in two batches (the panel rows with the first half of
	// Dinv, then the second half) rather than one round trip per MFMA
	double a[16], dv[2][16];
#pragma unroll
	for (int s = 0; s < 16; s++) {
		const int k = 4 * s + g;
		const double v = P[rc + (long)(k < w ? k : w - 1) * ldp];
		a[s] = k < w ? v : 0.0;
	}
	sa_d4 acc[4];
#pragma unroll
	for (int h = 0; h < 2; h++) {
#pragma unroll
		for (int c = 0; c < 2; c++)
#pragma unroll
			for (int s = 0; s < 16; s++)
				dv[c][s] = Dinv[(16 * (2 * h + c) + r) + (4 * s + g) * 64];
#pragma unroll
		for (int c = 0; c < 2; c++) {
			acc[2 * h + c] = sa_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
			for (int s = 0; s < 16; s++)
				acc[2 * h + c] = sa_mfma(a[s], dv[c][s], acc[2 * h + c]);
		}
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

} // namespace sa64

#endif
