// Batched factor-and-solve of many independent small SPD systems of uneven
// sizes in one persistent kernel (gfx950).
//
// The grid is a fixed set of workgroups that stay resident; each pulls the
// next system from a device-side queue (an atomic counter over a list of
// systems, largest first, so the big ones start early and the small ones
// fill the tail) and processes it completely:
//   factor   left-looking over 64-column panels, in place in the system's
//            own column-major storage: panel -= L(panel rows, left) *
//            L(panel cols, left)^T (MFMA, 64x64 accumulator block per
//            wave, the panel's own rows staged in LDS -- each element of
//            the left part is read once per panel, where a right-looking
//            trailing update re-reads operands per tile and rewrites the
//            whole trailing triangle per panel), then factor-and-invert
//            the 64x64 diagonal block in LDS (block_potrf_inv64) and
//            panel <- panel * inv(L_jj)^T (MFMA, rows_apply_inv); the block
//            inverses are kept for the solve;
//   solve    L y = b then L^T x = y by 64-row blocks: partial products over
//            the solved part (coalesced column reads, 64 loads in flight
//            per lane, sums through LDS), then the block inverse.
// k_batched_posv_ir is the mixed-precision variant: the same factor and
// solve in FP32 on an FP32 copy of the matrix, wrapped in FP64 iterative
// refinement against the untouched FP64 matrix (the scheme of LAPACK
// dsposv); it runs two workgroups per CU.
// Every workgroup leaves the loop when the queue is empty, so the grid
// drains before the kernel ends.  k_queue_reset zeroes the counter with a
// kernel (not a memset), so the pair can be captured in a hipGraph and
// replayed.
#include "sa_block64.h"
#include "starpu_amd_kernels.h"

namespace {
using namespace sa64;

// Left-looking update of the 64-column panel at column j0 (width w <= 64)
// of one system (n x n, ld n), by one 256-thread workgroup:
//   A(j0:n, j0:j0+w) -= L(j0:n, 0:j0) L(j0:j0+w, 0:j0)^T
// (lower part of the diagonal block only).  Rows in groups of 64 RB, wave
// v taking the 16 RB rows from 16 RB v (RB = 2 or 4 for FP64, chosen by the
// autotuner; 2 for FP32: with 4 the FP32 version spills within the 256
// VGPRs of two workgroups per CU and runs 1.6x slower); K = j0 in chunks of
// 64, the chunk of the panel's own rows L(j0:j0+w, kc:kc+64) staged in LDS
// (Bs, 64 x LD) and shared by the four waves, the wave's rows streamed from
// memory once per panel.  Each wave keeps a (16 RB)x64 block of MFMA
// accumulators and computes the transposed block, so a lane's results are
// consecutive rows.  The next chunk's staging loads and row operands are in
// flight during the current chunk's MFMAs.
template <typename T, int RB>
__device__ void wg_panel_update(int n, int j0, int w, T *__restrict__ A, T *__restrict__ Bs)
{
	typedef typename V4<T>::type V;
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r = lane & 15, g = lane >> 4;
	for (int rg = j0; rg < n; rg += 64 * RB) {
		const int row0 = rg + 16 * RB * wave;
		const bool active = row0 < n;
		V acc[RB][4];
#pragma unroll
		for (int a = 0; a < RB; a++)
#pragma unroll
			for (int b = 0; b < 4; b++)
				acc[a][b] = zero4<T>();
		// operands of the wave's rows for one chunk: bo[s][a] = L(row, kc + 4s + g)
		int rc[RB];
#pragma unroll
		for (int a = 0; a < RB; a++)
			rc[a] = (row0 + 16 * a + r) < n ? row0 + 16 * a + r : n - 1;
		T nxt[16], bo[16][RB];
#pragma unroll
		for (int q = 0; q < 16; q++) {
			const int c = (tid + 256 * q) & 63, k = (tid + 256 * q) >> 6;
			nxt[q] = c < w ? A[(j0 + c) + (long)k * n] : T(0);
		}
		if (active)
#pragma unroll
			for (int s = 0; s < 16; s++)
#pragma unroll
				for (int a = 0; a < RB; a++)
					bo[s][a] = A[rc[a] + (long)(4 * s + g) * n];
		for (int kc = 0; kc < j0; kc += 64) {
			__syncthreads();                 // the previous chunk's readers are done
#pragma unroll
			for (int q = 0; q < 16; q++) {
				const int c = (tid + 256 * q) & 63, k = (tid + 256 * q) >> 6;
				Bs[c + k * LD] = nxt[q];
			}
			__syncthreads();
			const bool more = kc + 64 < j0;
			if (more)
#pragma unroll
				for (int q = 0; q < 16; q++) {
					const int c = (tid + 256 * q) & 63, k = (tid + 256 * q) >> 6;
					nxt[q] = c < w ? A[(j0 + c) + (long)(kc + 64 + k) * n] : T(0);
				}
			if (active) {
				// the next chunk's row operands load into the registers each
				// k-step frees, while this chunk's MFMAs run
#pragma unroll
				for (int s = 0; s < 16; s++) {
					const int k = 4 * s + g;
					T ao[4];
#pragma unroll
					for (int b = 0; b < 4; b++)
						ao[b] = Bs[(16 * b + r) + k * LD];
#pragma unroll
					for (int a = 0; a < RB; a++)
#pragma unroll
						for (int b = 0; b < 4; b++)
							acc[a][b] = sa_mfma(ao[b], bo[s][a], acc[a][b]);
					if (more)
#pragma unroll
						for (int a = 0; a < RB; a++)
							bo[s][a] = A[rc[a] + (long)(kc + 64 + k) * n];
				}
			}
		}
		// acc[a][b][q] = (L L^T)(row0 + 16a + r, j0 + 16b + drow(q, g))
		if (active)
#pragma unroll
			for (int a = 0; a < RB; a++)
#pragma unroll
				for (int b = 0; b < 4; b++)
#pragma unroll
					for (int q = 0; q < 4; q++) {
						const int row = row0 + 16 * a + r, c = 16 * b + drow<T>(q, g);
						if (row < n && c < w && row - j0 >= c)
							A[row + (long)(j0 + c) * n] -= acc[a][b][q];
					}
	}
}

// Sum over the 64 lanes of each of a[0..15]: a halving butterfly (8 + 4 + 2
// + 1 exchanges, then two plain ones) instead of sixteen full reductions.
// Lane l returns the sum of a[c] for c = col16(l) = 8 b5 + 4 b4 + 2 b3 + b2
// (bk: bit k of l); the four lanes that differ in bits 0-1 agree.
__device__ __forceinline__ int col16(int lane)
{
	return ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
}

template <typename T>
__device__ __forceinline__ T reduce16(const T (&a)[16], int lane)
{
	T b[8], c[4], d[2];
	const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4;
#pragma unroll
	for (int i = 0; i < 8; i++) {
		const T keep = h5 ? a[8 + i] : a[i], send = h5 ? a[i] : a[8 + i];
		b[i] = keep + __shfl_xor(send, 32);
	}
#pragma unroll
	for (int i = 0; i < 4; i++) {
		const T keep = h4 ? b[4 + i] : b[i], send = h4 ? b[i] : b[4 + i];
		c[i] = keep + __shfl_xor(send, 16);
	}
#pragma unroll
	for (int i = 0; i < 2; i++) {
		const T keep = h3 ? c[2 + i] : c[i], send = h3 ? c[i] : c[2 + i];
		d[i] = keep + __shfl_xor(send, 8);
	}
	T e = (h2 ? d[1] : d[0]) + __shfl_xor(h2 ? d[0] : d[1], 4);
	e += __shfl_xor(e, 2);
	e += __shfl_xor(e, 1);
	return e;
}

// In-place solve of L L^T x = b for one right-hand side of length n
// (x overwrites b), L lower with its 64x64 block inverses Dinv (block j at
// Dinv + 4096 j, ld 64, identity past the block's width), all in precision
// T.  One 256-thread workgroup; Tl is LDS scratch of 512 elements.  The
// sweeps are bound by memory latency, not arithmetic, so each 64-row block
// costs about one round trip to memory: the block's inverse is fetched
// together with the panel products, and x lives in LDS (wg_solve1, for
// n <= 1024), so the only global traffic is L and Dinv.
//   forward, block j:  t = b_j - L(j, 0:j0) y(0:j0)  (lane = row, the waves
//                      split the columns, 64 loads in flight per lane), then
//                      y_j = Dinv_j t  (wave v: columns 16v..16v+15 of Dinv);
//   backward, block j: t = y_j - L(j0+w:n, j)^T x(j0+w:n)  (wave v: the 16
//                      columns 16v.. of the panel at once, lanes down four
//                      64-row chunks), then x_j = Dinv_j^T t  (a column of
//                      Dinv per dot, coalesced).
template <typename T>
__device__ __forceinline__ void solve_core(int n, const T *__restrict__ L, long ldl, const T *__restrict__ Dinv,
                                           T *__restrict__ x, T *__restrict__ Tl)
{
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	T *part = Tl, *part2 = Tl + 256;
	const int nblk = (n + 63) / 64;
	for (int jb = 0; jb < nblk; jb++) {
		const int j0 = 64 * jb, w = (n - j0) < 64 ? (n - j0) : 64;
		const int rowc = j0 + (lane < w ? lane : 0);
		const T *Dj = Dinv + 4096L * jb;
		T dv[16];
#pragma unroll
		for (int qq = 0; qq < 16; qq++)
			dv[qq] = Dj[lane + 64 * (16 * wave + qq)];
		// wave v takes the columns [kv0, kv1) of the solved part, whole
		// chunks of 64 (j0 is a multiple of 64), x of the chunk in one load,
		// broadcast by v_readlane
		const int qc = ((j0 + 255) / 256) * 64;
		const int kv0 = wave * qc < j0 ? wave * qc : j0;
		const int kv1 = kv0 + qc < j0 ? kv0 + qc : j0;
		T sacc = T(0);
#pragma unroll 1
		for (int k = kv0; k < kv1; k += 64) {
			const T xv = x[k + lane];
			T a[64];
#pragma unroll
			for (int u = 0; u < 64; u++)
				a[u] = L[rowc + (long)(k + u) * ldl];
#pragma unroll
			for (int u = 0; u < 64; u++)
				sacc += a[u] * bcast_rl(xv, u);
		}
		part[wave * 64 + lane] = sacc;
		__syncthreads();
		T y = T(0);
#pragma unroll
		for (int qq = 0; qq < 16; qq++) {
			const int q = 16 * wave + qq;
			const T tq = q < w ? x[j0 + q] - ((part[q] + part[64 + q]) + (part[128 + q] + part[192 + q])) : T(0);
			y += dv[qq] * tq;
		}
		part2[wave * 64 + lane] = y;
		__syncthreads();
		if (wave == 0 && lane < w)
			x[j0 + lane] = (part2[lane] + part2[64 + lane]) + (part2[128 + lane] + part2[192 + lane]);
		__syncthreads();
	}
	for (int jb = nblk - 1; jb >= 0; jb--) {
		const int j0 = 64 * jb, w = (n - j0) < 64 ? (n - j0) : 64;
		const int kb = j0 + w;
		const int c0 = 16 * wave;
		const T *Dj = Dinv + 4096L * jb;
		T dv[16];
#pragma unroll
		for (int cc = 0; cc < 16; cc++)
			dv[cc] = Dj[lane + 64 * (c0 + cc)];
		T acc[16];
#pragma unroll
		for (int cc = 0; cc < 16; cc++)
			acc[cc] = T(0);
		// four 64-row chunks per step: 64 loads in flight per lane
#pragma unroll 1
		for (int k = kb + lane; k < n; k += 256) {
			T xk[4], a[4][16];
#pragma unroll
			for (int h = 0; h < 4; h++)
				xk[h] = k + 64 * h < n ? x[k + 64 * h] : T(0);
#pragma unroll
			for (int cc = 0; cc < 16; cc++) {
				const long col = (long)(j0 + (c0 + cc < w ? c0 + cc : 0)) * ldl;
#pragma unroll
				for (int h = 0; h < 4; h++)
					a[h][cc] = k + 64 * h < n ? L[k + 64 * h + col] : T(0);
			}
#pragma unroll
			for (int cc = 0; cc < 16; cc++)
				if (c0 + cc < w)
					acc[cc] += (a[0][cc] * xk[0] + a[1][cc] * xk[1]) + (a[2][cc] * xk[2] + a[3][cc] * xk[3]);
		}
		const int col = c0 + col16(lane);
		T s = reduce16(acc, lane);
		if ((lane & 3) == 0)
			part[col] = col < w ? x[j0 + col] - s : T(0);
		__syncthreads();
		const T tl = part[lane];
#pragma unroll
		for (int cc = 0; cc < 16; cc++)
			acc[cc] = dv[cc] * tl;
		s = reduce16(acc, lane);
		if ((lane & 3) == 0 && col < w)
			x[j0 + col] = s;
		__syncthreads();
	}
}

// solve_core with x held in LDS for n <= 1024 (one copy in, one copy out),
// in global memory beyond.
template <typename T>
__device__ __noinline__ void wg_solve1(int n, const T *__restrict__ L, long ldl, const T *__restrict__ Dinv,
                                       T *__restrict__ x, T *__restrict__ Tl)
{
	__shared__ T xs[1024];
	if (n <= 1024) {
		for (int i = threadIdx.x; i < n; i += 256)
			xs[i] = x[i];
		__syncthreads();
		solve_core<T>(n, L, ldl, Dinv, xs, Tl);
		for (int i = threadIdx.x; i < n; i += 256)
			x[i] = xs[i];
		__syncthreads();
	} else {
		solve_core<T>(n, L, ldl, Dinv, x, Tl);
	}
}

// Out-of-line steps of the factorisation: each gets its own register
// allocation, which keeps the FP32 refinement kernel within the 256 VGPRs
// of two resident workgroups per CU and the FP64 panel update free of
// spills.  (The FP64 kernel inlines the block factor and the panel apply:
// the block factor alone needs ~360 VGPRs.)
template <typename T>
__device__ __noinline__ int nb_block_potrf(int w, T *D, long ldd, T *Dinv, T *S, T *Li)
{
	return block_potrf_inv64_t<T, true>(w, D, ldd, Dinv, S, Li);
}

template <typename T>
__device__ __noinline__ void nb_rows_apply(int rows, int w, const T *Dj, T *Pn, long ldp)
{
	for (int row0 = (threadIdx.x >> 6) * 16; row0 < rows; row0 += 64)
		rows_apply_inv_t<T>(row0, rows, w, Dj, Pn, ldp);
}

template <typename T, int RB>
__device__ __noinline__ void nb_panel_update(int n, int j0, int w, T *A, T *Bs)
{
	wg_panel_update<T, RB>(n, j0, w, A, Bs);
}

// Left-looking factorisation of one n x n system (lower, ld n, in place)
// by the workgroup, 64 columns at a time: update the panel with all the
// columns to its left, factor-and-invert its diagonal block, apply the
// inverse to the rows below; block inverses into Di.  Returns in thread 0
// the 1-based column of the first non-positive pivot (or 0); *s_bad is LDS.
template <typename T, int RB = 2>
__device__ int wg_factor(int n, T *__restrict__ Ai, T *__restrict__ Di, T *__restrict__ S, T *__restrict__ Li,
                         int *s_bad)
{
	if (threadIdx.x == 0)
		*s_bad = 0;
	for (int j0 = 0; j0 < n; j0 += 64) {
		const int w = (n - j0) < 64 ? (n - j0) : 64;
		T *Dj = Di + 4096L * (j0 / 64);
		if (j0 > 0) {
			nb_panel_update<T, RB>(n, j0, w, Ai, S);
			__syncthreads();
		}
		int bad;
		if constexpr (sizeof(T) == 4)
			bad = nb_block_potrf<T>(w, Ai + j0 + (long)j0 * n, n, Dj, S, Li);
		else
			bad = block_potrf_inv64_t<T, true>(w, Ai + j0 + (long)j0 * n, n, Dj, S, Li);
		if (threadIdx.x == 0 && bad && bad <= w && *s_bad == 0)
			*s_bad = j0 + bad;
		const int rows = n - j0 - w;
		if (rows > 0) {
			T *Pn = Ai + (j0 + w) + (long)j0 * n;
			if constexpr (sizeof(T) == 4)
				nb_rows_apply<T>(rows, w, Dj, Pn, n);
			else
				for (int row0 = (threadIdx.x >> 6) * 16; row0 < rows; row0 += 64)
					rows_apply_inv_t<T>(row0, rows, w, Dj, Pn, n);
			__syncthreads();
		}
	}
	__syncthreads();
	return *s_bad;
}

// Sum of v over the workgroup, returned to every thread in the same order
// of additions (red4: 4 doubles of LDS).
__device__ double wg_sum(double v, double *red4)
{
	for (int off = 32; off >= 1; off >>= 1)
		v += __shfl_xor(v, off);
	__syncthreads();
	if ((threadIdx.x & 63) == 0)
		red4[threadIdx.x >> 6] = v;
	__syncthreads();
	return (red4[0] + red4[1]) + (red4[2] + red4[3]);
}

// One row chunk [r0, r1) (at most 64 NS rows) of wg_residual: lane l owns
// rows r0 + 64 m + l (m < NS); wave v walks the chunk's columns j = v,
// v + 4, ..., NC of them per step.  NS * NC = 32 loads are in flight per
// lane whatever the chunk's height.
template <int NS, int NC>
__device__ __noinline__ void residual_chunk(int r0, int r1, const double *__restrict__ A, long lda,
                                               const double *__restrict__ x, double *__restrict__ r,
                                               double *red)
{
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	double acc[NS], xr[NS];
#pragma unroll
	for (int m = 0; m < NS; m++) {
		const int i = r0 + 64 * m + lane;
		xr[m] = i < r1 ? x[i] : 0.0;
		acc[m] = 0.0;
	}
	const int nm = (r1 - r0 + 63) / 64;       // row slots in use
	for (int j = wave; j < r1; j += 4 * NC) {
		double dot[NC];
#pragma unroll
		for (int u = 0; u < NC; u++) {
			const int jj = j + 4 * u;
			const bool valid = jj < r1;
			const double xj = valid ? x[jj] : 0.0;
			const double *Aj = A + (long)(valid ? jj : 0) * lda;
			const int mlo = jj > r0 ? (jj - r0) / 64 : 0;   // slots wholly above the diagonal
			dot[u] = 0.0;
			// loads independent of the lanes' masks (rows clamped into the
			// stored matrix; entries above the diagonal are read and
			// discarded by the selects), only the slot range is a
			// wave-uniform condition, so all loads are in flight before the
			// first use
			double a[NS];
#pragma unroll
			for (int m = 0; m < NS; m++) {
				const int i = r0 + 64 * m + lane;
				a[m] = (m >= mlo && m < nm) ? Aj[i < r1 ? i : r1 - 1] : 0.0;
			}
#pragma unroll
			for (int m = 0; m < NS; m++) {
				const int i = r0 + 64 * m + lane;
				const bool lo = valid && i >= jj && i < r1;
				acc[m] = lo ? fma(a[m], xj, acc[m]) : acc[m];
				dot[u] = (lo && i > jj) ? fma(a[m], xr[m], dot[u]) : dot[u];
			}
		}
#pragma unroll
		for (int u = 0; u < NC; u++)
			for (int off = 32; off >= 1; off >>= 1)
				dot[u] += __shfl_xor(dot[u], off);
		if (lane == 0) {
#pragma unroll
			for (int u = 0; u < NC; u++)
				if (j + 4 * u < r1)
					r[j + 4 * u] -= dot[u];
		}
	}
#pragma unroll
	for (int m = 0; m < NS; m++)
		red[wave * 1024 + 64 * m + lane] = acc[m];
	__syncthreads();
	for (int i = r0 + tid; i < r1; i += 256) {
		const int q = i - r0;
		r[i] -= (red[q] + red[1024 + q]) + (red[2048 + q] + red[3072 + q]);
	}
	__syncthreads();
}

// r <- b - A x in FP64 for the symmetric A of order n given by its lower
// triangle (ld lda), one 256-thread workgroup; returns ||r||^2 to every
// thread.  The lower triangle is read once, by row chunks of up to 1024:
// in a chunk, lanes own rows and the waves walk the columns, adding A(i,j)
// x_j to their row accumulators (the lower product) and A(i,j) x_i to a dot
// product reduced over the wave into r_j (the upper product, A(j,i) =
// A(i,j)).  The row accumulators of the four waves meet in `red` (4 * 1024
// doubles of LDS).  Short chunks take more columns per step.
__device__ __noinline__ double wg_residual(int n, const double *__restrict__ A, long lda, const double *__restrict__ b,
                                          const double *__restrict__ x, double *__restrict__ r, double *red,
                                          double *red4)
{
	const int tid = threadIdx.x;
	for (int i = tid; i < n; i += 256)
		r[i] = b[i];
	__syncthreads();
	for (int r0 = 0; r0 < n; r0 += 1024) {
		const int r1 = (n - r0) < 1024 ? n : r0 + 1024;
		const int h = r1 - r0;
		if (h <= 256)
			residual_chunk<4, 8>(r0, r1, A, lda, x, r, red);
		else if (h <= 512)
			residual_chunk<8, 4>(r0, r1, A, lda, x, r, red);
		else
			residual_chunk<16, 2>(r0, r1, A, lda, x, r, red);
	}
	double ss = 0.0;
	for (int i = tid; i < n; i += 256)
		ss += r[i] * r[i];
	return wg_sum(ss, red4);
}

__global__ void k_queue_reset(int *counter) { *counter = 0; }

template <int RB>
__global__ __launch_bounds__(256)
void k_batched_potrs(int nsys, const int *__restrict__ order, const int *__restrict__ nvec,
                     const long *__restrict__ offA, const long *__restrict__ offB, const long *__restrict__ offD,
                     double *__restrict__ A, double *__restrict__ B, int nrhs, double *__restrict__ Dinv,
                     int *__restrict__ info, int *__restrict__ counter)
{
	__shared__ double S[64 * LD];
	__shared__ double Li[64 * LD];
	__shared__ double T[512];
	__shared__ int s_idx, s_bad;
	for (;;) {
		if (threadIdx.x == 0)
			s_idx = atomicAdd(counter, 1);
		__syncthreads();
		const int idx = s_idx;
		__syncthreads();
		if (idx >= nsys)
			break;                     // every workgroup ends here once the queue is empty
		const int sys = order[idx];
		const int n = nvec[sys];
		double *Ai = A + offA[sys];
		double *Di = Dinv + offD[sys];
		const int bad = wg_factor<double, RB>(n, Ai, Di, S, Li, &s_bad);
		if (threadIdx.x == 0)
			info[sys] = bad;
		if (nrhs > 0 && bad == 0)
			for (int c = 0; c < nrhs; c++)
				wg_solve1<double>(n, Ai, n, Di, B + offB[sys] + (long)c * n, T);
		__syncthreads();
	}
}

// lower(A) -> FP32 by one 256-thread workgroup, by 64-column panels: wave v
// takes the panel's columns 16v..16v+15, lanes run down the rows in two
// 64-row chunks (32 loads in flight per lane; only the lower triangle is
// read).  Returns this thread's share of ||A||_F^2 (off-diagonal entries
// twice).
__device__ __noinline__ double wg_lower_to_f32(int n, const double *__restrict__ Ai, float *__restrict__ Afi)
{
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	double ss = 0.0;
	for (int c0 = 0; c0 < n; c0 += 64) {
		const int jc = c0 + 16 * wave;
		for (int i = c0 + lane; i < n; i += 128) {
			const bool in2 = i + 64 < n;      // rows i + 64 > every column of the panel
			double a[16], a2[16];
#pragma unroll
			for (int q = 0; q < 16; q++) {
				const int j = jc + q;
				a[q] = (j < n && i >= j) ? Ai[i + (long)j * n] : 0.0;
				a2[q] = (j < n && in2) ? Ai[i + 64 + (long)j * n] : 0.0;
			}
#pragma unroll
			for (int q = 0; q < 16; q++) {
				const int j = jc + q;
				if (j < n && i >= j) {
					Afi[i + (long)j * n] = (float)a[q];
					ss += (i == j ? 1.0 : 2.0) * a[q] * a[q];
				}
				if (j < n && in2) {
					Afi[i + 64 + (long)j * n] = (float)a2[q];
					ss += 2.0 * a2[q] * a2[q];
				}
			}
		}
	}
	return ss;
}

// FP32 factorisation with FP64 iterative refinement, per system: lower(A)
// is rounded to FP32 into Af (same offsets as A, ld n), factored there in
// FP32 (block inverses into Dinvf), and each right-hand side b is solved by
//   x = 0, r = b;  repeat: d = (L L^T)^{-1} fl32(r) in FP32, x += d,
//   r = b - A x in FP64, berr = ||r|| / (||A||_F ||x|| + ||b||)
// until berr <= tol or maxiter steps.  A and B are only read; X, R (FP64)
// and Df (FP32) hold x, r and d at the offsets of B.  Per system and
// right-hand side iters[] gets the steps taken and berr[] the last
// backward error; a non-positive FP32 pivot sets info and leaves X of that
// system unwritten (iters 0, berr +inf).  Up to n = 1024, x and r stay in
// LDS (X receives x at the end; R is then unused).
// two workgroups per CU (<= 256 VGPRs): one's memory waits overlap the
// other's arithmetic
__global__ __launch_bounds__(256, 2)
void k_batched_posv_ir(int nsys, const int *__restrict__ order, const int *__restrict__ nvec,
                       const long *__restrict__ offA, const long *__restrict__ offB, const long *__restrict__ offD,
                       const double *__restrict__ A, const double *__restrict__ B, double *__restrict__ X, int nrhs,
                       float *__restrict__ Af, float *__restrict__ Dinvf, double *__restrict__ R,
                       float *__restrict__ Df, int *__restrict__ info, int *__restrict__ iters,
                       double *__restrict__ berr, int maxiter, double tol, int *__restrict__ counter)
{
	// S and Li of the FP32 factorisation; as doubles, the 4 x 1024 row
	// accumulators of the residual (the two phases never overlap)
	__shared__ __align__(16) float SL[2 * 64 * LD];
	__shared__ float Tf[512];
	__shared__ double xl[1024], rl[1024];
	__shared__ double red4[4];
	__shared__ int s_idx, s_bad;
	float *S = SL, *Li = SL + 64 * LD;
	double *red = reinterpret_cast<double *>(SL);
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	for (;;) {
		if (tid == 0)
			s_idx = atomicAdd(counter, 1);
		__syncthreads();
		const int idx = s_idx;
		__syncthreads();
		if (idx >= nsys)
			break;                     // every workgroup ends here once the queue is empty
		const int sys = order[idx];
		const int n = nvec[sys];
		const double *Ai = A + offA[sys];
		float *Afi = Af + offA[sys];
		float *Di = Dinvf + offD[sys];
		const double ss = wg_lower_to_f32(n, Ai, Afi);
		const double anorm = sqrt(wg_sum(ss, red4));
		const int bad = wg_factor<float>(n, Afi, Di, S, Li, &s_bad);
		for (int c = 0; c < nrhs; c++) {
			const long o = offB[sys] + (long)c * n;
			const double *b = B + o;
			// x and r in LDS up to n = 1024 (the residual updates r column by
			// column; in global memory each update was a dependent round trip)
			const bool inl = n <= 1024;
			double *x = inl ? xl : X + o, *r = inl ? rl : R + o;
			float *d = Df + o;
			int it = 0;
			double be = __builtin_inf();
			if (bad == 0) {
				double bb = 0.0;
				for (int i = tid; i < n; i += 256) {
					const double v = b[i];
					x[i] = 0.0;
					r[i] = v;
					bb += v * v;
				}
				const double bnorm = sqrt(wg_sum(bb, red4));
				while (it < maxiter) {
					for (int i = tid; i < n; i += 256)
						d[i] = (float)r[i];
					__syncthreads();
					wg_solve1<float>(n, Afi, n, Di, d, Tf);
					double xx = 0.0;
					for (int i = tid; i < n; i += 256) {
						const double v = x[i] + (double)d[i];
						x[i] = v;
						xx += v * v;
					}
					const double xnorm = sqrt(wg_sum(xx, red4));
					const double rr = wg_residual(n, Ai, n, b, x, r, red, red4);
					it++;
					be = rr == 0.0 ? 0.0 : sqrt(rr) / (anorm * xnorm + bnorm);
					if (be <= tol)
						break;
				}
				if (inl)
					for (int i = tid; i < n; i += 256)
						X[o + i] = xl[i];
				__syncthreads();
			}
			if (tid == 0) {
				iters[(long)sys * nrhs + c] = it;
				berr[(long)sys * nrhs + c] = be;
			}
		}
		if (tid == 0)
			info[sys] = bad;
		__syncthreads();
	}
}

} // namespace

extern "C" size_t sa_batched_dinv_elems(int n) { return (size_t)((n + 63) / 64) * 4096; }

// Panel-update shape of the FP64 batched kernel: 16-row blocks per wave (2
// or 4; 0 restores 2).  A process-wide setting read at launch (and frozen
// into a captured graph); returns the previous value, -EINVAL otherwise.
static int g_batched_rows = 2;
extern "C" int sa_set_batched_rows(int rb)
{
	if (rb != 0 && rb != 2 && rb != 4)
		return -EINVAL;
	const int prev = g_batched_rows;
	g_batched_rows = rb ? rb : 2;
	return prev;
}

extern "C" int sa_dpotrs_batched(hipStream_t s, int nsys, const int *order, const int *n, const long *offA,
                                 const long *offB, const long *offD, double *A, double *B, int nrhs, double *Dinv,
                                 int *info, int *counter, int grid)
{
	if (nsys < 0 || nrhs < 0 || grid <= 0 || !order || !n || !offA || !offD || !A || !Dinv || !info || !counter ||
	    (nrhs > 0 && (!B || !offB)))
		return -EINVAL;
	if (nsys == 0)
		return 0;
	hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, s, counter);
	int e = sa_launch_status();
	if (e)
		return e;
	if (g_batched_rows == 4)
		hipLaunchKernelGGL(k_batched_potrs<4>, dim3((unsigned)grid), dim3(256), 0, s, nsys, order, n, offA, offB, offD,
		                   A, B, nrhs, Dinv, info, counter);
	else
		hipLaunchKernelGGL(k_batched_potrs<2>, dim3((unsigned)grid), dim3(256), 0, s, nsys, order, n, offA, offB, offD,
		                   A, B, nrhs, Dinv, info, counter);
	return sa_launch_status();
}

extern "C" int sa_dposv_batched_ir(hipStream_t s, int nsys, const int *order, const int *n, const long *offA,
                                   const long *offB, const long *offD, const double *A, const double *B, double *X,
                                   int nrhs, float *Af, float *Dinvf, double *R, float *Df, int *info, int *iters,
                                   double *berr, int maxiter, double tol, int *counter, int grid)
{
	if (nsys < 0 || nrhs < 0 || grid <= 0 || maxiter < 1 || !(tol >= 0.0) || !order || !n || !offA || !offD ||
	    !A || !Af || !Dinvf || !info || !counter ||
	    (nrhs > 0 && (!B || !offB || !X || !R || !Df || !iters || !berr)))
		return -EINVAL;
	if (nsys == 0)
		return 0;
	hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, s, counter);
	int e = sa_launch_status();
	if (e)
		return e;
	hipLaunchKernelGGL(k_batched_posv_ir, dim3((unsigned)grid), dim3(256), 0, s, nsys, order, n, offA, offB, offD,
	                   A, B, X, nrhs, Af, Dinvf, R, Df, info, iters, berr, maxiter, tol, counter);
	return sa_launch_status();
}

// ---- hipGraph capture of a launch sequence --------------------------------

extern "C" int sa_graph_begin(hipStream_t s)
{
	return (int)hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
}

extern "C" int sa_graph_end(hipStream_t s, void **exec)
{
	hipGraph_t g;
	hipError_t e = hipStreamEndCapture(s, &g);
	if (e != hipSuccess)
		return (int)e;
	hipGraphExec_t x;
	e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
	hipGraphDestroy(g);
	if (e != hipSuccess)
		return (int)e;
	*exec = (void *)x;
	return 0;
}

extern "C" int sa_graph_launch(void *exec, hipStream_t s) { return (int)hipGraphLaunch((hipGraphExec_t)exec, s); }

extern "C" int sa_graph_destroy(void *exec) { return (int)hipGraphExecDestroy((hipGraphExec_t)exec); }
