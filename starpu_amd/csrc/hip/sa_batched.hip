// Batched factor-and-solve of many independent small SPD systems of uneven
// sizes in one persistent kernel (gfx950).
//
// The grid is a fixed set of workgroups that stay resident; each pulls the
// next system from a device-side queue (an atomic counter over a list of
// systems, largest first, so the big ones start early and the small ones
// fill the tail) and processes it completely:
//   factor   right-looking over 64-column panels, in place in the system's
//            own column-major storage: factor-and-invert the 64x64 diagonal
//            block in LDS (block_potrf_inv64), panel <- panel * inv(L_jj)^T
//            (MFMA, rows_apply_inv), lower(A22) -= P P^T (MFMA over 32x32
//            wave tiles, operands from L2); the block inverses are kept for
//            the solve;
//   solve    L y = b then L^T x = y by 64-row blocks: partial products over
//            the solved part (coalesced column reads, sums through LDS),
//            then the block inverse.
// Every workgroup leaves the loop when the queue is empty, so the grid
// drains before the kernel ends.  k_queue_reset zeroes the counter with a
// kernel (not a memset), so the pair can be captured in a hipGraph and
// replayed.
#include "sa_block64.h"
#include "starpu_amd_kernels.h"

namespace {
using namespace sa64;

// lower(C) (n x n, ld ldc) -= P P^T with P (n x w, w <= 64, ld ldp), by one
// 256-thread workgroup: wave tiles of 32x32 over the lower triangle.  The
// MFMA computes the transposed block so a lane's results are consecutive
// rows of C.
template <typename T>
__device__ void wg_syrk_lower(int n, int w, const T *__restrict__ P, long ldp, T *__restrict__ C, long ldc)
{
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
	const int nt = (n + 31) / 32;
	const int ntiles = nt * (nt + 1) / 2;
	for (int t = wave; t < ntiles; t += 4) {
		// tile (ti, tj), ti >= tj, enumerated row by row of the lower triangle
		int ti = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
		while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
		while (ti * (ti + 1) / 2 > t) ti--;
		const int tj = t - ti * (ti + 1) / 2;
		const int i0 = 32 * ti, j0 = 32 * tj;
		typename V4<T>::type acc[2][2];
#pragma unroll
		for (int a = 0; a < 2; a++)
#pragma unroll
			for (int b = 0; b < 2; b++)
				acc[a][b] = zero4<T>();
#pragma unroll
		for (int k0 = 0; k0 < 64; k0 += 4) {
			const int k = k0 + g;
			const bool kin = k < w;
			const long kc = (long)(kin ? k : w - 1) * ldp;
			T xa[2], xb[2];
#pragma unroll
			for (int b = 0; b < 2; b++) {
				const int rj = j0 + 16 * b + r, ri = i0 + 16 * b + r;
				const T va = P[(rj < n ? rj : n - 1) + kc];
				const T vb = P[(ri < n ? ri : n - 1) + kc];
				xa[b] = (kin && rj < n) ? va : T(0);
				xb[b] = (kin && ri < n) ? vb : T(0);
			}
#pragma unroll
			for (int bj = 0; bj < 2; bj++)
#pragma unroll
				for (int bi = 0; bi < 2; bi++)
					acc[bj][bi] = sa_mfma(xa[bj], xb[bi], acc[bj][bi]);
		}
		// acc[bj][bi][q] = (P P^T)(i0 + 16 bi + r, j0 + 16 bj + drow(q, g))
		T cv[2][4][2];
#pragma unroll
		for (int bj = 0; bj < 2; bj++)
#pragma unroll
			for (int q = 0; q < 4; q++)
#pragma unroll
				for (int bi = 0; bi < 2; bi++) {
					const int ic = i0 + 16 * bi + r, jc = j0 + 16 * bj + drow<T>(q, g);
					cv[bj][q][bi] = C[(ic < n ? ic : n - 1) + (long)(jc < n ? jc : n - 1) * ldc];
				}
#pragma unroll
		for (int bj = 0; bj < 2; bj++)
#pragma unroll
			for (int q = 0; q < 4; q++)
#pragma unroll
				for (int bi = 0; bi < 2; bi++) {
					const int ic = i0 + 16 * bi + r, jc = j0 + 16 * bj + drow<T>(q, g);
					if (ic < n && jc < n && ic >= jc)
						C[ic + (long)jc * ldc] = cv[bj][q][bi] - acc[bj][bi][q];
				}
	}
}

// In-place solve of L L^T x = b for one right-hand side of length n
// (x overwrites b), L lower with its 64x64 block inverses Dinv (block j at
// Dinv + 4096 j, ld 64), all in precision T.  One 256-thread workgroup;
// Tl is LDS scratch of 4 * 64 + 64 elements.
template <typename T>
__device__ void wg_solve1(int n, const T *__restrict__ L, long ldl, const T *__restrict__ Dinv,
                          T *__restrict__ x, T *__restrict__ Tl)
{
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	T *part = Tl, *t = Tl + 256;
	const int nblk = (n + 63) / 64;
	// forward: y_j = inv(L_jj) (b_j - L(j, 0:j0) y(0:j0))
	for (int jb = 0; jb < nblk; jb++) {
		const int j0 = 64 * jb, w = (n - j0) < 64 ? (n - j0) : 64;
		const int row = j0 + lane;
		const int rowc = lane < w ? row : j0;
		T s = T(0);
		for (int k = wave; k < j0; k += 4)
			s += L[rowc + (long)k * ldl] * x[k];
		part[wave * 64 + lane] = s;
		__syncthreads();
		if (wave == 0)
			t[lane] = lane < w ? x[row] - ((part[lane] + part[64 + lane]) + (part[128 + lane] + part[192 + lane]))
			                   : T(0);
		__syncthreads();
		if (wave == 0 && lane < w) {
			const T *Dj = Dinv + 4096L * jb;
			T y = T(0);
			for (int q = 0; q < w; q++)
				y += Dj[lane + 64 * q] * t[q];
			x[row] = y;
		}
		__syncthreads();
	}
	// backward: x_j = inv(L_jj)^T (y_j - L(j0+w:n, j)^T x(j0+w:n))
	for (int jb = nblk - 1; jb >= 0; jb--) {
		const int j0 = 64 * jb, w = (n - j0) < 64 ? (n - j0) : 64;
		const int kb = j0 + w;
		// wave v sums columns j0 + v, j0 + v + 4, ...: lanes run down the column
		for (int c = wave; c < w; c += 4) {
			T s = T(0);
			for (int k = kb + lane; k < n; k += 64)
				s += L[k + (long)(j0 + c) * ldl] * x[k];
			for (int off = 32; off >= 1; off >>= 1)
				s += __shfl_xor(s, off);
			if (lane == 0)
				t[c] = x[j0 + c] - s;
		}
		__syncthreads();
		if (wave == 0 && lane < w) {
			const T *Dj = Dinv + 4096L * jb;
			T v = T(0);
			for (int q = 0; q < w; q++)
				v += Dj[q + 64 * lane] * t[q];
			x[j0 + lane] = v;
		}
		__syncthreads();
	}
}

// Right-looking factorisation of one n x n system (lower, ld n, in place)
// by the workgroup, block inverses into Di.  Returns in thread 0 the
// 1-based column of the first non-positive pivot (or 0); *s_bad is LDS.
template <typename T>
__device__ int wg_factor(int n, T *__restrict__ Ai, T *__restrict__ Di, T *__restrict__ S, T *__restrict__ Li,
                         int *s_bad)
{
	if (threadIdx.x == 0)
		*s_bad = 0;
	for (int j0 = 0; j0 < n; j0 += 64) {
		const int w = (n - j0) < 64 ? (n - j0) : 64;
		T *Dj = Di + 4096L * (j0 / 64);
		const int bad = block_potrf_inv64_t<T, true>(w, Ai + j0 + (long)j0 * n, n, Dj, S, Li);
		if (threadIdx.x == 0 && bad && bad <= w && *s_bad == 0)
			*s_bad = j0 + bad;
		const int rows = n - j0 - w;
		if (rows > 0) {
			T *Pn = Ai + (j0 + w) + (long)j0 * n;
			for (int row0 = (threadIdx.x >> 6) * 16; row0 < rows; row0 += 64)
				rows_apply_inv_t<T>(row0, rows, w, Dj, Pn, n);
			__syncthreads();
			wg_syrk_lower<T>(rows, w, Pn, n, Ai + (j0 + w) + (long)(j0 + w) * n, n);
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

// r <- b - A x in FP64 for the symmetric A of order n given by its lower
// triangle (ld lda), one 256-thread workgroup; returns ||r||^2 to every
// thread.  The lower triangle is read once: rows in chunks of 1024, wave v
// walks columns j = v, v + 4, ... of the chunk; lane l owns rows
// r0 + 64 m + l (m < 16), adding A(i,j) x_j to its row accumulators (the
// lower product) and A(i,j) x_i to a dot product reduced over the wave
// into r_j (the upper product, A(j,i) = A(i,j)).  The row accumulators of
// the four waves meet in `red` (4 * 1024 doubles of LDS).
__device__ double wg_residual(int n, const double *__restrict__ A, long lda, const double *__restrict__ b,
                              const double *__restrict__ x, double *__restrict__ r, double *red, double *red4)
{
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	for (int i = tid; i < n; i += 256)
		r[i] = b[i];
	__syncthreads();
	for (int r0 = 0; r0 < n; r0 += 1024) {
		const int r1 = (n - r0) < 1024 ? n : r0 + 1024;
		double acc[16], xr[16];
#pragma unroll
		for (int m = 0; m < 16; m++) {
			const int i = r0 + 64 * m + lane;
			xr[m] = i < r1 ? x[i] : 0.0;
			acc[m] = 0.0;
		}
		for (int j = wave; j < r1; j += 4) {
			const double xj = x[j];
			const double *Aj = A + (long)j * lda;
			double dot = 0.0;
#pragma unroll
			for (int m = 0; m < 16; m++) {
				const int i = r0 + 64 * m + lane;
				if (i >= j && i < r1) {
					const double a = Aj[i];
					acc[m] = fma(a, xj, acc[m]);
					if (i > j)
						dot = fma(a, xr[m], dot);
				}
			}
			for (int off = 32; off >= 1; off >>= 1)
				dot += __shfl_xor(dot, off);
			if (lane == 0)
				r[j] -= dot;
		}
#pragma unroll
		for (int m = 0; m < 16; m++)
			red[wave * 1024 + 64 * m + lane] = acc[m];
		__syncthreads();
		for (int i = r0 + tid; i < r1; i += 256) {
			const int q = i - r0;
			r[i] -= (red[q] + red[1024 + q]) + (red[2048 + q] + red[3072 + q]);
		}
		__syncthreads();
	}
	double ss = 0.0;
	for (int i = tid; i < n; i += 256)
		ss += r[i] * r[i];
	return wg_sum(ss, red4);
}

__global__ void k_queue_reset(int *counter) { *counter = 0; }

__global__ __launch_bounds__(256)
void k_batched_potrs(int nsys, const int *__restrict__ order, const int *__restrict__ nvec,
                     const long *__restrict__ offA, const long *__restrict__ offB, const long *__restrict__ offD,
                     double *__restrict__ A, double *__restrict__ B, int nrhs, double *__restrict__ Dinv,
                     int *__restrict__ info, int *__restrict__ counter)
{
	__shared__ double S[64 * LD];
	__shared__ double Li[64 * LD];
	__shared__ double T[320];
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
		const int bad = wg_factor<double>(n, Ai, Di, S, Li, &s_bad);
		if (threadIdx.x == 0)
			info[sys] = bad;
		if (nrhs > 0 && bad == 0)
			for (int c = 0; c < nrhs; c++)
				wg_solve1<double>(n, Ai, n, Di, B + offB[sys] + (long)c * n, T);
		__syncthreads();
	}
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
// system unwritten (iters 0, berr +inf).
__global__ __launch_bounds__(256)
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
	__shared__ float Tf[320];
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
		// lower(A) -> FP32, and ||A||_F^2 (off-diagonal entries twice)
		double ss = 0.0;
		for (int j = wave; j < n; j += 4)
			for (int i = j + lane; i < n; i += 64) {
				const double a = Ai[i + (long)j * n];
				Afi[i + (long)j * n] = (float)a;
				ss += (i == j ? 1.0 : 2.0) * a * a;
			}
		const double anorm = sqrt(wg_sum(ss, red4));
		const int bad = wg_factor<float>(n, Afi, Di, S, Li, &s_bad);
		for (int c = 0; c < nrhs; c++) {
			const long o = offB[sys] + (long)c * n;
			const double *b = B + o;
			double *x = X + o, *r = R + o;
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
	hipLaunchKernelGGL(k_batched_potrs, dim3((unsigned)grid), dim3(256), 0, s, nsys, order, n, offA, offB, offD, A,
	                   B, nrhs, Dinv, info, counter);
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
