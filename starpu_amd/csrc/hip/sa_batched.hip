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
__device__ void wg_syrk_lower(int n, int w, const double *__restrict__ P, long ldp, double *__restrict__ C,
                              long ldc)
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
		sa_d4 acc[2][2];
#pragma unroll
		for (int a = 0; a < 2; a++)
#pragma unroll
			for (int b = 0; b < 2; b++)
				acc[a][b] = sa_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
		for (int k0 = 0; k0 < 64; k0 += 4) {
			const int k = k0 + g;
			const bool kin = k < w;
			const long kc = (long)(kin ? k : w - 1) * ldp;
			double xa[2], xb[2];
#pragma unroll
			for (int b = 0; b < 2; b++) {
				const int rj = j0 + 16 * b + r, ri = i0 + 16 * b + r;
				const double va = P[(rj < n ? rj : n - 1) + kc];
				const double vb = P[(ri < n ? ri : n - 1) + kc];
				xa[b] = (kin && rj < n) ? va : 0.0;
				xb[b] = (kin && ri < n) ? vb : 0.0;
			}
#pragma unroll
			for (int bj = 0; bj < 2; bj++)
#pragma unroll
				for (int bi = 0; bi < 2; bi++)
					acc[bj][bi] = sa_mfma(xa[bj], xb[bi], acc[bj][bi]);
		}
		// acc[bj][bi][q] = (P P^T)(i0 + 16 bi + r, j0 + 16 bj + 4q + g)
		double cv[2][4][2];
#pragma unroll
		for (int bj = 0; bj < 2; bj++)
#pragma unroll
			for (int q = 0; q < 4; q++)
#pragma unroll
				for (int bi = 0; bi < 2; bi++) {
					const int ic = i0 + 16 * bi + r, jc = j0 + 16 * bj + 4 * q + g;
					cv[bj][q][bi] = C[(ic < n ? ic : n - 1) + (long)(jc < n ? jc : n - 1) * ldc];
				}
#pragma unroll
		for (int bj = 0; bj < 2; bj++)
#pragma unroll
			for (int q = 0; q < 4; q++)
#pragma unroll
				for (int bi = 0; bi < 2; bi++) {
					const int ic = i0 + 16 * bi + r, jc = j0 + 16 * bj + 4 * q + g;
					if (ic < n && jc < n && ic >= jc)
						C[ic + (long)jc * ldc] = cv[bj][q][bi] - acc[bj][bi][q];
				}
	}
}

// In-place solve of L L^T x = b for one right-hand side of length n
// (x overwrites b), L lower with its 64x64 block inverses Dinv (block j at
// Dinv + 4096 j, ld 64).  One 256-thread workgroup; T is LDS scratch of
// 4 * 64 + 64 doubles.
__device__ void wg_solve1(int n, const double *__restrict__ L, long ldl, const double *__restrict__ Dinv,
                          double *__restrict__ x, double *__restrict__ T)
{
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	double *part = T, *t = T + 256;
	const int nblk = (n + 63) / 64;
	// forward: y_j = inv(L_jj) (b_j - L(j, 0:j0) y(0:j0))
	for (int jb = 0; jb < nblk; jb++) {
		const int j0 = 64 * jb, w = (n - j0) < 64 ? (n - j0) : 64;
		const int row = j0 + lane;
		const int rowc = lane < w ? row : j0;
		double s = 0.0;
		for (int k = wave; k < j0; k += 4)
			s += L[rowc + (long)k * ldl] * x[k];
		part[wave * 64 + lane] = s;
		__syncthreads();
		if (wave == 0)
			t[lane] = lane < w ? x[row] - ((part[lane] + part[64 + lane]) + (part[128 + lane] + part[192 + lane]))
			                   : 0.0;
		__syncthreads();
		if (wave == 0 && lane < w) {
			const double *Dj = Dinv + 4096L * jb;
			double y = 0.0;
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
			double s = 0.0;
			for (int k = kb + lane; k < n; k += 64)
				s += L[k + (long)(j0 + c) * ldl] * x[k];
			for (int off = 32; off >= 1; off >>= 1)
				s += __shfl_xor(s, off);
			if (lane == 0)
				t[c] = x[j0 + c] - s;
		}
		__syncthreads();
		if (wave == 0 && lane < w) {
			const double *Dj = Dinv + 4096L * jb;
			double v = 0.0;
			for (int q = 0; q < w; q++)
				v += Dj[q + 64 * lane] * t[q];
			x[j0 + lane] = v;
		}
		__syncthreads();
	}
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
		if (threadIdx.x == 0)
			s_bad = 0;
		for (int j0 = 0; j0 < n; j0 += 64) {
			const int w = (n - j0) < 64 ? (n - j0) : 64;
			double *Dj = Di + 4096L * (j0 / 64);
			const int bad = block_potrf_inv64<true>(w, Ai + j0 + (long)j0 * n, n, Dj, S, Li);
			if (threadIdx.x == 0 && bad && bad <= w && s_bad == 0)
				s_bad = j0 + bad;
			const int rows = n - j0 - w;
			if (rows > 0) {
				double *Pn = Ai + (j0 + w) + (long)j0 * n;
				for (int row0 = (threadIdx.x >> 6) * 16; row0 < rows; row0 += 64)
					rows_apply_inv(row0, rows, w, Dj, Pn, n);
				__syncthreads();
				wg_syrk_lower(rows, w, Pn, n, Ai + (j0 + w) + (long)(j0 + w) * n, n);
				__syncthreads();
			}
		}
		__syncthreads();
		if (threadIdx.x == 0)
			info[sys] = s_bad;
		if (nrhs > 0 && s_bad == 0)
			for (int c = 0; c < nrhs; c++)
				wg_solve1(n, Ai, n, Di, B + offB[sys] + (long)c * n, T);
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
