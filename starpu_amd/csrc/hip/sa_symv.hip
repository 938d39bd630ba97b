// Residual of a symmetric system from the stored lower triangle only:
//   R(n x r) = B - A X,  A symmetric, only A(i, j) with i >= j read.
// The refinement of the mixed-precision solve computes this FP64 residual
// once per step; reading both triangles (a general thin GEMM) moves twice
// the bytes of the matrix.  Here each element of the lower triangle is
// read once and used twice: A(i, j) X(j, :) adds to row i (u) and, below
// the diagonal, A(i, j) X(i, :) adds to row j (v).
//
// k_symm_lower_part: workgroup (run, I, z) takes the 256 rows of row block
// I (4 waves x 64 rows) and a run of up to rl column chunks of 128 columns
// of the lower triangle, right-hand sides 16z .. 16z+15.  A wave walks its
// 64 rows chunk by chunk in 16-column tiles; per 64 x 16 tile it issues 16
// MFMAs (v_mfma_f64_16x16x4f64) for u = tile * X(cols) and 16 for
// v = tile^T * X(rows), the second pass over the tile's columns served by
// the cache.  u stays in registers over the run (one partial per workgroup
// and row); v of a chunk is summed over the four waves in LDS, in wave
// order, and written as that chunk's partial for row block I.
// k_symm_lower_reduce then adds, for every row, its u partials (runs in
// order) and its v partials (row blocks in order) and writes B minus the
// sum, so the result never depends on timing.
#include <cstdlib>

#include "sa_common.h"
#include "starpu_amd_kernels.h"

namespace {

constexpr int SY_ROWS = 256;     // rows per row block (4 waves x 64)
constexpr int SY_COLS = 128;     // columns per chunk

struct symm_geom {
	int nI;      // row blocks
	int nJ;      // column chunks
	int nrun;    // runs of the longest row block
};

// run: column chunks per workgroup
__host__ __device__ inline symm_geom symm_geometry(int n, int run)
{
	symm_geom g;
	g.nI = (n + SY_ROWS - 1) / SY_ROWS;
	g.nJ = (n + SY_COLS - 1) / SY_COLS;
	g.nrun = (g.nJ + run - 1) / run;
	return g;
}

// chunks of row block I: those starting left of min(n, 256 (I + 1))
__host__ __device__ inline int symm_chunks(int n, int I)
{
	const long lim = (long)SY_ROWS * (I + 1) < n ? (long)SY_ROWS * (I + 1) : n;
	return (int)((lim + SY_COLS - 1) / SY_COLS);
}

// Row blocks holding chunk J: I >= J / 2 (chunk J starts left of 256 (I + 1)).
// v partials of chunk J for I = J/2 .. nI-1, chunk after chunk:
// base(J) = J nI - sum_{j < J} floor(j / 2).
__host__ __device__ inline long symm_vbase(int J, int nI)
{
	const long m = J / 2;
	const long h = (J % 2 == 0) ? m * (m - 1) : m * m;
	return (long)J * nI - h;
}

__host__ __device__ inline long symm_vcount(int n)
{
	const symm_geom g = symm_geometry(n, 1);
	return symm_vbase(g.nJ, g.nI);
}

typedef double sy_d4 __attribute__((ext_vector_type(4)));

// 4 doubles at p (VEC: one 32-byte vector load; the host checked the
// alignment) or element by element
template <bool VEC>
__device__ __forceinline__ sy_d4 load4(const double *p)
{
	if (VEC)
		return *reinterpret_cast<const sy_d4 *>(p);
	return sy_d4{p[0], p[1], p[2], p[3]};
}

// The lane layout.  MFMA operand A is 16 x 4 (lane l: row l%16, k = g),
// B is 4 x 16 (lane l: k = g, column l%16), D register q of lane l is
// D(4q + g, l%16).  For the wave's 64 rows rw0.. and the 16 columns c0..:
//   u (4 MFMAs per row block j = 0..3, rows rw0 + 4i + j, i = 0..15):
//       step q: A(rw0 + 4 l%16 + j, c0 + 4g + q) * X(c0 + 4g + q, rhs l%16)
//       -- lane l's A values for the four j are 4 consecutive rows (one
//       vector per q), its X values for the four q 4 consecutive rows;
//   v (16 MFMAs, k = the wave's 64 rows, row rw0 + 16h + 4g + e at step
//       s = 4h + e): X(row, rhs l%16) * A(row, c0 + l%16)
//       -- lane l's A values for step group h are 4 consecutive rows, and
//       the four lane groups g together read 16 consecutive rows (one
//       128-byte line) of each of the 16 columns.
// Elements are masked by index: u takes j <= i, v takes j < i, both
// i, j < n.
//
// Bound: at N = 32768 and 16 right-hand sides the kernel runs about as long
// as the general thin product over both triangles (1.44-1.58 against
// 1.45-1.64 ms) while reading half the bytes: every element feeds 32 FP64
// multiply-adds (u and v), so the MFMAs are 36 % busy in both kernels, and
// with two workgroups per CU (64 KB of LDS each for the v sums, about 210
// registers per lane) the tile loads' latency is not covered (waves wait on
// a dependency 45 % of their cycles); issuing the next tile's loads ahead
// needs registers that leave one workgroup per CU, and ran 1.5x slower.
struct sy_tile {
	sy_d4 au[4], av[4], xc;
};

template <bool VEC>
__global__ __launch_bounds__(256, 2)
void k_symm_lower_part(int n, int r, int rl, const double *__restrict__ A, long lda, const double *__restrict__ X,
                       long ldx, double *__restrict__ Pu, double *__restrict__ Pv)
{
	__shared__ sa_d4 vs[4][SY_COLS / 16][64];
	const symm_geom G = symm_geometry(n, rl);
	const int run = blockIdx.x, I = blockIdx.y, z = blockIdx.z;
	const int nch = symm_chunks(n, I);
	const int ch0 = run * rl;
	if (ch0 >= nch)
		return;                       // the whole workgroup: no barrier is reached
	const int ch1 = (ch0 + rl) < nch ? ch0 + rl : nch;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int l16 = lane & 15, g = lane >> 4;
	const int rw0 = I * SY_ROWS + 64 * wave;        // the wave's first row
	const int c = z * 16 + l16;
	const bool cin = c < r;
	const double *Xc = X + (long)(cin ? c : r - 1) * ldx;
	// X rows of the v steps: xr[4h + e] = X(rw0 + 16h + 4g + e, c)
	double xr[16];
#pragma unroll
	for (int h = 0; h < 4; h++) {
		const int row = rw0 + 16 * h + 4 * g;
		sy_d4 x;
		if (VEC) {
			x = load4<true>(Xc + (row < n ? row : n - 4));
		} else {
#pragma unroll
			for (int e = 0; e < 4; e++)
				x[e] = Xc[row + e < n ? row + e : n - 1];
		}
#pragma unroll
		for (int e = 0; e < 4; e++)
			xr[4 * h + e] = (cin && row + e < n) ? x[e] : 0.0;
	}
	// tiles wholly above the diagonal or past n contribute nothing
	auto active = [&](int c0) { return c0 < n && c0 <= rw0 + 63; };
	// One tile's operands.  VEC (n, lda, ldx multiples of 4, aligned
	// matrices): 32-byte vector loads, every 4-row run wholly inside or
	// wholly outside the matrix, a run outside read from rows n-4.. (in
	// bounds) and masked; otherwise element by element from clamped
	// addresses.  Every element is masked by index, so both forms give the
	// same values.
	auto load = [&](int c0, sy_tile &t) {
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int col = c0 + 4 * g + q, cc = col < n ? col : n - 1;
			const int row = rw0 + 4 * l16;
			sy_d4 x;
			if (VEC) {
				x = load4<true>(A + (long)cc * lda + (row < n ? row : n - 4));
			} else {
#pragma unroll
				for (int j = 0; j < 4; j++)
					x[j] = A[(row + j < n ? row + j : n - 1) + (long)cc * lda];
			}
#pragma unroll
			for (int j = 0; j < 4; j++)
				t.au[q][j] = (row + j < n && col < n && col <= row + j) ? x[j] : 0.0;
		}
		const int col = c0 + l16, cc = col < n ? col : n - 1;
#pragma unroll
		for (int h = 0; h < 4; h++) {
			const int row = rw0 + 16 * h + 4 * g;
			sy_d4 x;
			if (VEC) {
				x = load4<true>(A + (long)cc * lda + (row < n ? row : n - 4));
			} else {
#pragma unroll
				for (int e = 0; e < 4; e++)
					x[e] = A[(row + e < n ? row + e : n - 1) + (long)cc * lda];
			}
#pragma unroll
			for (int e = 0; e < 4; e++)
				t.av[h][e] = (row + e < n && col < n && col < row + e) ? x[e] : 0.0;
		}
		const int xrow = c0 + 4 * g;
		sy_d4 x;
		if (VEC) {
			x = load4<true>(Xc + (xrow < n ? xrow : n - 4));
		} else {
#pragma unroll
			for (int q = 0; q < 4; q++)
				x[q] = Xc[xrow + q < n ? xrow + q : n - 1];
		}
#pragma unroll
		for (int q = 0; q < 4; q++)
			t.xc[q] = (cin && xrow + q < n) ? x[q] : 0.0;
	};
	sa_d4 accu[4];
#pragma unroll
	for (int j = 0; j < 4; j++)
		accu[j] = sa_d4{0.0, 0.0, 0.0, 0.0};
	// u += tile * X(cols); returns v = tile^T * X(rows) (two accumulators,
	// even and odd steps, added at the end)
	auto mfmas = [&](const sy_tile &t) {
		sa_d4 v0 = sa_d4{0.0, 0.0, 0.0, 0.0}, v1 = v0;
#pragma unroll
		for (int q = 0; q < 4; q++) {
#pragma unroll
			for (int j = 0; j < 4; j++)
				accu[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(t.au[q][j], t.xc[q], accu[j], 0, 0, 0);
			v0 = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[4 * q], t.av[q][0], v0, 0, 0, 0);
			v1 = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[4 * q + 1], t.av[q][1], v1, 0, 0, 0);
			v0 = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[4 * q + 2], t.av[q][2], v0, 0, 0, 0);
			v1 = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[4 * q + 3], t.av[q][3], v1, 0, 0, 0);
		}
		return v0 + v1;
	};
	constexpr int NT = SY_COLS / 16;
	for (int ch = ch0; ch < ch1; ch++) {
		const int col0 = ch * SY_COLS;
		for (int ct = 0; ct < NT; ct++) {
			const int ca = col0 + 16 * ct;
			sa_d4 v = sa_d4{0.0, 0.0, 0.0, 0.0};
			if (active(ca)) {
				sy_tile t0;
				load(ca, t0);
				v = mfmas(t0);
			}
			vs[wave][ct][lane] = v;
		}
		__syncthreads();
		// v of this chunk: lane l holds v(c0 + l%16, rhs 4q + g) of tile ct;
		// wave w sums tiles 2w, 2w + 1 over the waves and stores them
#pragma unroll
		for (int h = 0; h < 2; h++) {
			const int ct = 2 * wave + h;
			sa_d4 sum = vs[0][ct][lane];
#pragma unroll
			for (int w = 1; w < 4; w++)
				sum += vs[w][ct][lane];
			double *P = Pv + ((long)z * symm_vcount(n) + symm_vbase(ch, G.nI) + (I - ch / 2)) * (16L * SY_COLS);
#pragma unroll
			for (int q = 0; q < 4; q++)
				P[(long)(4 * q + g) * SY_COLS + 16 * ct + l16] = sum[q];
		}
		__syncthreads();
	}
	// u: lane l holds u(rw0 + 4 (4q + g) + j, rhs l%16) of row block j
	double *P = Pu + (((long)z * G.nI + I) * G.nrun + run) * (16L * SY_ROWS);
#pragma unroll
	for (int j = 0; j < 4; j++)
#pragma unroll
		for (int q = 0; q < 4; q++)
			P[(long)l16 * SY_ROWS + 64 * wave + 4 * (4 * q + g) + j] = accu[j][q];
}

// R(i, c) = B(i, c) - (sum of the u partials of row i's block over its runs,
// then the v partials of row i's chunk over the row blocks I >= J / 2).
__global__ __launch_bounds__(256)
void k_symm_lower_reduce(int n, int r, int rl, const double *__restrict__ Pu, const double *__restrict__ Pv,
                         const double *__restrict__ B, long ldb, double *__restrict__ R, long ldr)
{
	const long t = (long)blockIdx.x * 256 + threadIdx.x;
	if (t >= (long)n * r)
		return;
	const int i = (int)(t % n), c = (int)(t / n), z = c / 16, cc = c % 16;
	const symm_geom G = symm_geometry(n, rl);
	const int I = i / SY_ROWS, J = i / SY_COLS;
	const int runs = (symm_chunks(n, I) + rl - 1) / rl;
	double s = 0.0;
	const double *pu = Pu + (((long)z * G.nI + I) * G.nrun) * (16L * SY_ROWS) + (long)cc * SY_ROWS + i % SY_ROWS;
	for (int q = 0; q < runs; q++)
		s += pu[(long)q * 16 * SY_ROWS];
	const double *pv = Pv + ((long)z * symm_vcount(n) + symm_vbase(J, G.nI)) * (16L * SY_COLS) + (long)cc * SY_COLS +
	                   i % SY_COLS;
	for (int q = 0; q < G.nI - J / 2; q++)
		s += pv[(long)q * 16 * SY_COLS];
	R[i + (long)c * ldr] = B[i + (long)c * ldb] - s;
}

// Column chunks per workgroup: 2 (SA_SYMM_RUN overrides, 1..64).  Shorter
// runs give more, smaller workgroups, so the last wave of workgroups is a
// smaller share of the launch (at n = 32768, 16 chunks ran 1.92 ms, 4
// chunks 1.57 ms, 2 chunks 1.44 ms), and more u partial sums.
int symm_run()
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_SYMM_RUN");
		const int x = e ? atoi(e) : 0;
		v = (x >= 1 && x <= 64) ? x : 2;
	}
	return v;
}

size_t symm_pu_elems(int n, int r)
{
	const symm_geom G = symm_geometry(n, symm_run());
	return (size_t)((r + 15) / 16) * G.nI * G.nrun * 16 * SY_ROWS;
}

size_t symm_pv_elems(int n, int r)
{
	return (size_t)((r + 15) / 16) * (size_t)symm_vcount(n) * 16 * SY_COLS;
}

} // namespace

extern "C" size_t sa_dsymm_lower_resid_worksize(int n, int r)
{
	if (n <= 0 || r <= 0)
		return 0;
	const size_t pu = (symm_pu_elems(n, r) * sizeof(double) + 255) & ~(size_t)255;
	return pu + symm_pv_elems(n, r) * sizeof(double);
}

extern "C" int sa_dsymm_lower_resid(hipStream_t s, int n, int r, const double *A, int lda, const double *X, int ldx,
                                    const double *B, int ldb, double *R, int ldr, void *work, size_t work_bytes)
{
	if (n < 0 || r < 0 || lda < (n > 1 ? n : 1) || ldx < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1) ||
	    ldr < (n > 1 ? n : 1))
		return -EINVAL;
	if (n == 0 || r == 0)
		return 0;
	if (!A || !X || !B || !R || !work || work_bytes < sa_dsymm_lower_resid_worksize(n, r))
		return -EINVAL;
	const int rl = symm_run();
	const symm_geom G = symm_geometry(n, rl);
	double *Pu = (double *)work;
	double *Pv = (double *)((char *)work + ((symm_pu_elems(n, r) * sizeof(double) + 255) & ~(size_t)255));
	const dim3 grid((unsigned)G.nrun, (unsigned)G.nI, (unsigned)((r + 15) / 16));
	// 32-byte vector loads when every 4-row run of A and X is so aligned
	// and lies wholly inside or outside the matrix
	const bool vec = n % 4 == 0 && lda % 4 == 0 && ldx % 4 == 0 && (uintptr_t)A % 32 == 0 && (uintptr_t)X % 32 == 0;
	if (vec)
		hipLaunchKernelGGL(k_symm_lower_part<true>, grid, dim3(256), 0, s, n, r, rl, A, (long)lda, X, (long)ldx, Pu,
		                   Pv);
	else
		hipLaunchKernelGGL(k_symm_lower_part<false>, grid, dim3(256), 0, s, n, r, rl, A, (long)lda, X, (long)ldx, Pu,
		                   Pv);
	int e = sa_launch_status();
	if (e)
		return e;
	const long cnt = (long)n * r;
	hipLaunchKernelGGL(k_symm_lower_reduce, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, n, r, rl,
	                   (const double *)Pu, (const double *)Pv, B, (long)ldb, R, (long)ldr);
	return sa_launch_status();
}
