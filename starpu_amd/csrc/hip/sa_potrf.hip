// Cholesky factorisation of one diagonal tile (n <= 2048), with the inverse
// of its factor, in a single persistent kernel (gfx950).
//
// Inside a task graph the diagonal-tile POTRF heads the critical path, and
// while it runs the other HIP streams keep every compute unit busy with
// tile GEMMs.  A POTRF built from one kernel launch per 64-column step (a
// 64x64 factor-and-invert, a panel apply, a trailing update: three
// dependent launches per step, 48 for a 1024 tile) then waits for the
// command processor and for free workgroup slots at every launch.  Here the
// whole tile is one launch: its workgroups wait for each other through
// counters in device memory instead.
//
// Work is split into items on 64x64 blocks (block b covers rows and columns
// 64b .. 64b+w_b-1 of the tile).  The factorisation:
//   D(j)      factor and invert the diagonal block (j, j)        (LDS, MFMA)
//   P(i,j)    L(i,j) = A(i,j) inv(L(j,j))^T                          (MFMA)
//   U(i,l,j)  A(i,l) -= L(i,j) L(l,j)^T, j < l <= i (lower part if i = l)
// and, when the inverse is asked for, W = L^-T (upper) block by block,
// W(j,i) = -(sum_{t=j}^{i-1} W(j,t) L(i,t)^T) inv(L(i,i))^T, W(j,j) =
// inv(L(j,j))^T:
//   WD(t)     W(t,t) = inv(L(t,t))^T, the inverse D(t) formed
//   WX(j,i)   the sum over t in registers, taking each product as soon as
//             W(j,t) and L(i,t) are final, then the scaling by
//             inv(L(i,i))^T (MFMA); one store and one publication per block
// with Linv = L^-1 = W^T written beside W by WD and WX when asked for.
// These are the inverse's flops (n^3/3) on the same matrix cores, overlapped
// with the factorisation's critical chain instead of following it.
//
// Ordering and progress.  Every item depends only on items that come earlier
// in one fixed global order: step by step, the critical chain D(j), P(j+1,j),
// U(j+1,j+1,j), then the other factor items of step j, then the inverse
// items of block column j - 1, WD(j-1) and WX(t, j-1) for t < j-1 (one step
// late, so that the factor items the chain waits for come first).  Workgroup 0 runs the critical chain; the other
// workgroups take the remaining items in that order from a ticket counter.
// Each workgroup processes its items in increasing order, so the lowest
// unfinished item always has all its inputs ready and a running workgroup
// holds it: the grid always drains.  Block (i, l) of A carries a counter:
// the number of updates applied, then +1 when the block is final; block
// (j, i) of W is raised to 1 once final.  The updates of a block are
// applied in step order, so the result does not depend on timing.  Inverting an already factored
// tile (sa_dtrtri_tile) runs the same inverse items after D items that only
// invert, so a solve on a resident factor multiplies by the same bits.
//
// Every wait is bounded: a counter that does not reach its value within
// seconds (which cannot happen unless the order above is broken) raises
// an abort flag that ends every workgroup's loop, and the factorisation then
// reports info = -1, so a fault of the ordering shows as an error, not as a
// hung device.
//
// Visibility across workgroups (possibly on different XCDs, each with its own
// L2) uses agent-scope release fences before a counter is raised and an
// agent-scope acquire after a wait.
#include <cstdlib>
#include <mutex>

#include "sa_block64.h"
#include "starpu_amd_kernels.h"

namespace {
using namespace sa64;

constexpr int PF_MAXN = 2048;

// Workgroups that take the bulk items (besides the chain's): more finish a
// tile sooner alone, fewer leave more compute units to the kernels running
// beside it.  SA_POTRF_WORKGROUPS overrides (1..255).
int max_bulk()
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("SA_POTRF_WORKGROUPS");
		const int x = e ? atoi(e) : 0;
		v = (x >= 1 && x <= 255) ? x : 31;
	}
	return v;
}

struct PfArgs {
	int n, nb, nbulk;
	int factor;         // 1: factor A in place; 0: A holds the factor already
	int inv;            // form W = L^-T (and Linv = L^-1 when Linv is given)
	int delay;          // steps the inverse items trail the factor items by
	double *A;
	long lda;
	double *W;
	long ldw;
	double *Linv;
	long ldi;
	double *Dinv;       // nb blocks of 64x64 (ld 64)
	unsigned *fa;       // counters of the blocks of A: index i + nb * l
	unsigned *ticket;   // next item of the bulk order to hand out
	unsigned *abort;    // raised by a wait that ran out
	unsigned *fw;       // counters of the blocks of W: index j + nb * i
	int *info;
};

__device__ __forceinline__ int bw(const PfArgs &p, int b) { return (p.n - 64 * b) < 64 ? (p.n - 64 * b) : 64; }

// A wait gives up after SPIN_LIMIT polls.  Each poll is two device-coherent
// loads of about a microsecond and a short sleep, so 2^22 polls are several
// seconds: far beyond any wait of a correct schedule (an item takes tens of
// microseconds).
constexpr unsigned SPIN_LIMIT = 1u << 22;

__device__ __forceinline__ bool spin(const unsigned *f, unsigned v, unsigned *abortf)
{
	unsigned it = 0;
	while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) {
		if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
			return false;
		if (++it > SPIN_LIMIT) {
			__hip_atomic_store(abortf, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			return false;
		}
		__builtin_amdgcn_s_sleep(1);
	}
	return true;
}

// Thread 0 waits until each given counter reaches its value; the whole
// workgroup then sees the data published before those counters were raised.
// False (in every thread) if the wait was abandoned.
__device__ __forceinline__ bool wait_for(const PfArgs &p, const unsigned *f1, unsigned v1,
                                         const unsigned *f2 = nullptr, unsigned v2 = 0,
                                         const unsigned *f3 = nullptr, unsigned v3 = 0)
{
	__shared__ int s_ok;
	if (threadIdx.x == 0) {
		bool ok = !f1 || spin(f1, v1, p.abort);
		ok = ok && (!f2 || spin(f2, v2, p.abort));
		ok = ok && (!f3 || spin(f3, v3, p.abort));
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
		s_ok = ok ? 1 : 0;
	}
	__syncthreads();
	const bool ok = s_ok != 0;
	__syncthreads();
	return ok;
}

// The abort flag as the whole workgroup sees it (one read, broadcast).
__device__ __forceinline__ bool aborted(const PfArgs &p)
{
	__shared__ int s_ab;
	if (threadIdx.x == 0)
		s_ab = __hip_atomic_load(p.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 1 : 0;
	__syncthreads();
	const bool a = s_ab != 0;
	__syncthreads();
	return a;
}

// Every thread's stores become visible at agent scope, then the counter is
// raised.  Each wave waits for its own stores to complete before the
// barrier; thread 0's agent-scope release then covers them all, so the L2
// write-back runs once per item rather than once per wave.
__device__ __forceinline__ void publish(unsigned *f, unsigned v)
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	if (threadIdx.x == 0)
		__hip_atomic_store(f, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// acc += Pr Pc^T over k < wk for the wave's 32x32 quadrant (rows i0.., cols
// j0.., i0 = 32 (wave & 1), j0 = 32 (wave >> 1)) of a wr x wc block; Pr
// (wr x wk, ld ldr), Pc (wc x wk, ld ldpc).  The MFMA computes the transposed
// product so a lane's results are consecutive rows: acc[bj][bi][q] =
// (Pr Pc^T)(i0 + 16 bi + r, j0 + 16 bj + 4q + g).  Operand loads in two
// batches of half the k range: the blocks sit in L2 and the item is
// latency-bound.
__device__ __forceinline__ void acc_nt(int wr, int wc, int wk, const double *__restrict__ Pr, long ldr,
                                       const double *__restrict__ Pc, long ldpc, sa_d4 (&acc)[2][2])
{
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
	const int i0 = 32 * (wave & 1), j0 = 32 * (wave >> 1);
	double xa[2][8], xb[2][8];
#pragma unroll
	for (int h = 0; h < 2; h++) {
#pragma unroll
		for (int s = 0; s < 8; s++) {
			const int k = 4 * (8 * h + s) + g;
			const bool kin = k < wk;
			const long kr = (long)(kin ? k : 0) * ldr, kc = (long)(kin ? k : 0) * ldpc;
#pragma unroll
			for (int b = 0; b < 2; b++) {
				const int rj = j0 + 16 * b + r, ri = i0 + 16 * b + r;
				const double va = Pc[(rj < wc ? rj : 0) + kc];
				const double vb = Pr[(ri < wr ? ri : 0) + kr];
				xa[b][s] = (kin && rj < wc) ? va : 0.0;
				xb[b][s] = (kin && ri < wr) ? vb : 0.0;
			}
		}
#pragma unroll
		for (int s = 0; s < 8; s++)
#pragma unroll
			for (int bj = 0; bj < 2; bj++)
#pragma unroll
				for (int bi = 0; bi < 2; bi++)
					acc[bj][bi] = sa_mfma(xa[bj][s], xb[bi][s], acc[bj][bi]);
	}
}

// C (wr x wc, ld ldc) -= Pr Pc^T over k < wk; lower part only if LOWER.
// Each wave a 32x32 quadrant (acc_nt); the quadrant of C is loaded with
// the operands and written once.
__device__ __forceinline__ void upd_nt(int wr, int wc, int wk, const double *__restrict__ Pr, long ldr,
                                       const double *__restrict__ Pc, long ldpc, double *__restrict__ C, long ldc,
                                       bool lower)
{
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
	const int i0 = 32 * (wave & 1), j0 = 32 * (wave >> 1);
	if (lower && j0 > i0 + 31)
		return;
	double cv[2][4][2];
#pragma unroll
	for (int bj = 0; bj < 2; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int bi = 0; bi < 2; bi++) {
				const int ic = i0 + 16 * bi + r, jc = j0 + 16 * bj + 4 * q + g;
				cv[bj][q][bi] = C[(ic < wr ? ic : 0) + (long)(jc < wc ? jc : 0) * ldc];
			}
	sa_d4 acc[2][2];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++)
			acc[a][b] = sa_d4{0.0, 0.0, 0.0, 0.0};
	acc_nt(wr, wc, wk, Pr, ldr, Pc, ldpc, acc);
#pragma unroll
	for (int bj = 0; bj < 2; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int bi = 0; bi < 2; bi++) {
				const int ic = i0 + 16 * bi + r, jc = j0 + 16 * bj + 4 * q + g;
				if (ic < wr && jc < wc && (!lower || ic >= jc))
					C[ic + (long)jc * ldc] = cv[bj][q][bi] - acc[bj][bi][q];
			}
}

// ---- factorisation items ---------------------------------------------------

// The chain applies the last update of block (j, j) itself just before
// D(j) (chain_U), so D(j) has no wait of its own.  Without FACTOR (the tile
// is factored already) D(j) forms the inverse of the diagonal block only.
template <bool FACTOR>
__device__ __noinline__ void item_D(const PfArgs &p, int j, double *S, double *Li, int &first_bad)
{
	const int wj = bw(p, j);
	const int bad = block_potrf_inv64<FACTOR>(wj, p.A + 64 * j + 64L * j * p.lda, p.lda, p.Dinv + 4096L * j, S, Li);
	if (FACTOR && threadIdx.x == 0 && bad && bad <= wj && first_bad == 0)
		first_bad = 64 * j + bad;
	publish(&p.fa[j + p.nb * j], (unsigned)j + 1);
}

__device__ __noinline__ void item_P(const PfArgs &p, int i, int j)
{
	if (!wait_for(p, &p.fa[j + p.nb * j], (unsigned)j + 1, &p.fa[i + p.nb * j], (unsigned)j))
		return;
	rows_apply_inv((threadIdx.x >> 6) * 16, bw(p, i), bw(p, j), p.Dinv + 4096L * j,
	               p.A + 64 * i + 64L * j * p.lda, p.lda);
	publish(&p.fa[i + p.nb * j], (unsigned)j + 1);
}

__device__ __noinline__ void item_U(const PfArgs &p, int i, int l, int j)
{
	if (!wait_for(p, &p.fa[i + p.nb * j], (unsigned)j + 1, &p.fa[l + p.nb * j], (unsigned)j + 1,
	              &p.fa[i + p.nb * l], (unsigned)j))
		return;
	const long lda = p.lda;
	upd_nt(bw(p, i), bw(p, l), bw(p, j), p.A + 64 * i + 64L * j * lda, lda, p.A + 64 * l + 64L * j * lda, lda,
	       p.A + 64 * i + 64L * l * lda, lda, i == l);
	publish(&p.fa[i + p.nb * l], (unsigned)j + 1);
}

// U(j+1,j+1,j) on the critical chain: after the bulk updates of steps < j
// (counter >= j), and not published, since only D(j+1) on this workgroup
// consumes it (the counter goes from j to j+2 when D(j+1) is done).
__device__ __noinline__ bool chain_U(const PfArgs &p, int j)
{
	if (!wait_for(p, &p.fa[(j + 1) + p.nb * (j + 1)], (unsigned)j))
		return false;
	const long lda = p.lda;
	upd_nt(bw(p, j + 1), bw(p, j + 1), bw(p, j), p.A + 64 * (j + 1) + 64L * j * lda, lda,
	       p.A + 64 * (j + 1) + 64L * j * lda, lda, p.A + 64 * (j + 1) + 64L * (j + 1) * lda, lda, true);
	__syncthreads();
	return true;
}

// ---- inverse items ---------------------------------------------------------

// W(t,t) = inv(L(t,t))^T (upper; zero below the diagonal) and Linv(t,t) =
// inv(L(t,t)) (lower; zero above), from the block inverse D(t) formed.
__device__ __noinline__ void item_WD(const PfArgs &p, int t)
{
	if (!wait_for(p, &p.fa[t + p.nb * t], (unsigned)t + 1))
		return;
	const int w = bw(p, t);
	const double *D = p.Dinv + 4096L * t;
	double *Wb = p.W + 64 * t + 64L * t * p.ldw;
	for (int idx = threadIdx.x; idx < 4096; idx += 256) {
		const int i = idx & 63, j = idx >> 6;
		if (i < w && j < w) {
			Wb[i + (long)j * p.ldw] = i <= j ? D[j + 64 * i] : 0.0;
			if (p.Linv)
				p.Linv[(64 * t + i) + (64L * t + j) * p.ldi] = i >= j ? D[i + 64 * j] : 0.0;
		}
	}
	publish(&p.fw[t + p.nb * t], 1u);
}

// W(j,i) for j < i: the sum of W(j,t) L(i,t)^T over t = j..i-1 in
// registers, each product taken as soon as its two blocks are final, then
// W(j,i) = -sum inv(L(i,i))^T in place, with Linv(i,j) = W(j,i)^T; the
// mirrored blocks W(i,j) and Linv(j,i) are zero.
__device__ __noinline__ void item_WX(const PfArgs &p, int j, int i)
{
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
	const int i0 = 32 * (wave & 1), j0 = 32 * (wave >> 1);
	const int wr = bw(p, j), wc = bw(p, i);
	sa_d4 acc[2][2];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++)
			acc[a][b] = sa_d4{0.0, 0.0, 0.0, 0.0};
	for (int t = j; t < i; t++) {
		if (!wait_for(p, &p.fw[j + p.nb * t], 1u, &p.fa[i + p.nb * t], (unsigned)t + 1))
			return;
		acc_nt(wr, wc, bw(p, t), p.W + 64 * j + 64L * t * p.ldw, p.ldw, p.A + 64 * i + 64L * t * p.lda, p.lda,
		       acc);
	}
	if (!wait_for(p, &p.fa[i + p.nb * i], (unsigned)i + 1))
		return;
	double *Wb = p.W + 64 * j + 64L * i * p.ldw;
#pragma unroll
	for (int bj = 0; bj < 2; bj++)
#pragma unroll
		for (int q = 0; q < 4; q++)
#pragma unroll
			for (int bi = 0; bi < 2; bi++) {
				const int ic = i0 + 16 * bi + r, jc = j0 + 16 * bj + 4 * q + g;
				if (ic < wr && jc < wc)
					Wb[ic + (long)jc * p.ldw] = -acc[bj][bi][q];
			}
	// the block's rows, written by all four waves, are read back by each
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
	__syncthreads();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
	rows_apply_inv_t<double>(wave * 16, wr, wc, p.Dinv + 4096L * i, Wb, p.ldw,
	                         p.Linv ? p.Linv + 64 * i + 64L * j * p.ldi : nullptr, p.ldi);
	for (int idx = threadIdx.x; idx < 4096; idx += 256) {
		const int a = idx & 63, c = idx >> 6;
		if (a < wc && c < wr)
			p.W[(64 * i + a) + (64L * j + c) * p.ldw] = 0.0;
		if (p.Linv && a < wr && c < wc)
			p.Linv[(64 * j + a) + (64L * i + c) * p.ldi] = 0.0;
	}
	publish(&p.fw[j + p.nb * i], 1u);
}

__device__ void run_chain(const PfArgs &p, double *S, double *Li)
{
	int first_bad = 0;
	for (int j = 0; j < p.nb; j++) {
		item_D<true>(p, j, S, Li, first_bad);
		if (j + 1 < p.nb) {
			item_P(p, j + 1, j);
			if (!chain_U(p, j))
				break;
		}
	}
	if (aborted(p))
		first_bad = -1;
	if (threadIdx.x == 0)
		*p.info = first_bad;
}

// ---- the bulk order --------------------------------------------------------
// Step s: P(a, s) for a = s+2..nb-1 and U(b, a, s) for a = s+1..nb-1,
// b = a..nb-1 without the chain's (s+1, s+1) (factorisation, s < nb), then
// for t = s - delay: WD(t), WX(j, t) for j < t (inverse).  Without the factorisation, step 0 starts with D(a), a < nb.
// Host and device walk it with the same code (the host counts the items).

enum { PH_D = 0, PH_P, PH_U, PH_WD, PH_WX, PH_END };

struct PfOrder {
	int nb, factor, inv, delay;
};

struct Cursor {
	int s, ph, a, b;
};

__host__ __device__ inline bool phase_on(const PfOrder &o, int s, int ph)
{
	const int t = s - o.delay;
	switch (ph) {
	case PH_D:
		return !o.factor && s == 0;
	case PH_P:
	case PH_U:
		return o.factor && s < o.nb;
	default:
		return o.inv && t >= 0 && t < o.nb;
	}
}

__host__ __device__ inline void phase_begin(Cursor &c, const PfOrder &o)
{
	(void)o;
	c.a = c.b = 0;
	if (c.ph == PH_P)
		c.a = c.s + 2;
	else if (c.ph == PH_U)
		c.a = c.b = c.s + 1;
}

// Moves c to the first item at or after its position; false past the end.
__host__ __device__ inline bool cursor_norm(Cursor &c, const PfOrder &o)
{
	const int last = o.nb - 1 + (o.factor && o.inv ? o.delay : 0);
	for (;;) {
		if (c.s > last)
			return false;
		if (c.ph >= PH_END) {
			c.s++;
			c.ph = 0;
			phase_begin(c, o);
			continue;
		}
		if (phase_on(o, c.s, c.ph)) {
			const int s = c.s, t = s - o.delay;
			switch (c.ph) {
			case PH_D:
			case PH_P:
				if (c.a < o.nb)
					return true;
				break;
			case PH_U:
				if (c.a < o.nb) {
					if (c.b >= o.nb) {
						c.a++;
						c.b = c.a;
						continue;
					}
					if (c.a == s + 1 && c.b == s + 1) {
						c.b++;
						continue;
					}
					return true;
				}
				break;
			case PH_WD:
				if (c.a < 1)
					return true;
				break;
			case PH_WX:
				if (c.a < t)
					return true;
				break;
			}
		}
		c.ph++;
		phase_begin(c, o);
	}
}

__host__ __device__ inline void cursor_next(Cursor &c)
{
	if (c.ph == PH_U)
		c.b++;
	else
		c.a++;
}

__host__ __device__ inline Cursor cursor_first(const PfOrder &o, bool *valid)
{
	Cursor c = {0, 0, 0, 0};
	phase_begin(c, o);
	*valid = cursor_norm(c, o);
	return c;
}

// The bulk workgroups take items in the bulk order from a shared ticket
// counter: whichever workgroup is free takes the next item, so an item is
// never queued behind another workgroup's unrelated work.  Tickets only grow
// and every item waits only for items earlier in the order, so the earliest
// unfinished item is always held by a running workgroup with its inputs
// ready; an item not yet handed out blocks nobody.
__device__ void run_bulk(const PfArgs &p, double *S, double *Li, int *s_ticket)
{
	const PfOrder o = {p.nb, p.factor, p.inv, p.delay};
	bool valid;
	Cursor c = cursor_first(o, &valid);
	int pos = 0, unused = 0;
	for (;;) {
		if (threadIdx.x == 0)
			*s_ticket = (int)atomicAdd(p.ticket, 1u);
		__syncthreads();
		const int t = *s_ticket;
		__syncthreads();
		while (valid && pos < t) {
			cursor_next(c);
			valid = cursor_norm(c, o);
			pos++;
		}
		if (!valid || aborted(p))
			break;                        // past the last item: this workgroup is done
		const int s = c.s, ts = c.s - p.delay;
		switch (c.ph) {
		case PH_D:
			item_D<false>(p, c.a, S, Li, unused);
			break;
		case PH_P:
			item_P(p, c.a, s);
			break;
		case PH_U:
			item_U(p, c.b, c.a, s);
			break;
		case PH_WD:
			item_WD(p, ts);
			break;
		default:
			item_WX(p, c.a, ts);
			break;
		}
	}
}

__global__ __launch_bounds__(256)
void k_potrf_tile(PfArgs p)
{
	__shared__ double S[64 * LD];
	__shared__ double Li[64 * LD];
	__shared__ int s_ticket;
	if (p.factor && blockIdx.x == 0)
		run_chain(p, S, Li);
	else
		run_bulk(p, S, Li, &s_ticket);
}

// Counters of a launch: those of A's blocks zero when the tile is factored
// here; when it is factored already, the off-diagonal blocks of L are final
// (count l + 1) and the diagonal ones wait for their D item.  W's counters,
// the ticket and the abort flag zero.  A kernel, not a memset, so that a
// captured graph replays it correctly.
__global__ void k_init_counters(unsigned *fa, unsigned *fw, int nb, int factor)
{
	for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
		const int i = idx % nb, l = idx / nb;
		fa[idx] = (!factor && i > l) ? (unsigned)l + 1 : 0u;
		fw[idx] = 0u;
	}
	if (threadIdx.x == 0) {
		fa[nb * nb] = 0u;       // ticket
		fa[nb * nb + 1] = 0u;   // abort
	}
}

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace of a tile: the 64x64 block inverses, A's block counters with
// the ticket and the abort flag, W's block counters.
constexpr size_t tile_bytes(int n)
{
	return align256((size_t)(n > 0 ? (n + 63) / 64 : 1) * 4096 * sizeof(double)) +
	       align256(((size_t)(n > 0 ? (n + 63) / 64 : 1) * (n > 0 ? (n + 63) / 64 : 1) + 2) * sizeof(unsigned)) +
	       align256((size_t)(n > 0 ? (n + 63) / 64 : 1) * (n > 0 ? (n + 63) / 64 : 1) * sizeof(unsigned));
}

long count_items(const PfOrder &o)
{
	bool valid;
	Cursor c = cursor_first(o, &valid);
	long k = 0;
	while (valid) {
		k++;
		cursor_next(c);
		valid = cursor_norm(c, o);
	}
	return k;
}

// factor: A is factored in place (info receives the pivot report); else A
// holds the factor.  W (and Linv, optional) receive the inverse when W is
// given.  bulk: workgroups for the bulk items (<= 0: max_bulk()).
int launch_tile(hipStream_t s, int n, double *A, int lda, int *dev_info, void *work, double *W, int ldw,
                double *Linv, int ldi, int factor, int bulk = 0)
{
	const int nb = n > 0 ? (n + 63) / 64 : 0;
	char *ws = (char *)work;
	PfArgs p;
	p.n = n;
	p.nb = nb;
	p.factor = factor;
	p.inv = W != nullptr;
	p.delay = factor ? 1 : 0;
	p.A = A;
	p.lda = lda;
	p.W = W;
	p.ldw = ldw;
	p.Linv = Linv;
	p.ldi = ldi;
	const size_t dinv = align256((size_t)(nb ? nb : 1) * 4096 * sizeof(double));
	const size_t cnt = align256(((size_t)(nb ? nb : 1) * (nb ? nb : 1) + 2) * sizeof(unsigned));
	p.Dinv = (double *)ws;
	p.fa = (unsigned *)(ws + dinv);
	p.ticket = p.fa + (long)nb * nb;
	p.abort = p.ticket + 1;
	p.fw = (unsigned *)(ws + dinv + cnt);
	p.info = dev_info;
	if (nb == 0) {
		if (!factor)
			return 0;
		// n = 0: reports info = 0
		p.nbulk = 0;
		hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(64), 0, s, p.fa, p.fw, 1, 1);
		int e = sa_launch_status();
		if (e)
			return e;
		hipLaunchKernelGGL(k_potrf_tile, dim3(1), dim3(256), 0, s, p);
		return sa_launch_status();
	}
	const long items = count_items(PfOrder{nb, factor, p.inv, p.delay});
	const int want = bulk > 0 ? (bulk < 255 ? bulk : 255) : max_bulk();
	p.nbulk = (int)(items < want ? items : want);
	hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(256), 0, s, p.fa, p.fw, nb, factor);
	int e = sa_launch_status();
	if (e)
		return e;
	const int grid = (factor ? 1 : 0) + p.nbulk;
	if (grid == 0)
		return 0;
	hipLaunchKernelGGL(k_potrf_tile, dim3((unsigned)grid), dim3(256), 0, s, p);
	return sa_launch_status();
}

} // namespace

// Workspace: the block inverses and the counters, or for tiles past the
// single-launch limit the substitution inverse's.
extern "C" size_t sa_dpotrf_tile_worksize(int n)
{
	if (n <= PF_MAXN)
		return tile_bytes(n);
	return sa_dtrtri_lt_worksize(n);
}

extern "C" int sa_dpotrf_tile_maxn(void) { return PF_MAXN; }

extern "C" int sa_dpotrf_tile_wg(hipStream_t s, int n, double *A, int lda, double *Linv, int ldl, double *W,
                                 int ldw, int *dev_info, void *work, size_t work_bytes, int workgroups)
{
	if (n < 0 || n > PF_MAXN || lda < (n > 1 ? n : 1) || !dev_info || !work ||
	    work_bytes < sa_dpotrf_tile_worksize(n) || (Linv && !W) ||
	    (W && ldw < (n > 1 ? n : 1)) || (Linv && ldl < (n > 1 ? n : 1)))
		return -EINVAL;
	return launch_tile(s, n, A, lda, dev_info, work, W, ldw, Linv, ldl, 1, workgroups);
}

extern "C" int sa_dpotrf_tile(hipStream_t s, int n, double *A, int lda, double *Linv, int ldl, double *W, int ldw,
                              int *dev_info, void *work, size_t work_bytes)
{
	return sa_dpotrf_tile_wg(s, n, A, lda, Linv, ldl, W, ldw, dev_info, work, work_bytes, 0);
}

extern "C" int sa_dtrtri_tile(hipStream_t s, int n, const double *L, int ldl, double *Linv, int ldx, double *W,
                              int ldw, void *work, size_t work_bytes)
{
	if (n < 0 || ldl < (n > 1 ? n : 1) || !Linv || !W || ldx < (n > 1 ? n : 1) || ldw < (n > 1 ? n : 1) ||
	    !work || work_bytes < sa_dpotrf_tile_worksize(n > 0 ? n : 1))
		return -EINVAL;
	if (n == 0)
		return 0;
	if (n <= PF_MAXN)
		return launch_tile(s, n, const_cast<double *>(L), ldl, nullptr, work, W, ldw, Linv, ldx, 0);
	int e = sa_dtrtri_lt(s, n, L, ldl, W, ldw, work, work_bytes);
	return e ? e : sa_dtranspose(s, n, n, W, ldw, Linv, ldx);
}

// ---------------------------------------------------------------------------
// Allocation-free tile factorisation (sa_dpotrf_tile_noalloc).  Its
// workspace -- the 64x64 block inverses and the counters of the largest
// tile accepted -- is this file-scope device array: each GPU the library
// runs on holds one copy until the process ends.  Host code never passes
// `scratch` itself to a kernel (on the host the name is only a stand-in);
// it uses the device address hipGetSymbolAddress returns, looked up once per
// device.  Overlapping calls on one GPU share the array and race; the
// counters are cleared by a kernel (k_init_counters) before every
// factorisation, so a call does not depend on what the previous one left
// there.

constexpr int NOALLOC_MAXN = PF_MAXN;
constexpr size_t NOALLOC_BYTES = tile_bytes(NOALLOC_MAXN);
constexpr int NOALLOC_MAXDEV = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[NOALLOC_MAXDEV];

// Device address of `scratch` on the current device (looked up on the first
// call per device), or null.
static void *scratch_address(int *err)
{
	int dev = -1;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess || dev < 0 || dev >= NOALLOC_MAXDEV) {
		*err = e != hipSuccess ? (int)e : -ENODEV;
		return nullptr;
	}
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_dev[dev]) {
		void *p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) {
			*err = e != hipSuccess ? (int)e : -ENODEV;
			return nullptr;
		}
		g_scratch_dev[dev] = p;
	}
	return g_scratch_dev[dev];
}

extern "C" int sa_dpotrf_tile_noalloc_maxn(void) { return NOALLOC_MAXN; }

extern "C" int sa_dpotrf_tile_noalloc(hipStream_t s, int n, double *A, int lda, double *Linv, int ldl, double *W,
                                      int ldw, int *dev_info)
{
	if (n > NOALLOC_MAXN || sa_dpotrf_tile_worksize(n) > NOALLOC_BYTES)
		return -EOVERFLOW;
	int err = 0;
	void *ws = scratch_address(&err);
	if (!ws)
		return err;
	return sa_dpotrf_tile(s, n, A, lda, Linv, ldl, W, ldw, dev_info, ws, NOALLOC_BYTES);
}
