// Cholesky factorisation of one diagonal tile (n <= 2048) in a single
// persistent kernel (gfx950).
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
// 64b .. 64b+w_b-1 of the tile):
//   D(j)      factor and invert the diagonal block (j, j)        (LDS, MFMA)
//   P(i,j)    L(i,j) = A(i,j) inv(L(j,j))^T                          (MFMA)
//   U(i,l,j)  A(i,l) -= L(i,j) L(l,j)^T, j < l <= i (lower part if i = l)
//
// Ordering and progress.  Every item depends only on items that come earlier
// in one fixed global order: step by step, the critical chain D(j), P(j+1,j),
// U(j+1,j+1,j), then the other items of step j.  Workgroup 0 runs the
// critical chain; the other workgroups take the remaining items in that order
// from a ticket counter.  Each workgroup processes its items in increasing
// order, so the lowest unfinished item always has all its inputs ready and a
// running workgroup holds it: the grid always drains.  Block (i, l) carries a
// counter: the number of updates applied, then +1 when the block is final.
// The updates of a block are applied in step order, so the result does not
// depend on timing.
//
// Visibility across workgroups (possibly on different XCDs, each with its own
// L2) uses agent-scope release fences before a counter is raised and an
// agent-scope acquire after a wait.
//
// The inverse of the factor (sa_dpotrf_tile with Linv and W, sa_dtrtri_tile)
// is formed afterwards by the substitution kernel of sa_dtrtri_lt and a
// transpose, the same kernels a solve on a resident factor uses.
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
	double *A;
	long lda;
	double *Dinv;       // nb blocks of 64x64 (ld 64)
	unsigned *fa;       // counters of the blocks of A: index i + nb * l
	unsigned *ticket;   // next item of the bulk order to hand out
	int *info;
};

__device__ __forceinline__ int bw(const PfArgs &p, int b) { return (p.n - 64 * b) < 64 ? (p.n - 64 * b) : 64; }

__device__ __forceinline__ void spin(const unsigned *f, unsigned v)
{
	while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v)
		__builtin_amdgcn_s_sleep(1);
}

// Thread 0 waits until each given counter reaches its value; the whole
// workgroup then sees the data published before those counters were raised.
__device__ __forceinline__ void wait_for(const unsigned *f1, unsigned v1, const unsigned *f2 = nullptr,
                                         unsigned v2 = 0, const unsigned *f3 = nullptr, unsigned v3 = 0)
{
	if (threadIdx.x == 0) {
		if (f1)
			spin(f1, v1);
		if (f2)
			spin(f2, v2);
		if (f3)
			spin(f3, v3);
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
	}
	__syncthreads();
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

// C (wr x wc, ld ldc) -= Pr Pc^T over k < wk, Pr (wr x wk, ld ldr), Pc
// (wc x wk, ld ldc2); lower part only if LOWER.  Each wave a 32x32 quadrant;
// the MFMA computes the transposed product so a lane's results are
// consecutive rows of C.
__device__ __forceinline__ void upd_nt(int wr, int wc, int wk, const double *__restrict__ Pr, long ldr,
                                       const double *__restrict__ Pc, long ldpc, double *__restrict__ C, long ldc,
                                       bool lower)
{
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
	const int i0 = 32 * (wave & 1), j0 = 32 * (wave >> 1);
	if (lower && j0 > i0 + 31)
		return;
	// operand loads in two batches of half the k range, the C loads with the
	// first: the blocks sit in L2 and the item is latency-bound
	double xa[2][8], xb[2][8], cv[2][4][2];
	sa_d4 acc[2][2];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++)
			acc[a][b] = sa_d4{0.0, 0.0, 0.0, 0.0};
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
		if (h == 0) {
#pragma unroll
			for (int bj = 0; bj < 2; bj++)
#pragma unroll
				for (int q = 0; q < 4; q++)
#pragma unroll
					for (int bi = 0; bi < 2; bi++) {
						const int ic = i0 + 16 * bi + r, jc = j0 + 16 * bj + 4 * q + g;
						cv[bj][q][bi] = C[(ic < wr ? ic : 0) + (long)(jc < wc ? jc : 0) * ldc];
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
	// acc[bj][bi][q] = (Pr Pc^T)(i0 + 16 bi + r, j0 + 16 bj + 4q + g)
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

// ---- items -----------------------------------------------------------------

// The chain applies the last update of block (j, j) itself just before
// D(j) (chain_U), so D(j) has no wait of its own.
__device__ __noinline__ void item_D(const PfArgs &p, int j, double *S, double *Li, int &first_bad)
{
	const int wj = bw(p, j);
	const int bad = block_potrf_inv64<true>(wj, p.A + 64 * j + 64L * j * p.lda, p.lda, p.Dinv + 4096L * j, S, Li);
	if (threadIdx.x == 0 && bad && bad <= wj && first_bad == 0)
		first_bad = 64 * j + bad;
	publish(&p.fa[j + p.nb * j], (unsigned)j + 1);
}

__device__ __noinline__ void item_P(const PfArgs &p, int i, int j)
{
	wait_for(&p.fa[j + p.nb * j], (unsigned)j + 1, &p.fa[i + p.nb * j], (unsigned)j);
	rows_apply_inv((threadIdx.x >> 6) * 16, bw(p, i), bw(p, j), p.Dinv + 4096L * j,
	               p.A + 64 * i + 64L * j * p.lda, p.lda);
	publish(&p.fa[i + p.nb * j], (unsigned)j + 1);
}

__device__ __noinline__ void item_U(const PfArgs &p, int i, int l, int j)
{
	wait_for(&p.fa[i + p.nb * j], (unsigned)j + 1, &p.fa[l + p.nb * j], (unsigned)j + 1, &p.fa[i + p.nb * l],
	         (unsigned)j);
	const long lda = p.lda;
	upd_nt(bw(p, i), bw(p, l), bw(p, j), p.A + 64 * i + 64L * j * lda, lda, p.A + 64 * l + 64L * j * lda, lda,
	       p.A + 64 * i + 64L * l * lda, lda, i == l);
	publish(&p.fa[i + p.nb * l], (unsigned)j + 1);
}

// U(j+1,j+1,j) on the critical chain: after the bulk updates of steps < j
// (counter >= j), and not published, since only D(j+1) on this workgroup
// consumes it (the counter goes from j to j+2 when D(j+1) is done).
__device__ __noinline__ void chain_U(const PfArgs &p, int j)
{
	wait_for(&p.fa[(j + 1) + p.nb * (j + 1)], (unsigned)j);
	const long lda = p.lda;
	upd_nt(bw(p, j + 1), bw(p, j + 1), bw(p, j), p.A + 64 * (j + 1) + 64L * j * lda, lda,
	       p.A + 64 * (j + 1) + 64L * j * lda, lda, p.A + 64 * (j + 1) + 64L * (j + 1) * lda, lda, true);
	__syncthreads();
}

__device__ void run_chain(const PfArgs &p, double *S, double *Li)
{
	int first_bad = 0;
	for (int j = 0; j < p.nb; j++) {
		item_D(p, j, S, Li, first_bad);
		if (j + 1 < p.nb) {
			item_P(p, j + 1, j);
			chain_U(p, j);
		}
	}
	if (threadIdx.x == 0)
		*p.info = first_bad;
}

// Position in the bulk order: phase 0 P(a, j) for a = j+2..nb-1; phase 1
// U(b, a, j) for columns a = j+1..nb-1 and rows b = a..nb-1, without the
// chain's (j+1, j+1); steps j = 0..nb-1.
struct Cursor {
	int j, ph, a, b;
};

// Moves c to the first item at or after its position; false past the end.
__device__ bool cursor_norm(Cursor &c, int nb)
{
	for (;;) {
		if (c.j >= nb)
			return false;
		if (c.ph == 0) {
			if (c.a < nb)
				return true;
			c.ph = 1;
			c.a = c.j + 1;
			c.b = c.j + 1;
		} else if (c.a < nb) {
			if (c.b >= nb) {
				c.a++;
				c.b = c.a;
			} else if (c.a == c.j + 1 && c.b == c.j + 1) {
				c.b++;
			} else {
				return true;
			}
		} else {
			c.j++;
			c.ph = 0;
			c.a = c.j + 2;
		}
	}
}

__device__ __forceinline__ void cursor_next(Cursor &c)
{
	if (c.ph == 0)
		c.a++;
	else
		c.b++;
}

// The bulk workgroups take items in the bulk order from a shared ticket
// counter: whichever workgroup is free takes the next item, so an item is
// never queued behind another workgroup's unrelated work.  Tickets only grow
// and every item waits only for items earlier in the order, so the earliest
// unfinished item is always held by a running workgroup with its inputs
// ready; an item not yet handed out blocks nobody.
__device__ void run_bulk(const PfArgs &p, int *s_ticket)
{
	Cursor c = {0, 0, 2, 0};
	bool valid = cursor_norm(c, p.nb);
	int pos = 0;
	for (;;) {
		if (threadIdx.x == 0)
			*s_ticket = (int)atomicAdd(p.ticket, 1u);
		__syncthreads();
		const int t = *s_ticket;
		__syncthreads();
		while (valid && pos < t) {
			cursor_next(c);
			valid = cursor_norm(c, p.nb);
			pos++;
		}
		if (!valid)
			break;                        // past the last item: this workgroup is done
		if (c.ph == 0)
			item_P(p, c.a, c.j);
		else
			item_U(p, c.b, c.a, c.j);
	}
}

__global__ __launch_bounds__(256)
void k_potrf_tile(PfArgs p)
{
	__shared__ double S[64 * LD];
	__shared__ double Li[64 * LD];
	__shared__ int s_ticket;
	if (blockIdx.x == 0)
		run_chain(p, S, Li);
	else
		run_bulk(p, &s_ticket);
}

__global__ void k_zero_u32(unsigned *f, int count)
{
	for (int i = threadIdx.x; i < count; i += blockDim.x)
		f[i] = 0u;
}

// number of items outside the critical chain
long bulk_items(int nb)
{
	long c = 0;
	for (int j = 0; j < nb; j++) {
		const long m = nb - j - 1;
		c += (nb - j - 2 > 0 ? nb - j - 2 : 0) + (m > 0 ? m * (m + 1) / 2 - 1 : 0);
	}
	return c;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t tile_factor_bytes(int n)
{
	const int nb = n > 0 ? (n + 63) / 64 : 1;
	return align256((size_t)nb * 4096 * sizeof(double)) + align256(((size_t)nb * nb + 1) * sizeof(unsigned));
}

int launch_tile(hipStream_t s, int n, double *A, int lda, int *dev_info, void *work)
{
	const int nb = n > 0 ? (n + 63) / 64 : 0;
	char *ws = (char *)work;
	PfArgs p;
	p.n = n;
	p.nb = nb;
	p.A = A;
	p.lda = lda;
	p.Dinv = (double *)ws;
	p.fa = (unsigned *)(ws + align256((size_t)(nb ? nb : 1) * 4096 * sizeof(double)));
	p.ticket = p.fa + (long)nb * nb;
	p.info = dev_info;
	const long items = bulk_items(nb);
	p.nbulk = (int)(items < max_bulk() ? items : max_bulk());
	if (nb == 0) {
		// n = 0: reports info = 0
		hipLaunchKernelGGL(k_potrf_tile, dim3(1), dim3(256), 0, s, p);
		return sa_launch_status();
	}
	hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, p.fa, nb * nb + 1);
	int e = sa_launch_status();
	if (e)
		return e;
	hipLaunchKernelGGL(k_potrf_tile, dim3((unsigned)(1 + p.nbulk)), dim3(256), 0, s, p);
	return sa_launch_status();
}

} // namespace

// Workspace: the block inverses, the block counters and the ticket counter
// (cleared by k_zero_u32 before each launch), or the inverse's workspace.
extern "C" size_t sa_dpotrf_tile_worksize(int n)
{
	const size_t f = tile_factor_bytes(n), t = sa_dtrtri_lt_worksize(n > 0 ? n : 1);
	return f > t ? f : t;
}

extern "C" int sa_dpotrf_tile(hipStream_t s, int n, double *A, int lda, double *Linv, int ldl, double *W, int ldw,
                              int *dev_info, void *work, size_t work_bytes)
{
	if (n < 0 || n > PF_MAXN || lda < (n > 1 ? n : 1) || !dev_info || !work ||
	    work_bytes < sa_dpotrf_tile_worksize(n) || (!Linv) != (!W) ||
	    (Linv && (ldl < (n > 1 ? n : 1) || ldw < (n > 1 ? n : 1))))
		return -EINVAL;
	int e = launch_tile(s, n, A, lda, dev_info, work);
	if (e || !Linv || n == 0)
		return e;
	return sa_dtrtri_tile(s, n, A, lda, Linv, ldl, W, ldw, work, work_bytes);
}

extern "C" int sa_dtrtri_tile(hipStream_t s, int n, const double *L, int ldl, double *Linv, int ldx, double *W,
                              int ldw, void *work, size_t work_bytes)
{
	if (n < 0 || ldl < (n > 1 ? n : 1) || !Linv || !W || ldx < (n > 1 ? n : 1) || ldw < (n > 1 ? n : 1) ||
	    !work || work_bytes < sa_dtrtri_lt_worksize(n > 0 ? n : 1))
		return -EINVAL;
	if (n == 0)
		return 0;
	int e = sa_dtrtri_lt(s, n, L, ldl, W, ldw, work, work_bytes);
	return e ? e : sa_dtranspose(s, n, n, W, ldw, Linv, ldx);
}

// ---------------------------------------------------------------------------
// Allocation-free tile factorisation (sa_dpotrf_tile_noalloc).  Its
// workspace -- the 64x64 block inverses, the block counters and the ticket
// of the largest tile accepted, or the inverse's staging -- is this
// file-scope device array: each GPU the library runs on holds one copy
// until the process ends.  Host code never passes `scratch` itself to a
// kernel (on the host the name is only a stand-in); it uses the device
// address hipGetSymbolAddress returns, looked up once per device.
// Overlapping calls on one GPU share the array and race; the counters are
// cleared by a kernel (k_zero_u32) before every factorisation, so a call
// does not depend on what the previous one left there.

constexpr int NOALLOC_MAXN = PF_MAXN;
constexpr size_t NOALLOC_BYTES =
	((size_t)(NOALLOC_MAXN / 64) * 4096 * sizeof(double) + 255) / 256 * 256 +
	(((size_t)(NOALLOC_MAXN / 64) * (NOALLOC_MAXN / 64) + 1) * sizeof(unsigned) + 255) / 256 * 256;
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
