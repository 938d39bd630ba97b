// Device-memory helpers, timers and synthetic-matrix generators behind the
// C ABI of starpu_amd_kernels.h.
#include "sa_common.h"
#include "starpu_amd_kernels.h"

namespace {

// Counter-based generator: splitmix64 finaliser of (seed, i, j).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t h)   // in (0, 1)
{
	return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__global__ void k_fill_random(int m, int n, double *X, long ld, uint64_t seed, int dist)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)m * n)
		return;
	const int i = (int)(idx % m), j = (int)(idx / m);
	const uint64_t key = mix64(seed ^ mix64(((uint64_t)(uint32_t)j << 32) | (uint32_t)i));
	double v;
	if (dist == 0) {
		v = 2.0 * u01(mix64(key)) - 1.0;
	} else {
		const double u1 = u01(mix64(key));
		const double u2 = u01(mix64(key ^ 0xD1B54A32D192ED03ull));
		v = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
	}
	X[i + (long)j * ld] = v;
}

__global__ void k_add_diag(int n, double *X, long ld, double shift)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		X[i + (long)i * ld] += shift;
}

__global__ void k_symmetrize(int n, double *X, long ld)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)n * n)
		return;
	const int i = (int)(idx % n), j = (int)(idx / n);
	if (i < j)
		X[i + (long)j * ld] = X[j + (long)i * ld];
}

// B(m x n) <- A(m x n), both column-major: workgroup (x, j) copies rows
// 256 V x .. of column j, V consecutive elements per thread (one 16-byte
// load and store when both columns are so aligned), so both sides move in
// coalesced runs and no thread divides by m.
template <typename T, int V>
__global__ __launch_bounds__(256) void k_lacpy(int m, const T *__restrict__ A, long lda, T *__restrict__ B, long ldb)
{
	typedef T TV __attribute__((ext_vector_type(V)));
	const int j = blockIdx.y;
	const int i = (blockIdx.x * 256 + threadIdx.x) * V;
	if (i >= m)
		return;
	const T *a = A + (long)j * lda + i;
	T *b = B + (long)j * ldb + i;
	if (V > 1 && i + V <= m) {
		*reinterpret_cast<TV *>(b) = *reinterpret_cast<const TV *>(a);
	} else {
		for (int e = 0; e < V && i + e < m; e++)
			b[e] = a[e];
	}
}

// columns in launches of at most 65535 (the grid's y limit)
template <typename T>
int lacpy_launch(hipStream_t s, int m, int n, const T *A, int lda, T *B, int ldb)
{
	constexpr int V = 16 / sizeof(T);
	const bool vec = (uintptr_t)A % 16 == 0 && (uintptr_t)B % 16 == 0 && lda % V == 0 && ldb % V == 0;
	for (int j0 = 0; j0 < n; j0 += 65535) {
		const int nj = (n - j0) < 65535 ? n - j0 : 65535;
		const T *a = A + (long)j0 * lda;
		T *b = B + (long)j0 * ldb;
		if (vec)
			hipLaunchKernelGGL((k_lacpy<T, V>), dim3((unsigned)((m + 256 * V - 1) / (256 * V)), (unsigned)nj), dim3(256),
			                   0, s, m, a, (long)lda, b, (long)ldb);
		else
			hipLaunchKernelGGL((k_lacpy<T, 1>), dim3((unsigned)((m + 255) / 256), (unsigned)nj), dim3(256), 0, s, m, a,
			                   (long)lda, b, (long)ldb);
		const int e = sa_launch_status();
		if (e)
			return e;
	}
	return 0;
}

// B (n x m) = A^T for A (m x n), through 64x64 LDS blocks.
__global__ __launch_bounds__(256)
void k_dtranspose(int m, int n, const double *__restrict__ A, long lda, double *__restrict__ B, long ldb)
{
	__shared__ double S[64 * 65];
	const int bi = blockIdx.x, bj = blockIdx.y;     // block of A: rows 64 bi, cols 64 bj
	const int tid = threadIdx.x;
	for (int idx = tid; idx < 64 * 64; idx += 256) {
		const int i = idx & 63, j = idx >> 6;
		const int gi = 64 * bi + i, gj = 64 * bj + j;
		S[i + 65 * j] = (gi < m && gj < n) ? A[gi + (long)gj * lda] : 0.0;
	}
	__syncthreads();
	for (int idx = tid; idx < 64 * 64; idx += 256) {
		const int i = idx & 63, j = idx >> 6;       // B block (bj, bi): row i, col j
		const int gi = 64 * bj + i, gj = 64 * bi + j;
		if (gi < n && gj < m)
			B[gi + (long)gj * ldb] = S[j + 65 * i];
	}
}

// A(i, j) = (i == j) ? diag : offdiag over an m x n region.
__global__ void k_dlaset(int m, int n, double offdiag, double diag, double *A, long lda)
{
	const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (long)m * n)
		return;
	const int i = (int)(idx % m), j = (int)(idx / m);
	A[i + (long)j * lda] = (i == j) ? diag : offdiag;
}

// out[j] = sum_i X(i, j)^2: one workgroup of 1024 threads per column, 16
// loads in flight per thread; partial sums are combined in a fixed order.
__global__ __launch_bounds__(1024)
void k_colsumsq(int m, const double *__restrict__ X, long ld, double *__restrict__ out)
{
	__shared__ double part[16];
	const int j = blockIdx.x, t = threadIdx.x;
	const double *x = X + (long)j * ld;
	double s[4] = {0.0, 0.0, 0.0, 0.0};
	int i = t;
	for (; i + 15 * 1024 < m; i += 16 * 1024) {
		double v[16];
#pragma unroll
		for (int u = 0; u < 16; u++)
			v[u] = x[i + u * 1024];
#pragma unroll
		for (int u = 0; u < 16; u++)
			s[u & 3] += v[u] * v[u];
	}
	for (; i < m; i += 1024)
		s[0] += x[i] * x[i];
	double v = (s[0] + s[1]) + (s[2] + s[3]);
	for (int off = 32; off >= 1; off >>= 1)
		v += __shfl_xor(v, off);
	if ((t & 63) == 0)
		part[t >> 6] = v;
	__syncthreads();
	if (t == 0) {
		double a = 0.0;
		for (int w = 0; w < 16; w++)
			a += part[w];
		out[j] = a;
	}
}

} // namespace

// ------------------------------------------------------------ helpers ----

extern "C" int sa_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess)
		return 0;
	return n;
}

extern "C" int sa_set_device(int dev) { return (int)hipSetDevice(dev); }

extern "C" void *sa_dev_malloc(size_t bytes)
{
	void *p = nullptr;
	if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess)
		return nullptr;
	return p;
}

extern "C" int sa_dev_free(void *p) { return (int)hipFree(p); }
extern "C" int sa_memcpy_h2d(void *d, const void *s, size_t b) { return (int)hipMemcpy(d, s, b, hipMemcpyHostToDevice); }
extern "C" int sa_memcpy_d2h(void *d, const void *s, size_t b) { return (int)hipMemcpy(d, s, b, hipMemcpyDeviceToHost); }
extern "C" int sa_memcpy_d2d(void *d, const void *s, size_t b) { return (int)hipMemcpy(d, s, b, hipMemcpyDeviceToDevice); }

extern "C" int sa_memcpy2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                                 size_t width, size_t height, int kind, hipStream_t s)
{
	return (int)hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, (hipMemcpyKind)kind, s);
}

extern "C" int sa_stream_create(hipStream_t *s) { return (int)hipStreamCreateWithFlags(s, hipStreamNonBlocking); }
extern "C" int sa_stream_destroy(hipStream_t s) { return (int)hipStreamDestroy(s); }
extern "C" int sa_stream_sync(hipStream_t s) { return (int)hipStreamSynchronize(s); }
extern "C" int sa_device_sync(void) { return (int)hipDeviceSynchronize(); }
extern "C" int sa_event_create(void **ev)
{
	hipEvent_t e;
	const hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
	if (r != hipSuccess)
		return (int)r;
	*ev = (void *)e;
	return 0;
}
extern "C" int sa_event_record(void *ev, hipStream_t s) { return (int)hipEventRecord((hipEvent_t)ev, s); }
extern "C" int sa_stream_wait_event(hipStream_t s, void *ev) { return (int)hipStreamWaitEvent(s, (hipEvent_t)ev, 0); }
extern "C" int sa_event_destroy(void *ev) { return (int)hipEventDestroy((hipEvent_t)ev); }
extern "C" const char *sa_error_string(int e) { return hipGetErrorString((hipError_t)e); }

struct sa_timer {
	hipEvent_t a, b;
	hipStream_t s;
};

extern "C" int sa_timer_start(hipStream_t s, void **timer)
{
	sa_timer *t = new sa_timer;
	t->s = s;
	hipError_t e = hipEventCreate(&t->a);
	if (e == hipSuccess)
		e = hipEventCreate(&t->b);
	if (e == hipSuccess)
		e = hipEventRecord(t->a, s);
	if (e != hipSuccess) {
		delete t;
		return (int)e;
	}
	*timer = t;
	return 0;
}

extern "C" int sa_timer_stop(void *timer, float *ms)
{
	sa_timer *t = (sa_timer *)timer;
	hipError_t e = hipEventRecord(t->b, t->s);
	if (e == hipSuccess)
		e = hipEventSynchronize(t->b);
	if (e == hipSuccess)
		e = hipEventElapsedTime(ms, t->a, t->b);
	hipEventDestroy(t->a);
	hipEventDestroy(t->b);
	delete t;
	return (int)e;
}

extern "C" int sa_fill_random(hipStream_t s, int m, int n, double *X, int ld,
                              uint64_t seed, int dist)
{
	if (m < 0 || n < 0 || ld < (m > 1 ? m : 1) || (dist != 0 && dist != 1))
		return -EINVAL;
	const long cnt = (long)m * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_fill_random, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
	                   m, n, X, (long)ld, seed, dist);
	return sa_launch_status();
}

extern "C" int sa_add_diagonal(hipStream_t s, int n, double *X, int ld, double shift)
{
	if (n < 0 || ld < (n > 1 ? n : 1))
		return -EINVAL;
	if (n == 0)
		return 0;
	hipLaunchKernelGGL(k_add_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, X, (long)ld, shift);
	return sa_launch_status();
}

extern "C" int sa_symmetrize_lower(hipStream_t s, int n, double *X, int ld)
{
	if (n < 0 || ld < (n > 1 ? n : 1))
		return -EINVAL;
	const long cnt = (long)n * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, n, X, (long)ld);
	return sa_launch_status();
}

extern "C" int sa_dlacpy(hipStream_t s, int m, int n, const double *A, int lda, double *B, int ldb)
{
	if (m < 0 || n < 0 || lda < (m > 1 ? m : 1) || ldb < (m > 1 ? m : 1))
		return -EINVAL;
	const long cnt = (long)m * n;
	if (cnt == 0)
		return 0;
	return lacpy_launch<double>(s, m, n, A, lda, B, ldb);
}

extern "C" int sa_slacpy(hipStream_t s, int m, int n, const float *A, int lda, float *B, int ldb)
{
	if (m < 0 || n < 0 || lda < (m > 1 ? m : 1) || ldb < (m > 1 ? m : 1))
		return -EINVAL;
	if ((long)m * n == 0)
		return 0;
	return lacpy_launch<float>(s, m, n, A, lda, B, ldb);
}

extern "C" int sa_dcolsumsq(hipStream_t s, int m, int n, const double *X, int ld, double *out)
{
	if (m < 0 || n < 0 || ld < (m > 1 ? m : 1) || (n > 0 && !out))
		return -EINVAL;
	if (n == 0)
		return 0;
	hipLaunchKernelGGL(k_colsumsq, dim3((unsigned)n), dim3(1024), 0, s, m, X, (long)ld, out);
	return sa_launch_status();
}

extern "C" int sa_dlaset(hipStream_t s, int m, int n, double offdiag, double diag, double *A, int lda)
{
	if (m < 0 || n < 0 || lda < (m > 1 ? m : 1))
		return -EINVAL;
	const long cnt = (long)m * n;
	if (cnt == 0)
		return 0;
	hipLaunchKernelGGL(k_dlaset, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m, n, offdiag, diag, A,
	                   (long)lda);
	return sa_launch_status();
}

extern "C" int sa_dtranspose(hipStream_t s, int m, int n, const double *A, int lda, double *B, int ldb)
{
	if (m < 0 || n < 0 || lda < (m > 1 ? m : 1) || ldb < (n > 1 ? n : 1))
		return -EINVAL;
	if (m == 0 || n == 0)
		return 0;
	hipLaunchKernelGGL(k_dtranspose, dim3((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64)), dim3(256), 0, s, m,
	                   n, A, (long)lda, B, (long)ldb);
	return sa_launch_status();
}
