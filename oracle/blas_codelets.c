/*
 * CPU tile codelets on an external BLAS/LAPACK (TEST INFRASTRUCTURE ONLY:
 * the CPU-baseline leg of the benchmark and the tests that check it).
 *
 * The reference's CPU codelets call BLAS and LAPACK through
 * examples/common/blas.c:240-427 (STARPU_DGEMM / DSYRK / DTRSM / DPOTRF):
 *   GEMM  cholesky_kernels.c:75-76   dgemm("N","T", -1, A_mk, A_nk, 1, A_mn)
 *   SYRK  cholesky_kernels.c:169-170 dsyrk("L","N", -1, A_nk, 1, A_nn)
 *   TRSM  cholesky_kernels.c:254     dtrsm("R","L","T","N", 1, L_kk, A_mk)
 *   POTRF cholesky_kernels.c:329-342 dpotrf("L", A_kk) (LAPACK form of the
 *         unblocked loop there)
 * These codelets make the same calls, on the same tiles, with the library
 * resolved at run time (orb_load) -- in this image scipy's bundled OpenBLAS,
 * whose Fortran entry points carry a scipy_ prefix.  Each codelet runs on
 * one worker thread and the library is set to one thread, as StarPU runs
 * sequential BLAS per CPU worker.
 *
 * The DAG's POTRF task also writes W_k = L_kk^-T and Linv_k = L_kk^-1 (the
 * inputs of the GPU TRSM and solve tasks): dtrsm of the identity and a
 * transpose, nb^3 more flops per diagonal tile than the reference's POTRF.
 * The solve tasks use dtrsm/dgemm on L_kk and the panels, as LAPACK dpotrs.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stddef.h>
#include <string.h>

#include <starpu.h>

typedef void (*dgemm_t)(const char *, const char *, const int *, const int *, const int *, const double *,
                        const double *, const int *, const double *, const int *, const double *, double *,
                        const int *, size_t, size_t);
typedef void (*dsyrk_t)(const char *, const char *, const int *, const int *, const double *, const double *,
                        const int *, const double *, double *, const int *, size_t, size_t);
typedef void (*dtrsm_t)(const char *, const char *, const char *, const char *, const int *, const int *,
                        const double *, const double *, const int *, double *, const int *, size_t, size_t,
                        size_t, size_t);
typedef void (*dpotrf_t)(const char *, const int *, double *, const int *, int *, size_t);
typedef void (*set_threads_t)(int);

static dgemm_t p_dgemm;
static dsyrk_t p_dsyrk;
static dtrsm_t p_dtrsm;
static dpotrf_t p_dpotrf;

static void *sym(void *h, const char *name)
{
	char buf[64];
	void *f = dlsym(h, name);
	if (!f) {
		strcpy(buf, "scipy_");
		strncat(buf, name, sizeof(buf) - 8);
		f = dlsym(h, buf);
	}
	return f;
}

/* Loads the BLAS/LAPACK library at `path` (entry points dgemm_ or
 * scipy_dgemm_ and so on) and sets it to one thread.  Returns 0, or -1 if
 * the library or one of the four routines is missing. */
int orb_load(const char *path)
{
	void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
	if (!h)
		return -1;
	p_dgemm = (dgemm_t)sym(h, "dgemm_");
	p_dsyrk = (dsyrk_t)sym(h, "dsyrk_");
	p_dtrsm = (dtrsm_t)sym(h, "dtrsm_");
	p_dpotrf = (dpotrf_t)sym(h, "dpotrf_");
	set_threads_t st = (set_threads_t)dlsym(h, "scipy_openblas_set_num_threads");
	if (!st)
		st = (set_threads_t)dlsym(h, "openblas_set_num_threads");
	if (st)
		st(1);
	return (p_dgemm && p_dsyrk && p_dtrsm && p_dpotrf) ? 0 : -1;
}

#define MPTR(i) ((double *)STARPU_MATRIX_GET_PTR(descr[i]))
#define MNX(i) ((int)STARPU_MATRIX_GET_NX(descr[i]))
#define MNY(i) ((int)STARPU_MATRIX_GET_NY(descr[i]))
#define MLD(i) ((int)STARPU_MATRIX_GET_LD(descr[i]))

static const double one = 1.0, mone = -1.0;

/* A_kk <- L_kk (dpotrf), W_k <- L_kk^-T (dtrsm R,L,T,N of the identity),
 * Linv_k <- W_k^T (RW A_kk, W info_k, W W_k, W Linv_k, scratch) */
void orb_cl_potrf(void *descr[], void *arg)
{
	(void)arg;
	const int n = MNY(0), lda = MLD(0), ldw = MLD(2), ldi = MLD(3);
	double *A = MPTR(0), *W = MPTR(2), *Li = MPTR(3);
	int *info = (int *)STARPU_VECTOR_GET_PTR(descr[1]);
	int inf = 0;
	p_dpotrf("L", &n, A, &lda, &inf, 1);
	*info = inf;
	for (int j = 0; j < n; j++)
		for (int i = 0; i < n; i++)
			W[i + (size_t)j * ldw] = i == j ? 1.0 : 0.0;
	if (inf != 0)
		return;
	p_dtrsm("R", "L", "T", "N", &n, &n, &one, A, &lda, W, &ldw, 1, 1, 1, 1);
	for (int j = 0; j < n; j++)
		for (int i = 0; i < n; i++)
			Li[i + (size_t)j * ldi] = W[j + (size_t)i * ldw];
}

/* A_mk <- A_mk L_kk^-T (R L_kk, R Linv_k, RW A_mk, scratch); m = NX and
 * n = NY of A_mk (the reference passes NY, NX, right only for square tiles) */
void orb_cl_trsm(void *descr[], void *arg)
{
	(void)arg;
	const int m = MNX(2), n = MNY(2), ldl = MLD(0), ldb = MLD(2);
	p_dtrsm("R", "L", "T", "N", &m, &n, &one, MPTR(0), &ldl, MPTR(2), &ldb, 1, 1, 1, 1);
}

/* lower A_nn -= A_nk A_nk^T (R A_nk, RW A_nn) */
void orb_cl_syrk(void *descr[], void *arg)
{
	(void)arg;
	const int n = MNY(1), k = MNY(0), lda = MLD(0), ldc = MLD(1);
	p_dsyrk("L", "N", &n, &k, &mone, MPTR(0), &lda, &one, MPTR(1), &ldc, 1, 1);
}

/* A_mn -= A_mk A_nk^T (R A_mk, R A_nk, RW A_mn) */
void orb_cl_gemm(void *descr[], void *arg)
{
	(void)arg;
	const int m = MNX(2), n = MNY(2), k = MNY(0), lda = MLD(0), ldb = MLD(1), ldc = MLD(2);
	p_dgemm("N", "T", &m, &n, &k, &mone, MPTR(0), &lda, MPTR(1), &ldb, &one, MPTR(2), &ldc, 1, 1);
}

/* W_k <- L_kk^-T, Linv_k <- L_kk^-1 (R L_kk, W W_k, W Linv_k, scratch) */
void orb_cl_trtri(void *descr[], void *arg)
{
	(void)arg;
	const int n = MNY(0), ldl = MLD(0), ldw = MLD(1), ldi = MLD(2);
	double *W = MPTR(1), *Li = MPTR(2);
	for (int j = 0; j < n; j++)
		for (int i = 0; i < n; i++)
			W[i + (size_t)j * ldw] = i == j ? 1.0 : 0.0;
	p_dtrsm("R", "L", "T", "N", &n, &n, &one, MPTR(0), &ldl, W, &ldw, 1, 1, 1, 1);
	for (int j = 0; j < n; j++)
		for (int i = 0; i < n; i++)
			Li[i + (size_t)j * ldi] = W[j + (size_t)i * ldw];
}

/* Forward sweep of block row k (R L_kk, R W_k, RW b_k, then pairs R L_mk,
 * RW b_m, scratch; cl_arg -> number of pairs): b_k <- L_kk^-1 b_k, then
 * b_m -= L_mk b_k */
void orb_cl_solve_fwd(void *descr[], void *arg)
{
	const int np = *(const int *)arg;
	const int m = MNX(2), r = MNY(2), ldl = MLD(0), ldb = MLD(2);
	p_dtrsm("L", "L", "N", "N", &m, &r, &one, MPTR(0), &ldl, MPTR(2), &ldb, 1, 1, 1, 1);
	for (int q = 0; q < np; q++) {
		const int t = 3 + 2 * q;
		const int mm = MNX(t + 1), nn = MNY(t + 1), kk = MNY(t), lt = MLD(t), lb = MLD(t + 1);
		p_dgemm("N", "N", &mm, &nn, &kk, &mone, MPTR(t), &lt, MPTR(2), &ldb, &one, MPTR(t + 1), &lb, 1, 1);
	}
}

/* Backward sweep of block row k (R L_kk, R W_k, RW b_k, then pairs R L_km,
 * RW b_m, scratch): b_k <- L_kk^-T b_k, then b_m -= L_km^T b_k */
void orb_cl_solve_bwd(void *descr[], void *arg)
{
	const int np = *(const int *)arg;
	const int m = MNX(2), r = MNY(2), ldl = MLD(0), ldb = MLD(2);
	p_dtrsm("L", "L", "T", "N", &m, &r, &one, MPTR(0), &ldl, MPTR(2), &ldb, 1, 1, 1, 1);
	for (int q = 0; q < np; q++) {
		const int t = 3 + 2 * q;
		const int mm = MNX(t + 1), nn = MNY(t + 1), kk = MNX(t), lt = MLD(t), lb = MLD(t + 1);
		p_dgemm("T", "N", &mm, &nn, &kk, &mone, MPTR(t), &lt, MPTR(2), &ldb, &one, MPTR(t + 1), &lb, 1, 1);
	}
}

/* dgemm through the loaded library, for the tests (C <- beta C + alpha A B^T) */
int orb_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda, const double *B, int ldb,
                 double beta, double *C, int ldc)
{
	if (!p_dgemm)
		return -1;
	p_dgemm("N", "T", &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc, 1, 1);
	return 0;
}
