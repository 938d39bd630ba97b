/*
 * CPU oracle for the tiled Cholesky factor-and-solve path.
 * TEST INFRASTRUCTURE ONLY -- see chol_oracle.h.  Plain FP64 C; every
 * function cites the reference lines whose arithmetic it restates.
 */
#include "chol_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* examples/cholesky/cholesky_kernels.c:75-76: SGEMM("N","T", dy, dx, dz,
 * -1, left, ld21, right, ld12, 1, center, ld22).  Column-oriented loop
 * (j, p, i) so the inner axpy runs down contiguous columns. */
void or_dgemm_nt(int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc)
{
	for (int j = 0; j < n; j++) {
		double *c = C + IDX(0, j, ldc);
		if (beta == 0.0)
			for (int i = 0; i < m; i++) c[i] = 0.0;
		else if (beta != 1.0)
			for (int i = 0; i < m; i++) c[i] *= beta;
		for (int p = 0; p < k; p++) {
			const double s = alpha * B[IDX(j, p, ldb)];
			const double *a = A + IDX(0, p, lda);
			for (int i = 0; i < m; i++) c[i] += s * a[i];
		}
	}
}

/* cholesky_kernels.c:169-170: SSYRK("L","N", dx, dz, -1, left, ld21, 1,
 * center, ld22).  Only i >= j of C is read or written. */
void or_dsyrk_ln(int n, int k, double alpha, const double *A, int lda,
                 double beta, double *C, int ldc)
{
	for (int j = 0; j < n; j++) {
		double *c = C + IDX(0, j, ldc);
		if (beta == 0.0)
			for (int i = j; i < n; i++) c[i] = 0.0;
		else if (beta != 1.0)
			for (int i = j; i < n; i++) c[i] *= beta;
		for (int p = 0; p < k; p++) {
			const double s = alpha * A[IDX(j, p, lda)];
			const double *a = A + IDX(0, p, lda);
			for (int i = j; i < n; i++) c[i] += s * a[i];
		}
	}
}

/* cholesky_kernels.c:254: STRSM("R","L","T","N", m, n, 1, L, ld11, B, ld21):
 * solve X * L^T = alpha*B.  Column j of X depends on columns 0..j-1:
 *   X(:,j) = (alpha*B(:,j) - sum_{p<j} X(:,p) L(j,p)) / L(j,j). */
void or_dtrsm_rltn(int m, int n, double alpha, const double *L, int ldl,
                   double *B, int ldb)
{
	for (int j = 0; j < n; j++) {
		double *b = B + IDX(0, j, ldb);
		if (alpha != 1.0)
			for (int i = 0; i < m; i++) b[i] *= alpha;
		for (int p = 0; p < j; p++) {
			const double l = L[IDX(j, p, ldl)];
			if (l == 0.0) continue;
			const double *x = B + IDX(0, p, ldb);
			for (int i = 0; i < m; i++) b[i] -= l * x[i];
		}
		const double d = L[IDX(j, j, ldl)];
		for (int i = 0; i < m; i++) b[i] /= d;
	}
}

/* cholesky_kernels.c:329-342, the unblocked CPU POTRF:
 *   lambda11 = sqrt(alpha11); l21 = alpha21 / lambda11;
 *   A22 <- A22 - l21 l21^T (SSYR lower).
 * The reference asserts lambda11 != 0 (:335); here a non-positive or NaN
 * pivot returns its 1-based column instead, as LAPACK dpotrf does. */
int or_dpotrf_l(int n, double *A, int lda)
{
	for (int z = 0; z < n; z++) {
		const double a = A[IDX(z, z, lda)];
		if (!(a > 0.0)) return z + 1;
		const double lambda = sqrt(a);
		A[IDX(z, z, lda)] = lambda;
		const double inv = 1.0 / lambda;
		for (int i = z + 1; i < n; i++) A[IDX(i, z, lda)] *= inv;
		for (int j = z + 1; j < n; j++) {
			const double s = A[IDX(j, z, lda)];
			double *c = A + IDX(0, j, lda);
			const double *l = A + IDX(0, z, lda);
			for (int i = j; i < n; i++) c[i] -= s * l[i];
		}
	}
	return 0;
}

/* Chunk rule of starpu_filter_nparts_compute_chunk_size_and_offset
 * (src/datawizard/filters.c:1655-1679): the first n % nparts chunks get one
 * extra element. */
static void chunk(int n, int nparts, int id, int *size, int *off)
{
	int base = n / nparts, rem = n % nparts;
	*size = base + (id < rem ? 1 : 0);
	*off = id * base + (id < rem ? id : rem);
}

/* cholesky_implicit.c:244-318, executed sequentially in submission order:
 * for k: POTRF(kk); for m>k: TRSM(mk); for n>k: SYRK(nn); for m>n: GEMM(mn). */
int or_tiled_cholesky(int N, int nblocks, double *A, int lda)
{
	for (int k = 0; k < nblocks; k++) {
		int sk, ok;
		chunk(N, nblocks, k, &sk, &ok);
		double *Akk = A + IDX(ok, ok, lda);
		int info = or_dpotrf_l(sk, Akk, lda);
		if (info) return ok + info;
		for (int m = k + 1; m < nblocks; m++) {
			int sm, om;
			chunk(N, nblocks, m, &sm, &om);
			or_dtrsm_rltn(sm, sk, 1.0, Akk, lda, A + IDX(om, ok, lda), lda);
		}
		for (int n = k + 1; n < nblocks; n++) {
			int sn, on;
			chunk(N, nblocks, n, &sn, &on);
			const double *Ank = A + IDX(on, ok, lda);
			or_dsyrk_ln(sn, sk, -1.0, Ank, lda, 1.0, A + IDX(on, on, lda), lda);
			for (int m = n + 1; m < nblocks; m++) {
				int sm, om;
				chunk(N, nblocks, m, &sm, &om);
				or_dgemm_nt(sm, sn, sk, -1.0, A + IDX(om, ok, lda), lda,
				            Ank, lda, 1.0, A + IDX(om, on, lda), lda);
			}
		}
	}
	return 0;
}

/* LAPACK dpotrs('L'): L y = b (forward), then L^T x = y (backward).
 * Not in the reference (it never solves); pinned by scipy.linalg.cho_solve
 * fixtures. */
void or_dpotrs_l(int n, int nrhs, const double *L, int ldl, double *B, int ldb)
{
	for (int r = 0; r < nrhs; r++) {
		double *b = B + IDX(0, r, ldb);
		for (int j = 0; j < n; j++) {
			b[j] /= L[IDX(j, j, ldl)];
			const double bj = b[j];
			const double *l = L + IDX(0, j, ldl);
			for (int i = j + 1; i < n; i++) b[i] -= bj * l[i];
		}
		for (int j = n - 1; j >= 0; j--) {
			const double *l = L + IDX(0, j, ldl);
			double s = b[j];
			for (int i = j + 1; i < n; i++) s -= l[i] * b[i];
			b[j] = s / l[j];
		}
	}
}

void or_dtrsm_llnn(int m, int n, double alpha, const double *L, int ldl, double *B, int ldb)
{
	/* column by column of B: forward substitution, column-oriented */
	for (int r = 0; r < n; r++) {
		double *b = B + IDX(0, r, ldb);
		for (int i = 0; i < m; i++) b[i] *= alpha;
		for (int j = 0; j < m; j++) {
			b[j] /= L[IDX(j, j, ldl)];
			const double bj = b[j];
			const double *l = L + IDX(0, j, ldl);
			for (int i = j + 1; i < m; i++) b[i] -= bj * l[i];
		}
	}
}

void or_dtrsm_lltn(int m, int n, double alpha, const double *L, int ldl, double *B, int ldb)
{
	/* L^T x = alpha b: backward substitution with dot products down the
	 * columns of L (the rows of L^T) */
	for (int r = 0; r < n; r++) {
		double *b = B + IDX(0, r, ldb);
		for (int j = m - 1; j >= 0; j--) {
			const double *l = L + IDX(0, j, ldl);
			double s = alpha * b[j];
			for (int i = j + 1; i < m; i++) s -= l[i] * b[i];
			b[j] = s / l[j];
		}
	}
}

void or_dgemm_nn(int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc)
{
	for (int j = 0; j < n; j++) {
		double *c = C + IDX(0, j, ldc);
		for (int i = 0; i < m; i++) c[i] = (beta == 0.0) ? 0.0 : beta * c[i];
		for (int p = 0; p < k; p++) {
			const double t = alpha * B[IDX(p, j, ldb)];
			const double *a = A + IDX(0, p, lda);
			for (int i = 0; i < m; i++) c[i] += t * a[i];
		}
	}
}

void or_dgemm_tn(int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc)
{
	for (int j = 0; j < n; j++)
		for (int i = 0; i < m; i++) {
			const double *a = A + IDX(0, i, lda);
			const double *b = B + IDX(0, j, ldb);
			double s = 0.0;
			for (int p = 0; p < k; p++) s += a[p] * b[p];
			double *c = C + IDX(i, j, ldc);
			*c = ((beta == 0.0) ? 0.0 : beta * *c) + alpha * s;
		}
}

void or_tiled_potrs(int N, int nblocks, int nrhs, const double *L, int ldl, double *B, int ldb)
{
	int off[nblocks + 1];
	for (int b = 0; b < nblocks; b++) {
		int sz, o;
		chunk(N, nblocks, b, &sz, &o);
		off[b] = o;
	}
	off[nblocks] = N;
#define TSZ(b) ((int)(off[(b) + 1] - off[b]))
#define LT(i, j) (L + IDX(off[i], off[j], ldl))
#define BT(i) (B + off[i])
	for (int k = 0; k < nblocks; k++) {
		or_dtrsm_llnn(TSZ(k), nrhs, 1.0, LT(k, k), ldl, BT(k), ldb);
		for (int m = k + 1; m < nblocks; m++)
			or_dgemm_nn(TSZ(m), nrhs, TSZ(k), -1.0, LT(m, k), ldl, BT(k), ldb, 1.0, BT(m), ldb);
	}
	for (int k = nblocks - 1; k >= 0; k--) {
		or_dtrsm_lltn(TSZ(k), nrhs, 1.0, LT(k, k), ldl, BT(k), ldb);
		for (int m = 0; m < k; m++)
			or_dgemm_tn(TSZ(m), nrhs, TSZ(k), -1.0, LT(k, m), ldl, BT(k), ldb, 1.0, BT(m), ldb);
	}
#undef TSZ
#undef LT
#undef BT
}

void or_residual_sym_l(int n, int nrhs, const double *A, int lda,
                       const double *X, int ldx, const double *B, int ldb,
                       double *R, int ldr)
{
	for (int r = 0; r < nrhs; r++) {
		double *res = R + IDX(0, r, ldr);
		const double *x = X + IDX(0, r, ldx);
		for (int i = 0; i < n; i++) res[i] = B[IDX(i, r, ldb)];
		for (int j = 0; j < n; j++) {
			/* column j of the symmetric A: A(i,j) = lower(max,min) */
			const double xj = x[j];
			for (int i = 0; i < j; i++) res[i] -= A[IDX(j, i, lda)] * xj;
			for (int i = j; i < n; i++) res[i] -= A[IDX(i, j, lda)] * xj;
		}
	}
}

/* FP32 restatement of or_dpotrf_l (cholesky_kernels.c:329-342 in the
 * reference's float arithmetic). */
static int spotrf_l(int n, float *A, int lda)
{
	for (int z = 0; z < n; z++) {
		const float a = A[IDX(z, z, lda)];
		if (!(a > 0.0f)) return z + 1;
		const float lambda = sqrtf(a);
		A[IDX(z, z, lda)] = lambda;
		const float inv = 1.0f / lambda;
		for (int i = z + 1; i < n; i++) A[IDX(i, z, lda)] *= inv;
		for (int j = z + 1; j < n; j++) {
			const float s = A[IDX(j, z, lda)];
			float *c = A + IDX(0, j, lda);
			const float *l = A + IDX(0, z, lda);
			for (int i = j; i < n; i++) c[i] -= s * l[i];
		}
	}
	return 0;
}

/* or_dpotrs_l for one FP32 column. */
static void spotrs1_l(int n, const float *L, int ldl, float *b)
{
	for (int j = 0; j < n; j++) {
		b[j] /= L[IDX(j, j, ldl)];
		const float bj = b[j];
		const float *l = L + IDX(0, j, ldl);
		for (int i = j + 1; i < n; i++) b[i] -= bj * l[i];
	}
	for (int j = n - 1; j >= 0; j--) {
		const float *l = L + IDX(0, j, ldl);
		float s = b[j];
		for (int i = j + 1; i < n; i++) s -= l[i] * b[i];
		b[j] = s / l[j];
	}
}

int or_sposv_ir(int n, int nrhs, const double *A, int lda, const double *B, int ldb,
                double *X, int ldx, int maxiter, double tol, int *iters, double *berr)
{
	float *Lf = malloc(sizeof(float) * (size_t)(n > 0 ? n : 1) * (size_t)(n > 0 ? n : 1));
	float *d = malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
	double *r = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
	double anorm2 = 0.0;
	for (int j = 0; j < n; j++)
		for (int i = j; i < n; i++) {
			const double a = A[IDX(i, j, lda)];
			Lf[IDX(i, j, n)] = (float)a;
			anorm2 += (i == j ? 1.0 : 2.0) * a * a;
		}
	const double anorm = sqrt(anorm2);
	const int info = spotrf_l(n, Lf, n);
	for (int c = 0; c < nrhs; c++) {
		const double *b = B + IDX(0, c, ldb);
		double *x = X + IDX(0, c, ldx);
		int it = 0;
		double be = HUGE_VAL;
		if (info == 0) {
			double bb = 0.0;
			for (int i = 0; i < n; i++) {
				x[i] = 0.0;
				r[i] = b[i];
				bb += b[i] * b[i];
			}
			const double bnorm = sqrt(bb);
			while (it < maxiter) {
				for (int i = 0; i < n; i++) d[i] = (float)r[i];
				spotrs1_l(n, Lf, n, d);
				double xx = 0.0, rr = 0.0;
				for (int i = 0; i < n; i++) {
					x[i] += (double)d[i];
					xx += x[i] * x[i];
				}
				or_residual_sym_l(n, 1, A, lda, x, n, b, n, r, n);
				for (int i = 0; i < n; i++) rr += r[i] * r[i];
				it++;
				be = rr == 0.0 ? 0.0 : sqrt(rr) / (anorm * sqrt(xx) + bnorm);
				if (be <= tol) break;
			}
		}
		iters[c] = it;
		berr[c] = be;
	}
	free(Lf);
	free(d);
	free(r);
	return info;
}

/* ||A - L L^T||_F / ||A||_F on the lower triangle.  Mirrors the -check of
 * cholesky_implicit.c:530-585 (which rebuilds L L^T with SSYRK and compares
 * elementwise) as a normwise quantity. */
double or_backward_error_llt(int n, const double *A, int lda,
                             const double *L, int ldl)
{
	/* column j of L L^T below the diagonal, c(i) = sum_{p <= j} L(i,p) L(j,p),
	 * accumulated over p so that the inner loop runs down a column of L */
	double num = 0.0, den = 0.0;
	double *c = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
	if (!c)
		return NAN;
	for (int j = 0; j < n; j++) {
		for (int i = j; i < n; i++) c[i] = 0.0;
		for (int p = 0; p <= j; p++) {
			const double ljp = L[IDX(j, p, ldl)];
			const double *l = L + IDX(0, p, ldl);
			for (int i = j; i < n; i++) c[i] += l[i] * ljp;
		}
		for (int i = j; i < n; i++) {
			const double a = A[IDX(i, j, lda)];
			num += (a - c[i]) * (a - c[i]);
			den += a * a;
		}
	}
	free(c);
	return sqrt(num) / sqrt(den);
}

double or_rel_diff_lower(int n, const double *L, int ldl,
                         const double *Lref, int ldr)
{
	double num = 0.0, den = 0.0;
	for (int j = 0; j < n; j++)
		for (int i = j; i < n; i++) {
			double d = L[IDX(i, j, ldl)] - Lref[IDX(i, j, ldr)];
			num += d * d;
			den += Lref[IDX(i, j, ldr)] * Lref[IDX(i, j, ldr)];
		}
	return sqrt(num) / sqrt(den);
}
