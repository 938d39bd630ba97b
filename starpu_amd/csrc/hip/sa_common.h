// Internal helpers shared by the gfx950 kernel translation units.
#ifndef SA_COMMON_H
#define SA_COMMON_H

#include <hip/hip_runtime.h>
#include <errno.h>
#include <stdint.h>

typedef double sa_d2 __attribute__((ext_vector_type(2)));
typedef double sa_d4 __attribute__((ext_vector_type(4)));
typedef float sa_f4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64: lane l supplies A(l%16, l/16) and B(l/16, l%16) of
// a 16x4 by 4x16 product and accumulates D(4q + l/16, l%16) in register q.
__device__ __forceinline__ sa_d4 sa_mfma(double a, double b, sa_d4 c)
{
	return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

static inline int sa_launch_status()
{
	hipError_t e = hipGetLastError();
	return e == hipSuccess ? 0 : (int)e;
}

static inline bool sa_aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

// C = beta*C + alpha*A*B^T (lower: only the lower triangle of square C).
int sa_gemm_launch(hipStream_t s, bool lower, int m, int n, int k, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta,
                   double *C, int ldc);
// X <- a*X on an m x n region (lower: i >= j only); a == 0 stores zeros.
int sa_scale_launch(hipStream_t s, int m, int n, double *X, int ld, double a, bool lower);
// W = L^{-T} for n <= 1024 in two launches (sa_chol.hip); work holds
// sa_dtrsm_rltn_worksize(n) bytes.
int sa_trtri_lt_rows16(hipStream_t s, int n, const double *L, int ldl, double *W, int ldw, void *work,
                       size_t work_bytes);

#endif
