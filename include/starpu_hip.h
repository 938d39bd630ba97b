/*
 * starpu_hip.h -- HIP worker helpers (reference: include/starpu_hip.h:66
 * starpu_hip_get_local_stream, :48-53 STARPU_HIP_REPORT_ERROR).
 * Declares hipStream_t without the HIP headers so plain C codelets can use
 * it; it is the same type as HIP's.
 */
#ifndef STARPU_HIP_H
#define STARPU_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef STARPU_AMD_HIP_STREAM_T
#define STARPU_AMD_HIP_STREAM_T
typedef struct ihipStream_t *hipStream_t;
#endif

/* Stream of the HIP worker running the current codelet; codelets flagged
 * STARPU_HIP_ASYNC enqueue their kernels on it and return. */
hipStream_t starpu_hip_get_local_stream(void);
int starpu_hip_get_device(void);
void starpu_hip_report_error(const char *func, const char *file, int line, int status);
#define STARPU_HIP_REPORT_ERROR(status) starpu_hip_report_error(__func__, __FILE__, __LINE__, (int)(status))

/* The reference binds hipBLAS / hipSOLVER handles to worker streams
 * (src/drivers/hip/starpu_hipblas.c:56-150).  The tile kernels here are
 * hand-written, so these are no-ops kept for source compatibility. */
void starpu_hipblas_init(void);
void starpu_hipblas_shutdown(void);
void starpu_hipsolver_init(void);
void starpu_hipsolver_shutdown(void);
void starpu_cublas_init(void);
void starpu_cublas_shutdown(void);
void starpu_cusolver_init(void);
void starpu_cusolver_shutdown(void);

#ifdef __cplusplus
}
#endif
#endif
