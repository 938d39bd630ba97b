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

/* With STARPU_HIP_RESERVE_CUS=c at init (and three or more streams per
 * GPU), the first HIP worker of each GPU runs on c CUs of every XCD and the
 * device's other workers on the remaining CUs; with c = -1 it keeps a
 * stream on all CUs, like the others.  Either way the reserved worker runs
 * only tasks placed on it with STARPU_EXECUTE_ON_WORKER (a DAG's critical
 * chain, which then never queues in a stream behind bulk updates).  Returns
 * that worker's id on HIP device devid, or -1.  No reference counterpart. */
int starpu_hip_reserved_worker(int devid);

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
