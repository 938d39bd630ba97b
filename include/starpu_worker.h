/*
 * starpu_worker.h -- worker enumeration (reference: include/starpu_worker.h
 * :186,:200 and neighbours).
 */
#ifndef STARPU_WORKER_H
#define STARPU_WORKER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum starpu_worker_archtype {
	STARPU_CPU_WORKER = 0,
	STARPU_CUDA_WORKER = 1,
	STARPU_OPENCL_WORKER = 2,
	STARPU_MAX_FPGA_WORKER = 4,
	STARPU_MPI_SC_WORKER = 5,
	STARPU_TCPIP_SC_WORKER = 6,
	STARPU_HIP_WORKER = 7,
	STARPU_SYCL_WORKER = 8,
	STARPU_NARCH = 9,
	STARPU_ANY_WORKER = 255
};

unsigned starpu_worker_get_count(void);
unsigned starpu_cpu_worker_get_count(void);
unsigned starpu_hip_worker_get_count(void);
unsigned starpu_cuda_worker_get_count(void);
unsigned starpu_opencl_worker_get_count(void);
unsigned starpu_memory_nodes_get_count(void);
int starpu_worker_get_id(void);
unsigned starpu_worker_get_id_check(void);
enum starpu_worker_archtype starpu_worker_get_type(int id);
int starpu_worker_get_devid(int id);
unsigned starpu_worker_get_memory_node(unsigned workerid);
int starpu_worker_get_by_type(enum starpu_worker_archtype type, int num);
unsigned starpu_worker_get_count_by_type(enum starpu_worker_archtype type);
void starpu_worker_get_name(int id, char *dst, size_t maxlen);
const char *starpu_worker_get_type_as_string(enum starpu_worker_archtype type);
int starpu_combined_worker_get_size(void);
int starpu_combined_worker_get_rank(void);

#ifdef __cplusplus
}
#endif
#endif
