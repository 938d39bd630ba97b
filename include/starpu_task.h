/*
 * starpu_task.h -- codelets and tasks (reference: include/starpu_task.h;
 * codelet :370-727, task :752+, func types :214, STARPU_HIP_ASYNC :140,
 * submit :2012, wait_for_all :2073, iteration push/pop :2132-2140,
 * ft_failed :2277).
 */
#ifndef STARPU_TASK_H
#define STARPU_TASK_H

#include <stdint.h>

#include "starpu_config.h"
#include "starpu_data.h"
#include "starpu_perfmodel.h"

#ifdef __cplusplus
extern "C" {
#endif

#define STARPU_NOWHERE ((1ULL) << 0)
#define STARPU_CPU ((1ULL) << 1)
#define STARPU_CUDA ((1ULL) << 3)
#define STARPU_OPENCL ((1ULL) << 6)
#define STARPU_HIP ((1ULL) << 9)

#define STARPU_CUDA_ASYNC (1 << 0)
#define STARPU_HIP_ASYNC (1 << 0)
#define STARPU_OPENCL_ASYNC (1 << 0)

#define STARPU_MIN_PRIO (starpu_sched_get_min_priority())
#define STARPU_MAX_PRIO (starpu_sched_get_max_priority())
#define STARPU_DEFAULT_PRIO 0

#define STARPU_VARIABLE_NBUFFERS (-1)
#define STARPU_TASK_INVALID 0

typedef uint64_t starpu_tag_t;

typedef void (*starpu_cpu_func_t)(void **, void *);
typedef void (*starpu_cuda_func_t)(void **, void *);
typedef void (*starpu_opencl_func_t)(void **, void *);
typedef void (*starpu_hip_func_t)(void **, void *);

#define STARPU_MULTIPLE_CPU_IMPLEMENTATIONS ((starpu_cpu_func_t)-1)

enum starpu_codelet_type {
	STARPU_SEQ = 0,
	STARPU_SPMD,
	STARPU_FORKJOIN
};

enum starpu_task_status {
	STARPU_TASK_INIT,
	STARPU_TASK_BLOCKED,
	STARPU_TASK_READY,
	STARPU_TASK_RUNNING,
	STARPU_TASK_FINISHED,
	STARPU_TASK_BLOCKED_ON_TAG,
	STARPU_TASK_BLOCKED_ON_TASK,
	STARPU_TASK_BLOCKED_ON_DATA,
	STARPU_TASK_STOPPED
};

struct starpu_task;

struct starpu_codelet {
	uint32_t where;
	int (*can_execute)(unsigned workerid, struct starpu_task *task, unsigned nimpl);
	enum starpu_codelet_type type;
	int max_parallelism;
	starpu_cpu_func_t cpu_funcs[STARPU_MAXIMPLEMENTATIONS];
	/* accepted for source compatibility; no CUDA/OpenCL device exists here */
	starpu_cuda_func_t cuda_funcs[STARPU_MAXIMPLEMENTATIONS];
	char cuda_flags[STARPU_MAXIMPLEMENTATIONS];
	starpu_opencl_func_t opencl_funcs[STARPU_MAXIMPLEMENTATIONS];
	char opencl_flags[STARPU_MAXIMPLEMENTATIONS];
	starpu_hip_func_t hip_funcs[STARPU_MAXIMPLEMENTATIONS];
	char hip_flags[STARPU_MAXIMPLEMENTATIONS];
	const char *cpu_funcs_name[STARPU_MAXIMPLEMENTATIONS];
	int nbuffers;
	enum starpu_data_access_mode modes[STARPU_NMAXBUFS];
	enum starpu_data_access_mode *dyn_modes;
	unsigned specific_nodes;
	int nodes[STARPU_NMAXBUFS];
	int *dyn_nodes;
	struct starpu_perfmodel *model;
	struct starpu_perfmodel *energy_model;
	unsigned long per_worker_stats[STARPU_NMAXWORKERS];
	const char *name;
	unsigned color;
	void (*callback_func)(void *);
	int flags;
	int checked;
};

/* Timing of one execution, filled when profiling is on (STARPU_PROFILING=1
 * or starpu_profiling_status_set(1)); times in microseconds on the
 * starpu_timing_now clock. */
struct starpu_profiling_task_info {
	double submit_time;
	double push_start_time;
	double start_time;
	double end_time;
	int workerid;
	double kernel_us;        /* device execution time measured with events */
	size_t bytes_transferred;
};

struct starpu_task {
	const char *name;
	const char *file;
	int line;
	struct starpu_codelet *cl;
	int32_t where;
	int nbuffers;
	starpu_data_handle_t *dyn_handles;
	void **dyn_interfaces;
	enum starpu_data_access_mode *dyn_modes;
	starpu_data_handle_t handles[STARPU_NMAXBUFS];
	void *interfaces[STARPU_NMAXBUFS];
	enum starpu_data_access_mode modes[STARPU_NMAXBUFS];
	unsigned char *handles_sequential_consistency;
	void *cl_arg;
	size_t cl_arg_size;
	unsigned cl_arg_free;
	void (*callback_func)(void *);
	void *callback_arg;
	unsigned callback_arg_free;
	void (*prologue_callback_func)(void *);
	void *prologue_callback_arg;
	unsigned prologue_callback_arg_free;
	starpu_tag_t tag_id;
	unsigned use_tag;
	unsigned sequential_consistency;
	unsigned synchronous;
	unsigned execute_on_a_specific_worker;
	unsigned workerid;
	/* with workerids_len != 0: bit w % 32 of workerids[w / 32] allows
	 * worker w to run the task (workerid takes precedence);
	 * include/starpu_task.h:1410-1432 of the reference */
	uint32_t *workerids;
	unsigned workerids_len;
	int priority;
	enum starpu_task_status status;
	unsigned detach;
	unsigned destroy;
	unsigned regenerate;
	unsigned no_submitorder;
	unsigned failed;
	unsigned scheduled;
	unsigned prefetched;
	double flops;
	double predicted;
	double predicted_transfer;
	long iterations[2];
	struct starpu_profiling_task_info *profiling_info;
	unsigned sched_ctx;
	int magic;
	void *sched_data;
	/* runtime-private job */
	void *starpu_private;
};

void starpu_task_init(struct starpu_task *task);
void starpu_task_clean(struct starpu_task *task);
struct starpu_task *starpu_task_create(void);
void starpu_task_destroy(struct starpu_task *task);
int starpu_task_submit(struct starpu_task *task);
int starpu_task_wait(struct starpu_task *task);
int starpu_task_wait_array(struct starpu_task **tasks, unsigned nb_tasks);
int starpu_task_wait_for_all(void);
int starpu_task_wait_for_n_submitted(unsigned n);
int starpu_task_nsubmitted(void);
int starpu_task_nready(void);
struct starpu_task *starpu_task_get_current(void);

/* Handle and access mode of one buffer, the element type of
 * STARPU_DATA_MODE_ARRAY lists (reference include/starpu_task.h). */
struct starpu_data_descr {
	starpu_data_handle_t handle;
	enum starpu_data_access_mode mode;
};

/* Number of buffers of a task, variable-buffer codelets included. */
#define STARPU_TASK_GET_NBUFFERS(task) \
	((unsigned)((task)->cl->nbuffers == STARPU_VARIABLE_NBUFFERS ? (task)->nbuffers : (task)->cl->nbuffers))
const char *starpu_task_get_name(struct starpu_task *task);
const char *starpu_task_get_model_name(struct starpu_task *task);
void starpu_task_declare_deps(struct starpu_task *task, unsigned ndeps, ...);
void starpu_task_declare_deps_array(struct starpu_task *task, unsigned ndeps, struct starpu_task *task_array[]);
void starpu_iteration_push(unsigned long iteration);
void starpu_iteration_pop(void);
/* The running codelet reports that this execution failed: the task is not
 * considered complete and is scheduled again (reference src/core/task.c:2565). */
void starpu_task_ft_failed(struct starpu_task *task);
starpu_data_handle_t starpu_task_get_current_data_node_handle(unsigned i);

#define STARPU_TASK_GET_HANDLE(task, i) \
	(((task)->dyn_handles) ? (task)->dyn_handles[i] : (task)->handles[i])
#define STARPU_TASK_GET_MODE(task, i) \
	(((task)->cl->nbuffers == STARPU_VARIABLE_NBUFFERS || (task)->dyn_modes) ? \
	 (((task)->dyn_modes) ? (task)->dyn_modes[i] : (task)->modes[i]) : \
	 (((task)->cl->dyn_modes) ? (task)->cl->dyn_modes[i] : (task)->cl->modes[i]))
#define STARPU_CODELET_GET_MODE(codelet, i) \
	(((codelet)->dyn_modes) ? (codelet)->dyn_modes[i] : (codelet)->modes[i])

/* Scheduler priorities range of the current policy (INT_MIN..INT_MAX). */
int starpu_sched_get_min_priority(void);
int starpu_sched_get_max_priority(void);

#ifdef __cplusplus
}
#endif
#endif
