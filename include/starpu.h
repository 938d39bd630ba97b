/*
 * starpu.h -- public API of the starpu_amd runtime: the StarPU programming
 * model (codelets, tasks with implicit data dependencies, data handles with
 * MSI coherency, the dmda scheduler with history performance models) on
 * host CPU cores and AMD Instinct GPUs through HIP.
 *
 * Reference: include/starpu.h (starpu_init :679, starpu_shutdown :706,
 * starpu_pause/resume :715, struct starpu_conf :60-600).
 */
#ifndef STARPU_H
#define STARPU_H

#include <stddef.h>
#include <stdint.h>

#include "starpu_config.h"
#include "starpu_util.h"
#include "starpu_data.h"
#include "starpu_data_interfaces.h"
#include "starpu_data_filters.h"
#include "starpu_worker.h"
#include "starpu_perfmodel.h"
#include "starpu_task.h"
#include "starpu_task_util.h"
#include "starpu_task_dep.h"
#include "starpu_scheduler.h"
#include "starpu_hip.h"
#include "starpu_hash.h"
#include "starpu_stdlib.h"
#include "starpu_profiling.h"

#ifdef __cplusplus
extern "C" {
#endif

struct starpu_conf {
	int magic;
	const char *sched_policy_name;   /* NULL: $STARPU_SCHED or "dmda" */
	int ncpus;                       /* -1: $STARPU_NCPU or all cores */
	int nhip;                        /* -1: $STARPU_NHIP or all devices */
	int ncuda;                       /* accepted, always 0 here */
	int nopencl;
	int nworker_per_hip;             /* -1: $STARPU_NWORKER_PER_HIP or 8 */
	int hip_pipeline;                /* -1: $STARPU_HIP_PIPELINE or 2 */
	int calibrate;                   /* -1: $STARPU_CALIBRATE or 0 */
	int bus_calibrate;
	int disable_asynchronous_copy;
	int precedence_over_environment_variables;
	long long limit_hip_mem_mb;      /* -1: $STARPU_LIMIT_HIP_MEM or none */
	int watchdog_timeout_us;         /* 0: $STARPU_WATCHDOG_TIMEOUT or off */
};

int starpu_conf_init(struct starpu_conf *conf);
/* Returns 0, or -ENODEV when no worker can be created. */
int starpu_init(struct starpu_conf *conf);
int starpu_initialize(struct starpu_conf *user_conf, int *argc, char ***argv);
int starpu_is_initialized(void);
void starpu_shutdown(void);
/* Stop/resume handing tasks to workers (submission continues). */
void starpu_pause(void);
void starpu_resume(void);
void starpu_reset_scheduler(void);
void starpu_topology_print(FILE *f);
int starpu_get_env_number(const char *str);
int starpu_get_env_number_default(const char *str, int defval);

/* StarPU names the first HIP device's memory node with this helper in
 * examples; it is the node of the first HIP worker, or -1. */
int starpu_hip_memory_node(int devid);

#ifdef __cplusplus
}
#endif
#endif
