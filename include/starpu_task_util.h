/*
 * starpu_task_util.h -- variadic task construction (reference:
 * include/starpu_task_util.h:40-200 argument codes, starpu_task_insert
 * :520, starpu_task_build :477).  The argument list is terminated by 0.
 */
#ifndef STARPU_TASK_UTIL_H
#define STARPU_TASK_UTIL_H

#include <stdarg.h>

#include "starpu_task.h"

#ifdef __cplusplus
extern "C" {
#endif

#define STARPU_MODE_SHIFT 17
#define STARPU_VALUE (1 << STARPU_MODE_SHIFT)
#define STARPU_CALLBACK (2 << STARPU_MODE_SHIFT)
#define STARPU_CALLBACK_WITH_ARG (3 << STARPU_MODE_SHIFT)
#define STARPU_CALLBACK_ARG (4 << STARPU_MODE_SHIFT)
#define STARPU_PRIORITY (5 << STARPU_MODE_SHIFT)
#define STARPU_EXECUTE_ON_NODE (6 << STARPU_MODE_SHIFT)
#define STARPU_EXECUTE_ON_DATA (7 << STARPU_MODE_SHIFT)
#define STARPU_DATA_ARRAY (8 << STARPU_MODE_SHIFT)
#define STARPU_DATA_MODE_ARRAY (9 << STARPU_MODE_SHIFT)
#define STARPU_NODE_ARRAY (10 << STARPU_MODE_SHIFT)
#define STARPU_TAG (11 << STARPU_MODE_SHIFT)
#define STARPU_HYPERVISOR_TAG (12 << STARPU_MODE_SHIFT)
#define STARPU_FLOPS (13 << STARPU_MODE_SHIFT)
#define STARPU_SCHED_CTX (14 << STARPU_MODE_SHIFT)
#define STARPU_PROLOGUE_CALLBACK (15 << STARPU_MODE_SHIFT)
#define STARPU_PROLOGUE_CALLBACK_ARG (16 << STARPU_MODE_SHIFT)
#define STARPU_PROLOGUE_CALLBACK_POP (17 << STARPU_MODE_SHIFT)
#define STARPU_PROLOGUE_CALLBACK_POP_ARG (18 << STARPU_MODE_SHIFT)
#define STARPU_EXECUTE_ON_WORKER (19 << STARPU_MODE_SHIFT)
#define STARPU_EXECUTE_WHERE (20 << STARPU_MODE_SHIFT)
#define STARPU_TAG_ONLY (21 << STARPU_MODE_SHIFT)
#define STARPU_POSSIBLY_PARALLEL (22 << STARPU_MODE_SHIFT)
#define STARPU_WORKER_ORDER (23 << STARPU_MODE_SHIFT)
#define STARPU_NODE_SELECTION_POLICY (24 << STARPU_MODE_SHIFT)
#define STARPU_NAME (25 << STARPU_MODE_SHIFT)
#define STARPU_CL_ARGS (26 << STARPU_MODE_SHIFT)
#define STARPU_CL_ARGS_NFREE (27 << STARPU_MODE_SHIFT)
#define STARPU_TASK_DEPS_ARRAY (28 << STARPU_MODE_SHIFT)
#define STARPU_TASK_COLOR (29 << STARPU_MODE_SHIFT)
#define STARPU_HANDLES_SEQUENTIAL_CONSISTENCY (30 << STARPU_MODE_SHIFT)
#define STARPU_TASK_SYNCHRONOUS (31 << STARPU_MODE_SHIFT)
#define STARPU_TASK_END_DEPS_ARRAY (32 << STARPU_MODE_SHIFT)
#define STARPU_TASK_END_DEP (33 << STARPU_MODE_SHIFT)
#define STARPU_TASK_WORKERIDS (34 << STARPU_MODE_SHIFT)
#define STARPU_SEQUENTIAL_CONSISTENCY (35 << STARPU_MODE_SHIFT)
#define STARPU_TASK_PROFILING_INFO (36 << STARPU_MODE_SHIFT)
#define STARPU_TASK_NO_SUBMITORDER (37 << STARPU_MODE_SHIFT)
#define STARPU_CALLBACK_ARG_NFREE (38 << STARPU_MODE_SHIFT)
#define STARPU_CALLBACK_WITH_ARG_NFREE (39 << STARPU_MODE_SHIFT)
#define STARPU_PROLOGUE_CALLBACK_ARG_NFREE (40 << STARPU_MODE_SHIFT)
#define STARPU_PROLOGUE_CALLBACK_POP_ARG_NFREE (41 << STARPU_MODE_SHIFT)
#define STARPU_TASK_SCHED_DATA (42 << STARPU_MODE_SHIFT)
#define STARPU_TRANSACTION (43 << STARPU_MODE_SHIFT)
#define STARPU_SHIFTED_MODE_MAX (44 << STARPU_MODE_SHIFT)

int starpu_task_insert(struct starpu_codelet *cl, ...);
int starpu_insert_task(struct starpu_codelet *cl, ...);
struct starpu_task *starpu_task_build(struct starpu_codelet *cl, ...);
int starpu_task_set(struct starpu_task *task, struct starpu_codelet *cl, ...);
/* Packs STARPU_VALUE arguments the way starpu_codelet_unpack_args reads
 * them: an int count, then for each argument its size_t size and bytes. */
void starpu_codelet_unpack_args(void *cl_arg, ...);
void starpu_codelet_pack_args(void **arg_buffer, size_t *arg_buffer_size, ...);

#ifdef __cplusplus
}
#endif
#endif
