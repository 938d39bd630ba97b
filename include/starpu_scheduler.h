/*
 * starpu_scheduler.h -- scheduler queries (reference:
 * include/starpu_scheduler.h:430 starpu_task_expected_length and
 * neighbours).  Policies are chosen with STARPU_SCHED ("dmda", default;
 * "eager"; "prio"; "random").
 */
#ifndef STARPU_SCHEDULER_H
#define STARPU_SCHEDULER_H

#include "starpu_task.h"

#ifdef __cplusplus
extern "C" {
#endif

double starpu_task_expected_length(struct starpu_task *task, struct starpu_perfmodel_arch *arch, unsigned nimpl);
double starpu_task_worker_expected_length(struct starpu_task *task, unsigned workerid, unsigned sched_ctx_id, unsigned nimpl);
double starpu_task_expected_data_transfer_time_for(struct starpu_task *task, unsigned workerid);
double starpu_data_expected_transfer_time(starpu_data_handle_t handle, unsigned memory_node, enum starpu_data_access_mode mode);
double starpu_transfer_bandwidth(unsigned src_node, unsigned dst_node);
double starpu_transfer_latency(unsigned src_node, unsigned dst_node);
double starpu_transfer_predict(unsigned src_node, unsigned dst_node, size_t size);
const char *starpu_sched_get_policy_name(void);
/* Nonzero if worker `workerid` can run implementation nimpl of the task
 * (reference include/starpu_scheduler.h:308). */
int starpu_worker_can_execute_task(unsigned workerid, struct starpu_task *task, unsigned nimpl);
/* Expected end (microseconds, starpu_timing_now clock) of the work queued
 * on a worker, as dmda tracks it. */
double starpu_sched_worker_expected_end(unsigned workerid);

#ifdef __cplusplus
}
#endif
#endif
