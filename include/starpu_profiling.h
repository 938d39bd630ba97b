/*
 * starpu_profiling.h -- per-task timing records and trace export.
 * The reference dispatches FxT events and per-task profiling info
 * (src/profiling/profiling.c:293, src/profiling/starpu_tracing.c:97-700);
 * here each finished task leaves one record {codelet, worker, submit,
 * start, end, device-kernel time, bytes moved} that can be exported as a
 * Chrome trace (chrome://tracing, Perfetto).
 */
#ifndef STARPU_PROFILING_H
#define STARPU_PROFILING_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STARPU_PROFILING_DISABLE 0
#define STARPU_PROFILING_ENABLE 1

int starpu_profiling_status_set(int status);
int starpu_profiling_status_get(void);

struct starpu_trace_record {
	const char *name;       /* task name, or codelet name */
	int workerid;
	double submit_us, start_us, end_us;
	double kernel_us;       /* device time from HIP events (0 on CPU) */
	size_t bytes_in;
	int priority;
	double push_us;         /* when the scheduler placed the task (it was ready) */
};

/* Number of records collected since the last reset, and access to them. */
size_t starpu_trace_count(void);
int starpu_trace_get(size_t i, struct starpu_trace_record *rec);
void starpu_trace_reset(void);
int starpu_trace_dump_chrome(const char *path);
/* Per-worker totals since init: executed tasks, busy microseconds. */
int starpu_worker_stats(unsigned workerid, unsigned long *ntasks, double *busy_us, double *kernel_us);
void starpu_profiling_worker_helper_display_summary(void);

#ifdef __cplusplus
}
#endif
#endif
