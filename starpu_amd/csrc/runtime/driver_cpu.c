/*
 * CPU workers: one thread per core, each popping its own queue (or the
 * central queue), fetching the task's buffers to main memory and running
 * the codelet's cpu_funcs[impl].
 *
 * Reference: src/drivers/cpu/driver_cpu.c:374-470 (execute_job_on_cpu),
 * :772 (worker loop); feedback src/drivers/driver_common/driver_common.c
 * :234-330.
 */
#define _GNU_SOURCE
#include <sched.h>
#include <stdlib.h>
#include <string.h>

#include "sp_internal.h"

/* lock held on entry and on return */
void sp_execute_on_cpu(struct sp_worker *w, struct sp_job *j)
{
	struct starpu_task *t = j->task;
	j->running = 1;
	t->status = STARPU_TASK_RUNNING;
	for (int i = 0; i < j->nbuffers; i++) {
		int e = sp_fetch_buffer(j, i, w, NULL);
		STARPU_ASSERT_MSG(e == 0, "cannot fetch buffer %d of task %s on %s (%d)", i,
		                  t->name ? t->name : "?", w->name, e);
	}
	sp_sched_job_started(w, j);
	SP_UNLOCK();
	j->t_start = sp_now();
	sp_current_job = j;
	if (t->cl) {
		starpu_cpu_func_t f = t->cl->cpu_funcs[j->nimpl];
		if (f)
			f(j->ifaces, t->cl_arg);
	}
	sp_current_job = NULL;
	j->t_end = sp_now();
	const double dt = j->t_end - j->t_start;
	SP_LOCK();
	if (!t->failed)
		sp_perfmodel_feedback(j, w, dt);
	w->ntasks++;
	w->busy_us += dt;
	j->kernel_us = 0.0;
	if (sp.profiling) {
		if (!t->profiling_info)
			t->profiling_info = calloc(1, sizeof(*t->profiling_info));
		if (t->profiling_info) {
			t->profiling_info->submit_time = j->t_submit;
			t->profiling_info->push_start_time = j->t_push;
			t->profiling_info->start_time = j->t_start;
			t->profiling_info->end_time = j->t_end;
			t->profiling_info->workerid = w->id;
			t->profiling_info->bytes_transferred = j->bytes_in;
		}
		sp_trace_add(j, w);
	}
	sp_sched_job_done(w, j);
	SP_UNLOCK();
	sp_job_terminate(j);
	SP_LOCK();
}

static void *cpu_worker_main(void *arg)
{
	struct sp_worker *w = arg;
	sp_current_worker = w;
	SP_LOCK();
	for (;;) {
		struct sp_job *j = sp_sched_pop(w);
		if (j) {
			sp_execute_on_cpu(w, j);
			continue;
		}
		if (sp.shutting_down)
			break;
		pthread_cond_wait(&w->cond, &sp.lock);
	}
	SP_UNLOCK();
	return NULL;
}

int sp_cpu_workers_start(void)
{
	for (int i = 0; i < sp.nworkers; i++) {
		struct sp_worker *w = &sp.workers[i];
		if (w->arch != STARPU_CPU_WORKER)
			continue;
		pthread_attr_t attr;
		pthread_attr_init(&attr);
		pthread_attr_setstacksize(&attr, 8u << 20);
		int r = pthread_create(&w->thread, &attr, cpu_worker_main, w);
		pthread_attr_destroy(&attr);
		STARPU_ASSERT_MSG(r == 0, "cannot start CPU worker %d", i);
	}
	return 0;
}
