/*
 * Scheduling policies.
 *
 * dmda (default; reference src/sched_policies/deque_modeling_policy_data_aware.c):
 * for every worker able to run the task, predict its completion
 *     exp_end(w) = max(exp_start(w), now) + exp_len(w) + length(task, w)
 * from the history performance model and the data that would have to move
 * to w's memory node (bus model), then pick the minimum of
 *     alpha * (exp_end(w) - best_exp_end) + beta * transfer(w)
 * (:680-702).  While a model is not calibrated on some architecture, the
 * task goes to an uncalibrated worker with the fewest queued tasks
 * (:520-563).  Each worker owns a queue sorted by priority, FIFO among
 * equal priorities (:377 push_sorted_task).
 *
 * eager / prio: one central queue (FIFO / priority order) that every
 * worker pops the first task it can run from.  random: uniform choice.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sp_internal.h"

static struct sp_job *central_head;
static unsigned rnd_state = 12345u;

int sp_worker_can_execute(struct sp_worker *w, struct starpu_task *t, int *nimpl)
{
	struct starpu_codelet *cl = t->cl;
	if (t->execute_on_a_specific_worker && (int)t->workerid != w->id)
		return 0;
	if (w->reserved && !t->execute_on_a_specific_worker)
		return 0;
	if (!t->execute_on_a_specific_worker && t->workerids_len &&
	    ((unsigned)w->id / 32 >= t->workerids_len || !(t->workerids[w->id / 32] & (1u << (w->id % 32)))))
		return 0;
	if (!cl) {
		*nimpl = 0;
		return 1;
	}
	uint64_t where = t->where ? (uint64_t)(uint32_t)t->where : cl->where;
	for (int i = 0; i < STARPU_MAXIMPLEMENTATIONS; i++) {
		int ok = 0;
		if (w->arch == STARPU_CPU_WORKER)
			ok = cl->cpu_funcs[i] != NULL && (!where || (where & STARPU_CPU));
		else if (w->arch == STARPU_HIP_WORKER)
			ok = cl->hip_funcs[i] != NULL && (!where || (where & STARPU_HIP));
		if (!ok)
			continue;
		if (cl->can_execute && !cl->can_execute((unsigned)w->id, t, (unsigned)i))
			continue;
		*nimpl = i;
		return 1;
	}
	return 0;
}

static void queue_insert(struct sp_job **head, struct sp_job *j, int sorted)
{
	int prio = j->task->priority;
	struct sp_job **pp = head;
	if (sorted) {
		while (*pp && (*pp)->task->priority >= prio)
			pp = &(*pp)->qnext;
	} else {
		while (*pp)
			pp = &(*pp)->qnext;
	}
	j->qnext = *pp;
	*pp = j;
}

double sp_predict_length(struct starpu_task *t, struct sp_worker *w, int nimpl, struct sp_job *j)
{
	if (!t->cl || !t->cl->model)
		return NAN;
	return sp_perfmodel_predict(t->cl->model, t, j, &w->perf_arch, nimpl);
}

static double transfer_time(struct sp_job *j, struct sp_worker *w)
{
	double tt = 0.0;
	for (int i = 0; i < j->nbuffers; i++) {
		struct sp_buffer *b = &j->buf[i];
		if ((b->mode & STARPU_SCRATCH) || !(b->mode & STARPU_R))
			continue;
		starpu_data_handle_t h = b->handle;
		if (h->rep[w->node].state != SP_INVALID)
			continue;
		int src = sp_handle_valid_node(h);
		if (src < 0)
			continue;
		tt += sp.lat[src][w->node] + (double)h->size / sp.bw[src][w->node];
	}
	return tt;
}

static void wake_worker(struct sp_worker *w)
{
	if (w->arch == STARPU_CPU_WORKER)
		pthread_cond_signal(&w->cond);
	else {
		struct sp_hipthr *th = sp_hip_worker_thread(w);
		if (th)
			pthread_cond_signal(&th->cond);
	}
}

static void push_to_worker(struct sp_job *j, struct sp_worker *w, int nimpl, double len, double xfer)
{
	const double now = sp_now();
	j->workerid = w->id;
	j->nimpl = nimpl;
	j->predicted = isnan(len) ? 0.0 : len;
	j->predicted_transfer = xfer;
	j->task->predicted = len;
	j->task->predicted_transfer = xfer;
	j->t_push = now;
	if (w->exp_start < now)
		w->exp_start = now;
	w->exp_len += j->predicted;
	w->exp_end = w->exp_start + w->exp_len;
	queue_insert(&w->qhead, j, 1);
	w->qlen++;
	wake_worker(w);
}

/* Workers whose perfmodel arch descriptions are equal (the streams of one
 * GPU, say) get the same history-based prediction. */
static int same_perf_arch(const struct sp_worker *a, const struct sp_worker *b)
{
	if (a->perf_arch.ndevices != b->perf_arch.ndevices)
		return 0;
	for (int d = 0; d < a->perf_arch.ndevices; d++) {
		const struct starpu_perfmodel_device *x = &a->perf_arch.devices[d], *y = &b->perf_arch.devices[d];
		if (x->type != y->type || x->devid != y->devid || x->ncores != y->ncores)
			return 0;
	}
	return 1;
}

static void dmda_push(struct sp_job *j)
{
	struct starpu_task *t = j->task;
	const double now = sp_now();
	double len[STARPU_NMAXWORKERS], xfer[STARPU_NMAXWORKERS], end[STARPU_NMAXWORKERS];
	int impl[STARPU_NMAXWORKERS], ok[STARPU_NMAXWORKERS];
	int unknown = 0, any = 0;
	double best_end = INFINITY;
	/* one prediction per run of workers with the same arch (the lookup --
	 * footprint, arch and history entry -- runs under the runtime lock for
	 * every task), except for per-worker cost functions */
	const int per_worker = t->cl && t->cl->model && t->cl->model->type == STARPU_PER_WORKER;
	int prev = -1;
	for (int i = 0; i < sp.nworkers; i++) {
		struct sp_worker *w = &sp.workers[i];
		ok[i] = sp_worker_can_execute(w, t, &impl[i]);
		if (!ok[i])
			continue;
		any = 1;
		const struct sp_worker *pw = prev >= 0 ? &sp.workers[prev] : NULL;
		if (pw && !per_worker && impl[prev] == impl[i] && same_perf_arch(pw, w))
			len[i] = len[prev];
		else
			len[i] = sp_predict_length(t, w, impl[i], j);
		xfer[i] = (pw && pw->node == w->node) ? xfer[prev] : transfer_time(j, w);
		prev = i;
		if (isnan(len[i])) {
			unknown = 1;
			continue;
		}
		double start = w->exp_start > now ? w->exp_start : now;
		end[i] = start + w->exp_len + len[i];
		if (end[i] < best_end)
			best_end = end[i];
	}
	STARPU_ASSERT_MSG(any, "no worker can execute task %s", t->name ? t->name : "?");
	int best = -1;
	if (unknown) {
		/* calibration: an uncalibrated worker with the shortest queue */
		int bq = 0;
		for (int i = 0; i < sp.nworkers; i++) {
			if (!ok[i] || !isnan(len[i]))
				continue;
			int q = sp.workers[i].qlen;
			if (best < 0 || q < bq) {
				best = i;
				bq = q;
			}
		}
	} else {
		double bf = INFINITY;
		for (int i = 0; i < sp.nworkers; i++) {
			if (!ok[i])
				continue;
			double f = sp.alpha * (end[i] - best_end) + sp.beta * xfer[i];
			if (f < bf) {
				bf = f;
				best = i;
			}
		}
	}
	push_to_worker(j, &sp.workers[best], impl[best], len[best], xfer[best]);
	sp_prefetch_job(j, &sp.workers[best]);
}

static void random_push(struct sp_job *j)
{
	int cand[STARPU_NMAXWORKERS], impl[STARPU_NMAXWORKERS], n = 0;
	for (int i = 0; i < sp.nworkers; i++) {
		int im;
		if (sp_worker_can_execute(&sp.workers[i], j->task, &im)) {
			cand[n] = i;
			impl[n] = im;
			n++;
		}
	}
	STARPU_ASSERT(n > 0);
	rnd_state = rnd_state * 1103515245u + 12345u;
	int k = (int)((rnd_state >> 8) % (unsigned)n);
	struct sp_worker *w = &sp.workers[cand[k]];
	push_to_worker(j, w, impl[k], sp_predict_length(j->task, w, impl[k], j), 0.0);
}

void sp_sched_push(struct sp_job *j)
{
	struct starpu_task *t = j->task;
	if (t->prologue_callback_func) {
		/* reference src/core/sched_policy.c:583-588 calls the prologue at
		 * push time; it runs with the runtime lock dropped */
		void (*f)(void *) = t->prologue_callback_func;
		void *a = t->prologue_callback_arg;
		SP_UNLOCK();
		f(a);
		SP_LOCK();
	}
	if (!strcmp(sp.policy, "dmda")) {
		dmda_push(j);
	} else if (!strcmp(sp.policy, "random")) {
		random_push(j);
	} else {
		j->t_push = sp_now();
		queue_insert(&central_head, j, !strcmp(sp.policy, "prio"));
		sp_wake_all_workers();
	}
}

/* Hint for the HIP driver threads, read without the runtime lock: may a pop
 * for worker w find a task?  A stale answer only moves the pop to the next
 * poll, and a driver thread whose streams are all busy polls their events
 * without taking the lock at all. */
int sp_sched_may_pop(const struct sp_worker *w)
{
	return __atomic_load_n(&w->qhead, __ATOMIC_RELAXED) != NULL ||
	       __atomic_load_n(&central_head, __ATOMIC_RELAXED) != NULL;
}

struct sp_job *sp_sched_pop(struct sp_worker *w)
{
	if (sp.paused)
		return NULL;
	if (w->qhead) {
		struct sp_job *j = w->qhead;
		w->qhead = j->qnext;
		j->qnext = NULL;
		w->qlen--;
		sp.nready--;
		return j;
	}
	/* central queue (eager / prio) */
	for (struct sp_job **pp = &central_head; *pp; pp = &(*pp)->qnext) {
		int nimpl;
		if (sp_worker_can_execute(w, (*pp)->task, &nimpl)) {
			struct sp_job *j = *pp;
			*pp = j->qnext;
			j->qnext = NULL;
			j->workerid = w->id;
			j->nimpl = nimpl;
			sp.nready--;
			return j;
		}
	}
	return NULL;
}

void sp_sched_job_started(struct sp_worker *w, struct sp_job *j)
{
	(void)j;
	const double now = sp_now();
	if (w->exp_start < now)
		w->exp_start = now;
}

void sp_sched_job_done(struct sp_worker *w, struct sp_job *j)
{
	const double now = sp_now();
	w->exp_len -= j->predicted;
	if (w->exp_len < 0)
		w->exp_len = 0;
	w->exp_start = now;
	w->exp_end = w->exp_start + w->exp_len;
}

void sp_sched_reset(void)
{
	const double now = sp_now();
	for (int i = 0; i < sp.nworkers; i++) {
		struct sp_worker *w = &sp.workers[i];
		w->exp_start = now;
		if (!w->qhead && w->ninflight == 0)
			w->exp_len = 0;
		w->exp_end = w->exp_start + w->exp_len;
	}
}

const char *starpu_sched_get_policy_name(void) { return sp.policy; }

double starpu_sched_worker_expected_end(unsigned workerid)
{
	if ((int)workerid >= sp.nworkers)
		return NAN;
	return sp.workers[workerid].exp_end;
}

double starpu_task_expected_length(struct starpu_task *t, struct starpu_perfmodel_arch *arch, unsigned nimpl)
{
	if (!t->cl || !t->cl->model)
		return NAN;
	SP_LOCK();
	double r = sp_perfmodel_predict(t->cl->model, t, t->starpu_private, arch, (int)nimpl);
	SP_UNLOCK();
	return r;
}

int starpu_worker_can_execute_task(unsigned workerid, struct starpu_task *t, unsigned nimpl)
{
	if ((int)workerid >= sp.nworkers || !t)
		return 0;
	int impl = -1;
	return sp_worker_can_execute(&sp.workers[workerid], t, &impl) && impl == (int)nimpl;
}

double starpu_task_worker_expected_length(struct starpu_task *t, unsigned workerid, unsigned ctx, unsigned nimpl)
{
	(void)ctx;
	if ((int)workerid >= sp.nworkers)
		return NAN;
	return starpu_task_expected_length(t, &sp.workers[workerid].perf_arch, nimpl);
}

double starpu_transfer_bandwidth(unsigned src, unsigned dst)
{
	return sp.bw[src][dst];   /* bytes per microsecond = MB/s */
}

double starpu_transfer_latency(unsigned src, unsigned dst) { return sp.lat[src][dst]; }

double starpu_transfer_predict(unsigned src, unsigned dst, size_t size)
{
	if (src == dst)
		return 0.0;
	return sp.lat[src][dst] + (double)size / sp.bw[src][dst];
}

double starpu_data_expected_transfer_time(starpu_data_handle_t h, unsigned node, enum starpu_data_access_mode mode)
{
	if (!(mode & STARPU_R))
		return 0.0;
	SP_LOCK();
	double r = 0.0;
	if (h->rep[node].state == SP_INVALID) {
		int src = sp_handle_valid_node(h);
		if (src >= 0)
			r = starpu_transfer_predict((unsigned)src, node, h->size);
	}
	SP_UNLOCK();
	return r;
}

double starpu_task_expected_data_transfer_time_for(struct starpu_task *t, unsigned workerid)
{
	double r = 0.0;
	int n = t->cl ? (t->cl->nbuffers == STARPU_VARIABLE_NBUFFERS ? t->nbuffers : t->cl->nbuffers) : 0;
	unsigned node = starpu_worker_get_memory_node(workerid);
	for (int i = 0; i < n; i++)
		r += starpu_data_expected_transfer_time(STARPU_TASK_GET_HANDLE(t, i), node, STARPU_TASK_GET_MODE(t, i));
	return r;
}
