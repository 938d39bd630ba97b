/*
 * Tasks and jobs: creation, submission with implicit and explicit
 * dependencies, termination, waits, application data acquisition and the
 * fault-tolerance retry path.
 *
 * Reference call stack (src/core/task.c:1106 _starpu_task_submit ->
 * implicit_data_deps.c:435 -> task.c:662 _starpu_submit_job ->
 * jobs.c:793 enforce deps -> sched_policy.c:577 push), termination
 * (jobs.c:363 _starpu_handle_job_termination: release data, callback,
 * notify successors), ft_failed (task.c:2565).
 */
#define _GNU_SOURCE
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "sp_internal.h"

static __thread long sp_iteration[2] = {-1, -1};

void starpu_task_init(struct starpu_task *t)
{
	memset(t, 0, sizeof(*t));
	t->sequential_consistency = 1;
	t->detach = 1;
	t->destroy = 0;
	t->priority = STARPU_DEFAULT_PRIO;
	t->status = STARPU_TASK_INIT;
	t->workerid = 0;
	t->predicted = NAN;
	t->predicted_transfer = NAN;
	t->iterations[0] = sp_iteration[0];
	t->iterations[1] = sp_iteration[1];
	t->magic = 42;
}

struct starpu_task *starpu_task_create(void)
{
	struct starpu_task *t = malloc(sizeof(*t));
	STARPU_ASSERT(t);
	starpu_task_init(t);
	t->destroy = 1;
	return t;
}

static void job_free(struct sp_job *j)
{
	if (!j)
		return;
	if (j->buf != j->buf_inline)
		free(j->buf);
	if (j->ifaces != j->ifaces_inline)
		free(j->ifaces);
	free(j->succ);
	pthread_cond_destroy(&j->cond);
	free(j);
}

void starpu_task_clean(struct starpu_task *t)
{
	struct sp_job *j = t->starpu_private;
	if (j) {
		job_free(j);
		t->starpu_private = NULL;
	}
	if (t->cl_arg_free && t->cl_arg) {
		free(t->cl_arg);
		t->cl_arg = NULL;
	}
	if (t->callback_arg_free && t->callback_arg) {
		free(t->callback_arg);
		t->callback_arg = NULL;
	}
	if (t->prologue_callback_arg_free && t->prologue_callback_arg) {
		free(t->prologue_callback_arg);
		t->prologue_callback_arg = NULL;
	}
	free(t->dyn_handles);
	t->dyn_handles = NULL;
	free(t->dyn_modes);
	t->dyn_modes = NULL;
	free(t->dyn_interfaces);
	t->dyn_interfaces = NULL;
	free(t->profiling_info);
	t->profiling_info = NULL;
}

void starpu_task_destroy(struct starpu_task *t)
{
	if (!t)
		return;
	starpu_task_clean(t);
	free(t);
}

static struct sp_job *job_new(struct starpu_task *t, int nbuffers)
{
	struct sp_job *j = calloc(1, sizeof(*j));
	STARPU_ASSERT(j);
	j->task = t;
	j->nbuffers = nbuffers;
	if (nbuffers > STARPU_NMAXBUFS) {
		j->buf = calloc((size_t)nbuffers, sizeof(*j->buf));
		j->ifaces = calloc((size_t)nbuffers, sizeof(*j->ifaces));
	} else {
		j->buf = j->buf_inline;
		j->ifaces = j->ifaces_inline;
	}
	j->workerid = -1;
	j->nimpl = 0;
	j->predicted = NAN;
	pthread_cond_init(&j->cond, NULL);
	return j;
}

struct sp_job *sp_job_of(struct starpu_task *t) { return t->starpu_private; }

void sp_job_add_dependency(struct sp_job *pred, struct sp_job *succ)
{
	if (!pred || pred == succ || pred->terminated)
		return;
	for (int i = pred->nsucc - 1; i >= 0 && i >= pred->nsucc - 4; i--)
		if (pred->succ[i] == succ)
			return;
	if (pred->nsucc == pred->succ_cap) {
		pred->succ_cap = pred->succ_cap ? 2 * pred->succ_cap : 4;
		pred->succ = realloc(pred->succ, (size_t)pred->succ_cap * sizeof(*pred->succ));
	}
	pred->succ[pred->nsucc++] = succ;
	succ->ndeps++;
}

void sp_job_ready(struct sp_job *j)
{
	if (j->is_acquire) {
		j->acquire_ready = 1;
		pthread_cond_broadcast(&j->cond);
		return;
	}
	j->task->status = STARPU_TASK_READY;
	sp.nready++;
	sp_sched_push(j);
}

static int task_nbuffers(struct starpu_task *t)
{
	if (!t->cl)
		return 0;
	if (t->cl->nbuffers == STARPU_VARIABLE_NBUFFERS)
		return t->nbuffers;
	return t->cl->nbuffers;
}

/* Successors of j get one dependency fewer; those left without any are
 * handed to the scheduler (lock held). */
static void notify_successors(struct sp_job *j)
{
	for (int i = 0; i < j->nsucc; i++) {
		struct sp_job *s = j->succ[i];
		STARPU_ASSERT(s->ndeps > 0);
		if (--s->ndeps == 0 && s->submitted)
			sp_job_ready(s);
	}
	j->nsucc = 0;
}

int starpu_task_submit(struct starpu_task *t)
{
	STARPU_ASSERT_MSG(t->magic == 42, "task not initialised with starpu_task_init/create");
	if (!sp.initialized)
		return -ENODEV;
	struct starpu_codelet *cl = t->cl;
	int nbuf = task_nbuffers(t);
	if (cl) {
		/* a codelet must be runnable by at least one worker */
		int any = 0;
		for (int w = 0; w < sp.nworkers && !any; w++) {
			int nimpl;
			if (sp_worker_can_execute(&sp.workers[w], t, &nimpl))
				any = 1;
		}
		if (!any && !(cl->where & STARPU_NOWHERE))
			return -ENODEV;
	}
	/* A synchronous task is waited for here, so no worker may free it at
	 * termination (src/core/task.c:979-985 clears detach the same way).
	 * Once the lock is dropped below, a detached task may already have run
	 * and been destroyed: nothing of t is read after SP_UNLOCK. */
	const int is_sync = t->synchronous;
	if (is_sync)
		t->detach = 0;
	SP_LOCK();
	struct sp_job *j = t->starpu_private;
	if (!j) {
		j = job_new(t, nbuf);
		t->starpu_private = j;
	} else {
		/* a job created earlier by starpu_task_declare_deps (keep its
		 * edges), or the resubmission of a finished task */
		STARPU_ASSERT_MSG(j->terminated || !j->submitted, "task submitted twice");
		if (j->nbuffers != nbuf) {
			SP_UNLOCK();
			return -EINVAL;
		}
		if (j->terminated) {
			j->terminated = 0;
			j->running = 0;
			j->nsucc = 0;
		}
		j->bytes_in = 0;
		j->footprint_valid = 0;
	}
	for (int i = 0; i < nbuf; i++) {
		j->buf[i].handle = STARPU_TASK_GET_HANDLE(t, i);
		j->buf[i].mode = STARPU_TASK_GET_MODE(t, i);
		j->buf[i].rep = NULL;
		STARPU_ASSERT_MSG(j->buf[i].handle, "buffer %d of task %s has no handle", i,
		                  t->name ? t->name : "?");
		STARPU_ASSERT_MSG(!j->buf[i].handle->partitioned,
		                  "task %s accesses a partitioned handle", t->name ? t->name : "?");
	}
	j->t_submit = sp_now();
	j->submitted = 1;
	t->status = STARPU_TASK_BLOCKED;
	sp.nsubmitted++;
	/* keep a count while deps are being attached so that a predecessor
	 * finishing meanwhile cannot release the job early */
	j->ndeps++;
	if (cl)
		sp_detect_implicit_deps(j);
	sp_tag_job_submitted(j);
	j->ndeps--;
	if (j->ndeps == 0)
		sp_job_ready(j);
	SP_UNLOCK();
	if (is_sync)
		return starpu_task_wait(t);
	return 0;
}

void starpu_task_declare_deps_array(struct starpu_task *t, unsigned ndeps, struct starpu_task *arr[])
{
	SP_LOCK();
	struct sp_job *j = t->starpu_private;
	if (!j) {
		j = job_new(t, task_nbuffers(t));
		t->starpu_private = j;
	}
	for (unsigned i = 0; i < ndeps; i++) {
		struct starpu_task *p = arr[i];
		struct sp_job *pj = p->starpu_private;
		if (!pj) {
			pj = job_new(p, task_nbuffers(p));
			p->starpu_private = pj;
		}
		if (!pj->terminated || !pj->submitted)
			sp_job_add_dependency(pj, j);
	}
	SP_UNLOCK();
}

void starpu_task_declare_deps(struct starpu_task *t, unsigned ndeps, ...)
{
	struct starpu_task *arr[64];
	STARPU_ASSERT(ndeps <= 64);
	va_list ap;
	va_start(ap, ndeps);
	for (unsigned i = 0; i < ndeps; i++)
		arr[i] = va_arg(ap, struct starpu_task *);
	va_end(ap);
	starpu_task_declare_deps_array(t, ndeps, arr);
}

/* Called by a driver, without the lock, once the codelet of j has run to
 * completion (and its device work is done). */
void sp_job_terminate(struct sp_job *j)
{
	struct starpu_task *t = j->task;
	SP_LOCK();
	if (!j->is_acquire && t->failed) {
		/* The codelet reported a failure (starpu_task_ft_failed): run it
		 * again, keeping its dependencies and data locked. */
		t->failed = 0;
		for (int i = 0; i < j->nbuffers; i++)
			sp_release_buffer(j, i);
		j->running = 0;
		j->workerid = -1;
		sp_job_ready(j);
		SP_UNLOCK();
		return;
	}
	for (int i = 0; i < j->nbuffers; i++)
		sp_release_buffer(j, i);
	sp_remove_implicit_deps(j);
	j->terminated = 1;
	j->running = 0;
	notify_successors(j);
	sp_tag_job_done(j);
	sp.last_completion = sp_now();
	void (*cb)(void *) = j->is_acquire ? NULL : t->callback_func;
	void *cb_arg = t ? t->callback_arg : NULL;
	SP_UNLOCK();
	if (cb) {
		struct starpu_task *prev = sp_current_job ? sp_current_job->task : NULL;
		(void)prev;
		cb(cb_arg);
	}
	if (j->is_acquire) {
		job_free(j);
		return;
	}
	SP_LOCK();
	t->status = STARPU_TASK_FINISHED;
	int destroy = t->destroy && t->detach && !t->regenerate;
	int regenerate = t->regenerate;
	sp.nsubmitted--;
	pthread_cond_broadcast(&j->cond);
	pthread_cond_broadcast(&sp.done_cond);
	SP_UNLOCK();
	if (regenerate) {
		j->submitted = 0;
		starpu_task_submit(t);
	} else if (destroy) {
		starpu_task_destroy(t);
	}
}

int starpu_task_wait(struct starpu_task *t)
{
	STARPU_ASSERT_MSG(!t->detach || t->synchronous, "starpu_task_wait on a detached task");
	SP_LOCK();
	struct sp_job *j = t->starpu_private;
	if (!j || !j->submitted) {
		SP_UNLOCK();
		return -EINVAL;
	}
	while (t->status != STARPU_TASK_FINISHED)
		pthread_cond_wait(&j->cond, &sp.lock);
	SP_UNLOCK();
	if (t->destroy && !t->detach)
		starpu_task_destroy(t);
	return 0;
}

int starpu_task_wait_array(struct starpu_task **tasks, unsigned n)
{
	for (unsigned i = 0; i < n; i++) {
		int r = starpu_task_wait(tasks[i]);
		if (r)
			return r;
	}
	return 0;
}

int starpu_task_wait_for_all(void)
{
	SP_LOCK();
	while (sp.nsubmitted > 0)
		pthread_cond_wait(&sp.done_cond, &sp.lock);
	SP_UNLOCK();
	return 0;
}

int starpu_task_wait_for_n_submitted(unsigned n)
{
	SP_LOCK();
	while (sp.nsubmitted > (long)n)
		pthread_cond_wait(&sp.done_cond, &sp.lock);
	SP_UNLOCK();
	return 0;
}

int starpu_task_nsubmitted(void) { return (int)sp.nsubmitted; }
int starpu_task_nready(void) { return (int)sp.nready; }

struct starpu_task *starpu_task_get_current(void) { return sp_current_job ? sp_current_job->task : NULL; }

const char *starpu_task_get_name(struct starpu_task *t)
{
	if (t->name)
		return t->name;
	if (t->cl && t->cl->name)
		return t->cl->name;
	if (t->cl && t->cl->model && t->cl->model->symbol)
		return t->cl->model->symbol;
	return NULL;
}

const char *starpu_task_get_model_name(struct starpu_task *t)
{
	if (t->cl && t->cl->model && t->cl->model->symbol)
		return t->cl->model->symbol;
	return starpu_task_get_name(t);
}

void starpu_iteration_push(unsigned long it)
{
	sp_iteration[1] = sp_iteration[0];
	sp_iteration[0] = (long)it;
}

void starpu_iteration_pop(void)
{
	sp_iteration[0] = sp_iteration[1];
	sp_iteration[1] = -1;
}

void starpu_task_ft_failed(struct starpu_task *t)
{
	if (t)
		t->failed = 1;
}

starpu_data_handle_t starpu_task_get_current_data_node_handle(unsigned i)
{
	struct sp_job *j = sp_current_job;
	if (!j || (int)i >= j->nbuffers)
		return NULL;
	return j->buf[i].handle;
}

int starpu_sched_get_min_priority(void) { return INT_MIN; }
int starpu_sched_get_max_priority(void) { return INT_MAX; }

/* ---------------------------------------------------- data acquire -- */

/* A pseudo-job with one buffer goes through the same implicit
 * dependencies as a task; when it becomes ready the application thread
 * fetches the data and holds it until starpu_data_release (reference
 * src/datawizard/user_interactions.c). */
static __thread struct sp_acq {
	starpu_data_handle_t h;
	int node;
	struct sp_job *j;
	struct sp_acq *next;
} *sp_acq_list;

int starpu_data_acquire_on_node(starpu_data_handle_t h, int node, enum starpu_data_access_mode mode)
{
	STARPU_ASSERT_MSG(!h->partitioned, "acquiring a partitioned handle");
	if (node < 0)
		node = 0;
	SP_LOCK();
	struct sp_job *j = job_new(NULL, 1);
	j->is_acquire = 1;
	j->buf[0].handle = h;
	j->buf[0].mode = mode;
	j->buf[0].node = node;
	j->submitted = 1;
	j->ndeps++;
	sp_detect_implicit_deps(j);
	j->ndeps--;
	if (j->ndeps == 0)
		j->acquire_ready = 1;
	while (!j->acquire_ready)
		pthread_cond_wait(&j->cond, &sp.lock);
	int e = sp_fetch_buffer(j, 0, NULL, NULL);
	j->running = 1;
	SP_UNLOCK();
	if (e) {
		sp_job_terminate(j);
		return e;
	}
	struct sp_acq *a = malloc(sizeof(*a));
	a->h = h;
	a->node = node;
	a->j = j;
	a->next = sp_acq_list;
	sp_acq_list = a;
	return 0;
}

int starpu_data_acquire(starpu_data_handle_t h, enum starpu_data_access_mode mode)
{
	return starpu_data_acquire_on_node(h, STARPU_MAIN_RAM, mode);
}

void starpu_data_release_on_node(starpu_data_handle_t h, int node)
{
	struct sp_acq **pp = &sp_acq_list;
	if (node < 0)
		node = 0;
	while (*pp && !((*pp)->h == h && (*pp)->node == node))
		pp = &(*pp)->next;
	STARPU_ASSERT_MSG(*pp, "releasing a handle that this thread did not acquire");
	struct sp_acq *a = *pp;
	*pp = a->next;
	sp_job_terminate(a->j);
	free(a);
}

void starpu_data_release(starpu_data_handle_t h) { starpu_data_release_on_node(h, STARPU_MAIN_RAM); }
