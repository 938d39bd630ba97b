/*
 * Small CPU codelets used by tests/test_runtime_cpu.py to exercise the
 * runtime (ordering under implicit dependencies, fault-tolerant retry,
 * scheduler placement with fake cost models, tags, callbacks).  Built by
 * the test into a temporary directory with gcc.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <starpu.h>

/* ---- ordering log: each task appends its id to a shared log vector ---- */
static pthread_mutex_t log_lock = PTHREAD_MUTEX_INITIALIZER;
static int log_buf[65536];
static int log_len;

void tc_log_reset(void)
{
	pthread_mutex_lock(&log_lock);
	log_len = 0;
	pthread_mutex_unlock(&log_lock);
}

int tc_log_get(int *out, int max)
{
	pthread_mutex_lock(&log_lock);
	int n = log_len < max ? log_len : max;
	memcpy(out, log_buf, (size_t)n * sizeof(int));
	pthread_mutex_unlock(&log_lock);
	return n;
}

static void log_append(int id)
{
	pthread_mutex_lock(&log_lock);
	if (log_len < 65536)
		log_buf[log_len++] = id;
	pthread_mutex_unlock(&log_lock);
}

static void sleep_us(long us)
{
	struct timespec ts = {us / 1000000, (us % 1000000) * 1000};
	nanosleep(&ts, NULL);
}

/* buffer 0: vector of ints (R or RW); cl_arg: int id.  Writers append
 * their id to element 0..; readers log -id.  A small random sleep makes
 * out-of-order execution likely if dependencies were missing. */
static void cpu_access(void *descr[], void *arg)
{
	int id = *(int *)arg;
	sleep_us(50 + (id * 7919) % 200);
	int *v = (int *)STARPU_VECTOR_GET_PTR(descr[0]);
	size_t n = STARPU_VECTOR_GET_NX(descr[0]);
	if (n > 0)
		v[0] += 1;
	log_append(id);
}

static void cpu_read(void *descr[], void *arg)
{
	int id = *(int *)arg;
	sleep_us(50 + (id * 104729) % 200);
	(void)descr;
	log_append(-id);
}

struct starpu_codelet tc_cl_write = {.cpu_funcs = {cpu_access}, .nbuffers = 1, .modes = {STARPU_RW}, .name = "write"};
struct starpu_codelet tc_cl_read = {.cpu_funcs = {cpu_read}, .nbuffers = 1, .modes = {STARPU_R}, .name = "read"};

int tc_submit_write(starpu_data_handle_t h, int id)
{
	return starpu_task_insert(&tc_cl_write, STARPU_RW, h, STARPU_VALUE, &id, sizeof(id), 0) == 0
	       ? 0 : -1;
}

/* cl_arg via STARPU_VALUE is packed; unpack the id */
static void cpu_access_packed(void *descr[], void *arg)
{
	int id;
	starpu_codelet_unpack_args(arg, &id, 0);
	cpu_access(descr, &id);
}

static void cpu_read_packed(void *descr[], void *arg)
{
	int id;
	starpu_codelet_unpack_args(arg, &id, 0);
	cpu_read(descr, &id);
}

struct starpu_codelet tc_cl_write_p = {.cpu_funcs = {cpu_access_packed}, .nbuffers = 1, .modes = {STARPU_RW}, .name = "write"};
struct starpu_codelet tc_cl_read_p = {.cpu_funcs = {cpu_read_packed}, .nbuffers = 1, .modes = {STARPU_R}, .name = "read"};

int tc_insert_write(starpu_data_handle_t h, int id)
{
	return starpu_task_insert(&tc_cl_write_p, STARPU_RW, h, STARPU_VALUE, &id, sizeof(id), 0);
}

int tc_insert_read(starpu_data_handle_t h, int id)
{
	return starpu_task_insert(&tc_cl_read_p, STARPU_R, h, STARPU_VALUE, &id, sizeof(id), 0);
}

/* ---- fault tolerance: fail the first `nfail` executions ---- */
static int ft_runs;

static void cpu_ft(void *descr[], void *arg)
{
	int nfail = *(int *)arg;
	int run = __sync_fetch_and_add(&ft_runs, 1);
	double *x = (double *)STARPU_VARIABLE_GET_PTR(descr[0]);
	if (run < nfail) {
		*x = -1.0;    /* garbage the retry must overwrite */
		starpu_task_ft_failed(starpu_task_get_current());
		return;
	}
	*x = 42.0;
}

struct starpu_codelet tc_cl_ft = {.cpu_funcs = {cpu_ft}, .nbuffers = 1, .modes = {STARPU_W}, .name = "ft"};

int tc_ft_run(starpu_data_handle_t h, int nfail)
{
	ft_runs = 0;
	struct starpu_task *t = starpu_task_create();
	t->cl = &tc_cl_ft;
	t->handles[0] = h;
	int *arg = malloc(sizeof(int));
	*arg = nfail;
	t->cl_arg = arg;
	t->cl_arg_size = sizeof(int);
	t->cl_arg_free = 1;
	t->detach = 0;
	t->destroy = 1;
	int r = starpu_task_submit(t);
	if (r)
		return r;
	starpu_task_wait(t);
	return ft_runs;
}

/* ---- placement with fake per-worker cost models ---- */
static int placed[1024];
static int nplaced;

static void cpu_place(void *descr[], void *arg)
{
	(void)descr;
	(void)arg;
	int w = starpu_worker_get_id();
	int k = __sync_fetch_and_add(&nplaced, 1);
	if (k < 1024)
		placed[k] = w;
	sleep_us(200);
}

static int fast_worker = 0;

static double worker_cost(struct starpu_task *t, unsigned workerid, unsigned nimpl)
{
	(void)t;
	(void)nimpl;
	/* like tests/sched_policies/simple_cpu_gpu_sched.c: one worker is
	 * predicted 1 us, the others 1e7 us */
	return (int)workerid == fast_worker ? 1.0 : 1e7;
}

static struct starpu_perfmodel fake_model = {.type = STARPU_PER_WORKER, .worker_cost_function = worker_cost};
struct starpu_codelet tc_cl_place = {.cpu_funcs = {cpu_place}, .nbuffers = 0, .model = &fake_model, .name = "place"};

int tc_place(int ntasks, int fast, int *out)
{
	fast_worker = fast;
	nplaced = 0;
	for (int i = 0; i < ntasks; i++)
		if (starpu_task_insert(&tc_cl_place, 0))
			return -1;
	starpu_task_wait_for_all();
	int n = nplaced < 1024 ? nplaced : 1024;
	memcpy(out, placed, (size_t)n * sizeof(int));
	return n;
}

/* ---- tags: chain T3 <- {T1, T2}; each task logs its tag ---- */
static void cpu_tag(void *descr[], void *arg)
{
	(void)descr;
	int id = (int)(uintptr_t)arg;
	sleep_us(id == 1 ? 20000 : 100);
	log_append(id);
}

struct starpu_codelet tc_cl_tag = {.cpu_funcs = {cpu_tag}, .nbuffers = 0, .name = "tag"};

int tc_tags(void)
{
	starpu_tag_declare_deps((starpu_tag_t)3, 2, (starpu_tag_t)1, (starpu_tag_t)2);
	for (int id = 3; id >= 1; id--) {
		struct starpu_task *t = starpu_task_create();
		t->cl = &tc_cl_tag;
		t->cl_arg = (void *)(uintptr_t)id;
		t->use_tag = 1;
		t->tag_id = (starpu_tag_t)id;
		if (starpu_task_submit(t))
			return -1;
	}
	starpu_tag_wait((starpu_tag_t)3);
	return 0;
}

/* ---- callbacks and explicit task dependencies ---- */
static int cb_count;
static void cb(void *arg) { __sync_fetch_and_add(&cb_count, *(int *)arg); }

static void cpu_noop(void *descr[], void *arg)
{
	(void)descr;
	int id = (int)(uintptr_t)arg;
	sleep_us(id == 0 ? 20000 : 100);
	log_append(id);
}

struct starpu_codelet tc_cl_noop = {.cpu_funcs = {cpu_noop}, .nbuffers = 0, .name = "noop"};

int tc_task_deps(void)
{
	static int one = 1;
	cb_count = 0;
	struct starpu_task *a = starpu_task_create(), *b = starpu_task_create();
	a->cl = b->cl = &tc_cl_noop;
	a->cl_arg = (void *)(uintptr_t)0;
	b->cl_arg = (void *)(uintptr_t)1;
	a->callback_func = b->callback_func = cb;
	a->callback_arg = b->callback_arg = &one;
	a->detach = 0;
	a->destroy = 0;
	starpu_task_declare_deps_array(b, 1, &a);
	if (starpu_task_submit(b) || starpu_task_submit(a))
		return -1;
	starpu_task_wait_for_all();
	starpu_task_destroy(a);
	return cb_count;
}

/* ---- watchdog: one task that sleeps long ---- */
static void cpu_stall(void *descr[], void *arg)
{
	(void)descr;
	sleep_us((long)(uintptr_t)arg);
}

struct starpu_codelet tc_cl_stall = {.cpu_funcs = {cpu_stall}, .nbuffers = 0, .name = "stall"};

int tc_stall(long us)
{
	struct starpu_task *t = starpu_task_create();
	t->cl = &tc_cl_stall;
	t->cl_arg = (void *)(uintptr_t)us;
	int r = starpu_task_submit(t);
	starpu_task_wait_for_all();
	return r;
}

/* ---- synchronous tasks from starpu_task_create (detach left at its
 * default, destroy = 1), the pattern of the reference's basic examples
 * (examples/basic_examples/vector_scal.c:164-165): the submitter blocks in
 * starpu_task_submit and the task is freed once, after the wait ---- */
int tc_sync_tasks(starpu_data_handle_t h, int ntasks)
{
	static int id = 1;
	for (int i = 0; i < ntasks; i++) {
		struct starpu_task *t = starpu_task_create();
		t->cl = &tc_cl_write;
		t->handles[0] = h;
		t->cl_arg = &id;
		t->cl_arg_size = sizeof(id);
		t->synchronous = 1;
		if (starpu_task_submit(t))
			return -1;
	}
	return 0;
}

/* ---- placement restricted by STARPU_TASK_WORKERIDS: the cost model
 * prefers worker `fast`, the mask excludes it ---- */
int tc_place_mask(int ntasks, int fast, unsigned mask, int *out)
{
	static uint32_t bits;
	bits = mask;
	fast_worker = fast;
	nplaced = 0;
	for (int i = 0; i < ntasks; i++)
		if (starpu_task_insert(&tc_cl_place, STARPU_TASK_WORKERIDS, 1u, &bits, 0))
			return -1;
	starpu_task_wait_for_all();
	int n = nplaced < 1024 ? nplaced : 1024;
	memcpy(out, placed, (size_t)n * sizeof(int));
	return n;
}
