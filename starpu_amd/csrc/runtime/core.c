/*
 * Runtime bring-up and tear-down, configuration from the environment,
 * worker enumeration, timing and small utilities.
 *
 * Initialisation order mirrors the reference's starpu_initialize
 * (src/core/workers.c:1632-1990): configuration and environment, memory
 * nodes and devices, workers, scheduling policy, performance models, then
 * the worker threads.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "sp_internal.h"

struct sp_runtime sp = {
	.lock = PTHREAD_MUTEX_INITIALIZER,
	.done_cond = PTHREAD_COND_INITIALIZER,
	.mem_cond = PTHREAD_COND_INITIALIZER,
	.default_seq_consistency = 1,
};
__thread struct sp_worker *sp_current_worker;
__thread struct sp_job *sp_current_job;

double sp_now(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec * 1e6 + (double)ts.tv_nsec * 1e-3;
}

double starpu_timing_now(void) { return sp_now(); }

int sp_env_int(const char *name, int def)
{
	const char *s = getenv(name);
	if (!s || !*s)
		return def;
	return atoi(s);
}

double sp_env_double(const char *name, double def)
{
	const char *s = getenv(name);
	if (!s || !*s)
		return def;
	return atof(s);
}

int starpu_get_env_number(const char *str) { return sp_env_int(str, -1); }
int starpu_get_env_number_default(const char *str, int defval) { return sp_env_int(str, defval); }

int starpu_conf_init(struct starpu_conf *conf)
{
	if (!conf)
		return -EINVAL;
	memset(conf, 0, sizeof(*conf));
	conf->magic = 42;
	conf->sched_policy_name = NULL;
	conf->ncpus = -1;
	conf->nhip = -1;
	conf->ncuda = 0;
	conf->nopencl = 0;
	conf->nworker_per_hip = -1;
	conf->hip_pipeline = -1;
	conf->calibrate = -1;
	conf->bus_calibrate = -1;
	conf->limit_hip_mem_mb = -1;
	conf->watchdog_timeout_us = 0;
	return 0;
}

static int pick(int conf_val, const char *env, int def, int precedence)
{
	int e = sp_env_int(env, INT_MIN);
	if (precedence && conf_val >= 0)
		return conf_val;
	if (e != INT_MIN)
		return e;
	if (conf_val >= 0)
		return conf_val;
	return def;
}

void sp_wake_all_workers(void)
{
	for (int i = 0; i < sp.nworkers; i++)
		pthread_cond_broadcast(&sp.workers[i].cond);
	for (int d = 0; d < sp.nhipdevs; d++)
		for (int t = 0; t < sp.hipdevs[d].nthreads; t++)
			pthread_cond_broadcast(&sp.hipdevs[d].thr[t].cond);
}

static void *watchdog_main(void *arg)
{
	(void)arg;
	/* Reference src/core/task.c:2150-2218: warn when no task completed
	 * for STARPU_WATCHDOG_TIMEOUT microseconds while tasks are pending;
	 * STARPU_WATCHDOG_CRASH turns the warning into an abort. */
	const int crash = sp_env_int("STARPU_WATCHDOG_CRASH", 0);
	SP_LOCK();
	while (sp.watchdog_running) {
		struct timespec ts;
		clock_gettime(CLOCK_REALTIME, &ts);
		long long ns = ts.tv_nsec + (long long)(sp.watchdog_us / 4 + 1) * 1000LL;
		ts.tv_sec += ns / 1000000000LL;
		ts.tv_nsec = ns % 1000000000LL;
		pthread_cond_timedwait(&sp.done_cond, &sp.lock, &ts);
		if (!sp.watchdog_running)
			break;
		double now = sp_now();
		if (sp.nsubmitted > 0 && !sp.paused && now - sp.last_completion > sp.watchdog_us) {
			fprintf(stderr, "[starpu][watchdog] no task finished for %.0f us (%ld pending)\n",
			        now - sp.last_completion, sp.nsubmitted);
			if (crash)
				abort();
			sp.last_completion = now;
		}
	}
	SP_UNLOCK();
	return NULL;
}

int starpu_initialize(struct starpu_conf *user_conf, int *argc, char ***argv)
{
	(void)argc;
	(void)argv;
	struct starpu_conf conf;
	if (user_conf)
		conf = *user_conf;
	else
		starpu_conf_init(&conf);
	SP_LOCK();
	if (sp.initialized) {
		sp.initialized++;
		SP_UNLOCK();
		return 0;
	}
	SP_UNLOCK();

	const int prec = conf.precedence_over_environment_variables;
	const char *pol = conf.sched_policy_name;
	const char *env_pol = getenv("STARPU_SCHED");
	if (!(prec && pol) && env_pol && *env_pol)
		pol = env_pol;
	if (!pol)
		pol = "dmda";
	if (strcmp(pol, "dmda") && strcmp(pol, "eager") && strcmp(pol, "prio") &&
	    strcmp(pol, "random") && strcmp(pol, "dm") && strcmp(pol, "dmdas")) {
		fprintf(stderr, "[starpu] unknown scheduling policy '%s', using dmda\n", pol);
		pol = "dmda";
	}
	sp.policy = (!strcmp(pol, "dm") || !strcmp(pol, "dmdas")) ? "dmda" : pol;
	sp.alpha = sp_env_double("STARPU_SCHED_ALPHA", 1.0);
	sp.beta = sp_env_double("STARPU_SCHED_BETA", 1.0);
	sp.calibrate = pick(conf.calibrate, "STARPU_CALIBRATE", 0, prec);
	sp.calibrate_minimum = sp_env_int("STARPU_CALIBRATE_MINIMUM", 10);
	sp.profiling = sp_env_int("STARPU_PROFILING", 0);
	sp.prefetch = sp_env_int("STARPU_PREFETCH", 1) != 0;
	sp.hip_pipeline = pick(conf.hip_pipeline, "STARPU_HIP_PIPELINE", 2, prec);
	if (sp.hip_pipeline < 1)
		sp.hip_pipeline = 1;
	if (sp.hip_pipeline > 8)
		sp.hip_pipeline = 8;
	sp.watchdog_us = conf.watchdog_timeout_us ? conf.watchdog_timeout_us
	                                          : sp_env_int("STARPU_WATCHDOG_TIMEOUT", 0);

	long ncores = sysconf(_SC_NPROCESSORS_ONLN);
	if (ncores < 1)
		ncores = 1;
	int ncpu = pick(conf.ncpus, "STARPU_NCPU", -1, prec);
	if (ncpu < 0)
		ncpu = sp_env_int("STARPU_NCPUS", -1);
	int nhip = pick(conf.nhip, "STARPU_NHIP", -1, prec);
	int nwph = pick(conf.nworker_per_hip, "STARPU_NWORKER_PER_HIP", 8, prec);
	if (nwph < 1)
		nwph = 1;
	if (nwph > 16)
		nwph = 16;

	/* node 0: main RAM */
	sp.nnodes = 1;
	sp.nodes[0].kind = SP_NODE_CPU;
	sp.nodes[0].devid = -1;
	/* main RAM is unlimited unless STARPU_LIMIT_CPU_MEM (MiB) sets a budget
	 * (src/core/workers.c, memory_manager.c:45) */
	const long long cpu_limit_mb = sp_env_int("STARPU_LIMIT_CPU_MEM", -1);
	sp.nodes[0].total_bytes = cpu_limit_mb >= 0 ? cpu_limit_mb * (1LL << 20) : LLONG_MAX;
	sp.nworkers = 0;
	sp_memory_init();

	/* HIP devices and their workers (one memory node per device) */
	int nhip_workers = 0;
	if (nhip != 0) {
		int r = sp_hip_init(nhip, nwph, &nhip_workers);
		if (r < 0 && nhip > 0) {
			fprintf(stderr, "[starpu] %d HIP device(s) requested but the HIP driver could not start\n", nhip);
			return -ENODEV;
		}
	}
	sp.nhip = nhip_workers;
	/* Leave one core to each HIP driver thread, as the reference does. */
	if (ncpu < 0) {
		int ndrv = 0;
		for (int d = 0; d < sp.nhipdevs; d++)
			ndrv += sp.hipdevs[d].nthreads;
		ncpu = (int)ncores - ndrv;
		if (ncpu < 0)
			ncpu = 0;
	}
	if (ncpu > STARPU_NMAXWORKERS - sp.nworkers)
		ncpu = STARPU_NMAXWORKERS - sp.nworkers;
	for (int i = 0; i < ncpu; i++) {
		struct sp_worker *w = &sp.workers[sp.nworkers];
		memset(w, 0, sizeof(*w));
		w->id = sp.nworkers;
		w->arch = STARPU_CPU_WORKER;
		w->devid = i;
		w->node = 0;
		snprintf(w->name, sizeof(w->name), "CPU %d", i);
		pthread_cond_init(&w->cond, NULL);
		w->perf_dev.type = STARPU_CPU_WORKER;
		w->perf_dev.devid = 0;
		w->perf_dev.ncores = 1;
		w->perf_arch.ndevices = 1;
		w->perf_arch.devices = &w->perf_dev;
		sp.nworkers++;
	}
	sp.ncpu = ncpu;
	if (sp.nworkers == 0) {
		sp_hip_shutdown();
		return -ENODEV;
	}
	sp_data_init();

	SP_LOCK();
	sp.initialized = 1;
	sp.shutting_down = 0;
	sp.paused = 0;
	sp.nsubmitted = 0;
	sp.last_completion = sp_now();
	SP_UNLOCK();
	sp_cpu_workers_start();
	sp_hip_start();
	if (sp.watchdog_us > 0) {
		sp.watchdog_running = 1;
		pthread_create(&sp.watchdog_thread, NULL, watchdog_main, NULL);
	}
	return 0;
}

int starpu_init(struct starpu_conf *conf) { return starpu_initialize(conf, NULL, NULL); }
int starpu_is_initialized(void) { return sp.initialized > 0; }

void starpu_shutdown(void)
{
	SP_LOCK();
	if (!sp.initialized) {
		SP_UNLOCK();
		return;
	}
	if (--sp.initialized > 0) {
		SP_UNLOCK();
		return;
	}
	SP_UNLOCK();
	starpu_task_wait_for_all();
	if (sp.watchdog_running) {
		SP_LOCK();
		sp.watchdog_running = 0;
		pthread_cond_broadcast(&sp.done_cond);
		SP_UNLOCK();
		pthread_join(sp.watchdog_thread, NULL);
	}
	SP_LOCK();
	sp.shutting_down = 1;
	sp.paused = 0;
	sp_wake_all_workers();
	SP_UNLOCK();
	for (int i = 0; i < sp.nworkers; i++)
		if (sp.workers[i].arch == STARPU_CPU_WORKER)
			pthread_join(sp.workers[i].thread, NULL);
	sp_hip_stop();
	SP_LOCK();
	sp_perfmodel_save_all();
	sp_perfmodel_free_all();
	sp_tags_free();
	SP_UNLOCK();
	sp_data_shutdown();
	sp_hip_shutdown();
	SP_LOCK();
	free(sp.trace);
	sp.trace = NULL;
	sp.ntrace = sp.trace_cap = 0;
	sp.nworkers = sp.ncpu = sp.nhip = 0;
	sp.nhipdevs = 0;
	SP_UNLOCK();
}

void starpu_pause(void)
{
	SP_LOCK();
	sp.paused = 1;
	SP_UNLOCK();
}

void starpu_resume(void)
{
	SP_LOCK();
	sp.paused = 0;
	sp_wake_all_workers();
	SP_UNLOCK();
}

void starpu_reset_scheduler(void)
{
	SP_LOCK();
	sp_sched_reset();
	SP_UNLOCK();
}

/* ------------------------------------------------------------ workers -- */

unsigned starpu_worker_get_count(void) { return (unsigned)sp.nworkers; }
unsigned starpu_cpu_worker_get_count(void) { return (unsigned)sp.ncpu; }
unsigned starpu_hip_worker_get_count(void) { return (unsigned)sp.nhip; }
unsigned starpu_cuda_worker_get_count(void) { return 0; }
unsigned starpu_opencl_worker_get_count(void) { return 0; }
unsigned starpu_memory_nodes_get_count(void) { return (unsigned)sp.nnodes; }

int starpu_worker_get_id(void) { return sp_current_worker ? sp_current_worker->id : -1; }

unsigned starpu_worker_get_id_check(void)
{
	int id = starpu_worker_get_id();
	STARPU_ASSERT_MSG(id >= 0, "not called from a worker");
	return (unsigned)id;
}

enum starpu_worker_archtype starpu_worker_get_type(int id)
{
	if (id < 0 || id >= sp.nworkers)
		return STARPU_ANY_WORKER;
	return sp.workers[id].arch;
}

int starpu_worker_get_devid(int id)
{
	if (id < 0 || id >= sp.nworkers)
		return -1;
	return sp.workers[id].devid;
}

unsigned starpu_worker_get_memory_node(unsigned id)
{
	if ((int)id >= sp.nworkers)
		return 0;
	return (unsigned)sp.workers[id].node;
}

int starpu_worker_get_by_type(enum starpu_worker_archtype type, int num)
{
	for (int i = 0; i < sp.nworkers; i++)
		if (sp.workers[i].arch == type && num-- == 0)
			return i;
	return -1;
}

unsigned starpu_worker_get_count_by_type(enum starpu_worker_archtype type)
{
	unsigned n = 0;
	for (int i = 0; i < sp.nworkers; i++)
		if (sp.workers[i].arch == type)
			n++;
	return n;
}

void starpu_worker_get_name(int id, char *dst, size_t maxlen)
{
	if (id < 0 || id >= sp.nworkers) {
		snprintf(dst, maxlen, "?");
		return;
	}
	snprintf(dst, maxlen, "%s", sp.workers[id].name);
}

const char *starpu_worker_get_type_as_string(enum starpu_worker_archtype type)
{
	switch (type) {
	case STARPU_CPU_WORKER: return "STARPU_CPU_WORKER";
	case STARPU_HIP_WORKER: return "STARPU_HIP_WORKER";
	case STARPU_CUDA_WORKER: return "STARPU_CUDA_WORKER";
	default: return "STARPU_unknown_WORKER";
	}
}

int starpu_combined_worker_get_size(void) { return 1; }
int starpu_combined_worker_get_rank(void) { return 0; }

int starpu_hip_memory_node(int devid)
{
	for (int d = 0; d < sp.nhipdevs; d++)
		if (sp.hipdevs[d].devid == devid || devid < 0)
			return sp.hipdevs[d].node;
	return -1;
}

void starpu_topology_print(FILE *f)
{
	for (int i = 0; i < sp.nworkers; i++)
		fprintf(f, "worker %d\t%s\tnode %d\n", i, sp.workers[i].name, sp.workers[i].node);
}

struct starpu_perfmodel_arch *starpu_worker_get_perf_archtype(int workerid, unsigned ctx)
{
	(void)ctx;
	if (workerid < 0 || workerid >= sp.nworkers)
		return NULL;
	return &sp.workers[workerid].perf_arch;
}

/* ----------------------------------------------------------- profiling -- */

int starpu_profiling_status_set(int status)
{
	SP_LOCK();
	int old = sp.profiling;
	sp.profiling = status;
	SP_UNLOCK();
	return old;
}

int starpu_profiling_status_get(void) { return sp.profiling; }

void sp_trace_add(struct sp_job *j, struct sp_worker *w)
{
	if (!sp.profiling)
		return;
	if (sp.ntrace == sp.trace_cap) {
		size_t cap = sp.trace_cap ? sp.trace_cap * 2 : 4096;
		struct starpu_trace_record *t = realloc(sp.trace, cap * sizeof(*t));
		if (!t)
			return;
		sp.trace = t;
		sp.trace_cap = cap;
	}
	struct starpu_trace_record *r = &sp.trace[sp.ntrace++];
	struct starpu_task *task = j->task;
	r->name = task->name ? task->name : (task->cl && task->cl->name ? task->cl->name : "task");
	r->workerid = w ? w->id : -1;
	r->submit_us = j->t_submit;
	r->start_us = j->t_start;
	r->end_us = j->t_end;
	r->kernel_us = j->kernel_us;
	r->bytes_in = j->bytes_in;
	r->priority = task->priority;
	r->push_us = j->t_push;
}

size_t starpu_trace_count(void) { return sp.ntrace; }

int starpu_trace_get(size_t i, struct starpu_trace_record *rec)
{
	SP_LOCK();
	int ok = i < sp.ntrace;
	if (ok)
		*rec = sp.trace[i];
	SP_UNLOCK();
	return ok ? 0 : -EINVAL;
}

void starpu_trace_reset(void)
{
	SP_LOCK();
	sp.ntrace = 0;
	SP_UNLOCK();
}

int starpu_trace_dump_chrome(const char *path)
{
	FILE *f = fopen(path, "w");
	if (!f)
		return -errno;
	SP_LOCK();
	double t0 = sp.ntrace ? sp.trace[0].submit_us : 0.0;
	for (size_t i = 0; i < sp.ntrace; i++)
		if (sp.trace[i].submit_us < t0)
			t0 = sp.trace[i].submit_us;
	fprintf(f, "{\"traceEvents\":[\n");
	for (size_t i = 0; i < sp.ntrace; i++) {
		struct starpu_trace_record *r = &sp.trace[i];
		fprintf(f, "%s{\"name\":\"%s\",\"ph\":\"X\",\"pid\":0,\"tid\":%d,\"ts\":%.3f,\"dur\":%.3f,"
		           "\"args\":{\"kernel_us\":%.3f,\"bytes_in\":%zu,\"prio\":%d,\"submit_us\":%.3f}}\n",
		        i ? "," : "", r->name, r->workerid, r->start_us - t0, r->end_us - r->start_us,
		        r->kernel_us, r->bytes_in, r->priority, r->submit_us - t0);
	}
	fprintf(f, "]}\n");
	SP_UNLOCK();
	fclose(f);
	return 0;
}

int starpu_worker_stats(unsigned workerid, unsigned long *ntasks, double *busy_us, double *kernel_us)
{
	if ((int)workerid >= sp.nworkers)
		return -EINVAL;
	SP_LOCK();
	struct sp_worker *w = &sp.workers[workerid];
	if (ntasks)
		*ntasks = w->ntasks;
	if (busy_us)
		*busy_us = w->busy_us;
	if (kernel_us)
		*kernel_us = w->kernel_us;
	SP_UNLOCK();
	return 0;
}

void starpu_profiling_worker_helper_display_summary(void)
{
	for (int i = 0; i < sp.nworkers; i++) {
		struct sp_worker *w = &sp.workers[i];
		fprintf(stderr, "%-12s %8lu tasks  busy %.3f ms  kernel %.3f ms\n", w->name, w->ntasks,
		        w->busy_us / 1e3, w->kernel_us / 1e3);
	}
}

/* --------------------------------------------------------------- hash -- */

/* CRC32C (Castagnoli, polynomial 0x1EDC6F41), computed bit by bit from the
 * most significant bit, as the reference's footprint hash does
 * (src/common/hash.c:22-81); a table-free formulation. */
static uint32_t crc32c_be_8(uint8_t byte, uint32_t crc)
{
	crc ^= (uint32_t)byte << 24;
	for (int i = 0; i < 8; i++)
		crc = (crc & 0x80000000u) ? (crc << 1) ^ 0x1EDC6F41u : (crc << 1);
	return crc;
}

uint32_t starpu_hash_crc32c_be_n(const void *input, size_t n, uint32_t crc)
{
	const uint8_t *p = input;
	for (size_t i = 0; i < n; i++)
		crc = crc32c_be_8(p[i], crc);
	return crc;
}

uint32_t starpu_hash_crc32c_be(uint32_t input, uint32_t crc)
{
	const uint8_t *p = (const uint8_t *)&input;
	for (int i = 0; i < 4; i++)
		crc = crc32c_be_8(p[i], crc);
	return crc;
}

uint32_t starpu_hash_crc32c_be_ptr(void *input, uint32_t crc)
{
	return starpu_hash_crc32c_be_n(&input, sizeof(input), crc);
}

uint32_t starpu_hash_crc32c_string(const char *str, uint32_t crc)
{
	for (; *str; str++)
		crc = crc32c_be_8((uint8_t)*str, crc);
	return crc;
}

/* ------------------------------------------------------ host memory -- */

int starpu_malloc_flags(void **A, size_t dim, int flags)
{
	if (!A)
		return -EINVAL;
	if ((flags & STARPU_MALLOC_PINNED) && sp.nhipdevs > 0) {
		if (sp_hip_host_alloc(A, dim ? dim : 1) == 0)
			return 0;
	}
	void *p = NULL;
	if (posix_memalign(&p, 4096, dim ? dim : 1))
		return -ENOMEM;
	*A = p;
	return 0;
}

int starpu_free_flags(void *A, size_t dim, int flags)
{
	(void)dim;
	if (!A)
		return 0;
	if ((flags & STARPU_MALLOC_PINNED) && sp.nhipdevs > 0 && sp_hip_host_free(A) == 0)
		return 0;
	free(A);
	return 0;
}

int starpu_malloc(void **A, size_t dim) { return starpu_malloc_flags(A, dim, STARPU_MALLOC_PINNED); }
int starpu_free_noflag(void *A, size_t dim) { return starpu_free_flags(A, dim, STARPU_MALLOC_PINNED); }
void starpu_malloc_set_align(size_t align) { (void)align; }

/* no-op vendor-library hooks (source compatibility with the reference app) */
void starpu_hipblas_init(void) {}
void starpu_hipblas_shutdown(void) {}
void starpu_hipsolver_init(void) {}
void starpu_hipsolver_shutdown(void) {}
void starpu_cublas_init(void) {}
void starpu_cublas_shutdown(void) {}
void starpu_cusolver_init(void) {}
void starpu_cusolver_shutdown(void) {}
