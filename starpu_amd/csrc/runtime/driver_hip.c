/*
 * HIP worker driver and HIP memory nodes.
 *
 * One memory node per GPU.  Each GPU gets STARPU_NWORKER_PER_HIP workers
 * (default 8), one non-blocking HIP stream each, so that independent tile
 * tasks run concurrently on the device's compute units (a single 1024^3
 * tile GEMM fills only part of an MI355X).  One driver thread per GPU
 * serves all of its workers: it pops tasks, enqueues their data transfers
 * and codelets on the worker's stream (up to STARPU_HIP_PIPELINE tasks in
 * flight per stream), brackets each codelet with HIP events, and polls the
 * events to complete tasks.  The event pair gives the device-side
 * execution time fed to the performance models.
 *
 * Reference contract this driver fills: src/drivers/hip/driver_hip.c
 * (device init :511-601, malloc with margin :813-842, async pipeline with
 * events :1356-1456, run_once polling :1458-1621, STARPU_HIP_PIPELINE :715,
 * STARPU_NWORKER_PER_HIP :219, STARPU_LIMIT_HIP_MEM :456-480); codelets
 * enqueue on starpu_hip_get_local_stream() (include/starpu_hip.h:66).
 */
#define _GNU_SOURCE
#include <errno.h>
#include <sched.h>
#include <sys/prctl.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <hip/hip_runtime_api.h>

#include "sp_internal.h"

/* The pair of events around each task's launches times the task and, the
 * second one, tells the worker that the task is done.  Neither needs a
 * system-scope fence of its own: the kernels' dispatch packets already
 * release their results and acquire their inputs, the host only reads device
 * data through copies that synchronise, and a fence per event writes back
 * and invalidates the caches at every task boundary (slowing the kernels in
 * flight on the other streams and inflating the measured times).  Each
 * driver thread keeps its own pool. */
static hipEvent_t ev_get(struct sp_hipthr *th)
{
	if (th->nev > 0)
		return (hipEvent_t)th->ev[--th->nev];
	hipEvent_t e = NULL;
	if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess)
		return NULL;
	return e;
}

static void ev_put(struct sp_hipthr *th, hipEvent_t e)
{
	if (!e)
		return;
	if (th->nev < SP_EVPOOL)
		th->ev[th->nev++] = e;
	else
		(void)hipEventDestroy(e);
}

/* ------------------------------------------------ runtime wrappers -- */

int sp_hip_malloc(int devid, void **p, size_t bytes)
{
	if (hipSetDevice(devid) != hipSuccess)
		return -ENODEV;
	return hipMalloc(p, bytes) == hipSuccess ? 0 : -ENOMEM;
}

int sp_hip_free(int devid, void *p)
{
	(void)hipSetDevice(devid);
	return hipFree(p) == hipSuccess ? 0 : -EINVAL;
}

int sp_hip_memcpy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t w, size_t h,
                    int kind, void *stream)
{
	hipError_t e;
	/* one contiguous run (a vector, a variable, a whole matrix): the 1-D
	 * copy, which costs a fraction of the 2-D path for small transfers */
	const int flat = h == 1 || (dpitch == w && spitch == w);
	if (stream)
		e = flat ? hipMemcpyAsync(dst, src, w * h, (hipMemcpyKind)kind, (hipStream_t)stream)
		         : hipMemcpy2DAsync(dst, dpitch, src, spitch, w, h, (hipMemcpyKind)kind, (hipStream_t)stream);
	else
		e = flat ? hipMemcpy(dst, src, w * h, (hipMemcpyKind)kind)
		         : hipMemcpy2D(dst, dpitch, src, spitch, w, h, (hipMemcpyKind)kind);
	return e == hipSuccess ? 0 : (int)e;
}

int sp_hip_stream_sync(void *stream)
{
	return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : -EIO;
}

void *sp_hip_event_record(void *stream)
{
	hipEvent_t e = NULL;
	if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
		return NULL;
	if (hipEventRecord(e, (hipStream_t)stream) != hipSuccess) {
		(void)hipEventDestroy(e);
		return NULL;
	}
	return e;
}

int sp_hip_stream_wait_event(void *stream, void *event)
{
	return hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0) == hipSuccess ? 0 : -EIO;
}

int sp_hip_event_done(void *event)
{
	return !event || hipEventQuery((hipEvent_t)event) == hipSuccess;
}

void *sp_hip_transfer_stream(int devid)
{
	for (int d = 0; d < sp.nhipdevs; d++)
		if (sp.hipdevs[d].devid == devid)
			return sp.hipdevs[d].xfer_stream;
	return NULL;
}

void sp_hip_event_destroy(void *event)
{
	if (event)
		(void)hipEventDestroy((hipEvent_t)event);
}

long long sp_hip_total_mem(int devid)
{
	size_t fr = 0, tot = 0;
	if (hipSetDevice(devid) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess)
		return 0;
	return (long long)fr;
}

int sp_hip_host_alloc(void **p, size_t bytes)
{
	return hipHostMalloc(p, bytes, 0) == hipSuccess ? 0 : -ENOMEM;
}

int sp_hip_host_free(void *p) { return hipHostFree(p) == hipSuccess ? 0 : -EINVAL; }

hipStream_t starpu_hip_get_local_stream(void)
{
	struct sp_worker *w = sp_current_worker;
	if (!w || w->arch != STARPU_HIP_WORKER)
		return NULL;
	return (hipStream_t)w->stream;
}

int starpu_hip_get_device(void)
{
	struct sp_worker *w = sp_current_worker;
	return (w && w->arch == STARPU_HIP_WORKER) ? w->devid : -1;
}

void starpu_hip_report_error(const char *func, const char *file, int line, int status)
{
	fprintf(stderr, "[starpu] HIP error %d (%s) in %s (%s:%d)\n", status,
	        hipGetErrorString((hipError_t)status), func, file, line);
	abort();
}

/* ------------------------------------------------------------- init -- */

static void calibrate_bus(int node, int devid)
{
	/* Measure host<->device bandwidth once with a pinned 32 MiB buffer
	 * (the reference's perfmodel_bus.c does a longer sweep). */
	const size_t bytes = 32u << 20;
	void *h = NULL, *d = NULL;
	sp.bw[0][node] = sp.bw[node][0] = 25e3;
	sp.lat[0][node] = sp.lat[node][0] = 10.0;
	if (hipSetDevice(devid) != hipSuccess || hipHostMalloc(&h, bytes, 0) != hipSuccess)
		return;
	if (hipMalloc(&d, bytes) != hipSuccess) {
		(void)hipHostFree(h);
		return;
	}
	memset(h, 0, bytes);
	double best_h2d = 1e30, best_d2h = 1e30, best_lat = 1e30;
	for (int r = 0; r < 3; r++) {
		/* latency: a one-byte round trip */
		double l0 = sp_now();
		if (hipMemcpy(d, h, 1, hipMemcpyHostToDevice) != hipSuccess)
			break;
		double l1 = sp_now();
		if (l1 - l0 < best_lat) best_lat = l1 - l0;
		double t0 = sp_now();
		if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess)
			break;
		double t1 = sp_now();
		if (hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) != hipSuccess)
			break;
		double t2 = sp_now();
		if (t1 - t0 < best_h2d) best_h2d = t1 - t0;
		if (t2 - t1 < best_d2h) best_d2h = t2 - t1;
	}
	if (best_h2d < 1e29) {
		sp.bw[0][node] = (double)bytes / best_h2d;
		sp.bw[node][0] = (double)bytes / best_d2h;
	}
	if (best_lat < 1e29)
		sp.lat[0][node] = sp.lat[node][0] = best_lat;
	(void)hipFree(d);
	(void)hipHostFree(h);
}

/* CU masks of STARPU_HIP_RESERVE_CUS=c: c CUs of every XCD for the
 * reserved stream, the other CUs for the rest.  The reserved set takes, in
 * XCD x, the CUs x*cpx + ((x + j) mod 8), j < c, so that each XCD keeps c
 * of them whether the driver deals the flat mask out to the XCDs in blocks
 * of cpx bits, round robin bit by bit, or applies its first cpx bits to
 * every XCD: no XCD is ever left without CUs for either mask, which a
 * dispatch whose workgroups are dealt to all XCDs needs. */
#define SP_NXCD 8
/* Returns the number of 32-bit mask words in use (0: no reservation). */
static int reserve_masks(int devid, int c, uint32_t *res, uint32_t *rest, int nwords)
{
	hipDeviceProp_t prop;
	if (c <= 0 || hipGetDeviceProperties(&prop, devid) != hipSuccess)
		return 0;
	const int ncu = prop.multiProcessorCount;
	if (ncu % SP_NXCD || ncu > 32 * nwords)
		return 0;
	const int cpx = ncu / SP_NXCD;
	if (cpx < SP_NXCD || c > SP_NXCD || 2 * c > cpx)
		return 0;
	memset(res, 0, (size_t)nwords * sizeof(uint32_t));
	memset(rest, 0, (size_t)nwords * sizeof(uint32_t));
	for (int b = 0; b < ncu; b++)
		rest[b / 32] |= 1u << (b % 32);
	for (int x = 0; x < SP_NXCD; x++)
		for (int j = 0; j < c; j++) {
			const int b = x * cpx + (x + j) % SP_NXCD;
			res[b / 32] |= 1u << (b % 32);
			rest[b / 32] &= ~(1u << (b % 32));
		}
	return (ncu + 31) / 32;
}

int starpu_hip_reserved_worker(int devid)
{
	for (int d = 0; d < sp.nhipdevs; d++)
		if (sp.hipdevs[d].devid == devid)
			return sp.hipdevs[d].reserved_worker;
	return -1;
}

int sp_hip_init(int nhip, int nwph, int *nworkers_created)
{
	*nworkers_created = 0;
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
		return nhip > 0 ? -ENODEV : 0;
	if (nhip < 0 || nhip > count)
		nhip = count;
	if (nhip > STARPU_MAXHIPDEVS)
		nhip = STARPU_MAXHIPDEVS;
	const long long limit_mb = sp_env_int("STARPU_LIMIT_HIP_MEM", -1);
	for (int d = 0; d < nhip; d++) {
		if (hipSetDevice(d) != hipSuccess)
			continue;
		const int node = sp.nnodes++;
		struct sp_node *nd = &sp.nodes[node];
		memset(nd, 0, sizeof(*nd));
		nd->kind = SP_NODE_HIP;
		nd->devid = d;
		long long fr = sp_hip_total_mem(d);
		nd->total_bytes = limit_mb >= 0 ? limit_mb * (1LL << 20) : fr - fr / 10;
		struct sp_hipdev *hd = &sp.hipdevs[sp.nhipdevs++];
		memset(hd, 0, sizeof(*hd));
		hd->devid = d;
		hd->node = node;
		hd->first_worker = sp.nworkers;
		hd->reserved_worker = -1;
		/* STARPU_HIP_RESERVE_CUS=c (c CUs per XCD, 0 = off): the first
		 * stream of the device runs on CUs of its own, for a DAG's
		 * critical chain, and only the tasks placed on it explicitly;
		 * c = -1 keeps that stream for such tasks but on all the CUs (no
		 * split: the chain's tasks just never queue behind updates) */
		uint32_t mres[16], mrest[16];
		const int want_cus = sp_env_int("STARPU_HIP_RESERVE_CUS", 0);
		const int dedicated = nwph >= 3 && want_cus < 0;
		const int reserve = nwph >= 3 && want_cus > 0 ? reserve_masks(d, want_cus, mres, mrest, 16) : 0;
		{
			hipStream_t xs = NULL;   /* default priority, like the workers' */
			if (hipStreamCreateWithFlags(&xs, hipStreamNonBlocking) == hipSuccess)
				hd->xfer_stream = xs;
		}
		for (int k = 0; k < nwph && sp.nworkers < STARPU_NMAXWORKERS; k++) {
			struct sp_worker *w = &sp.workers[sp.nworkers];
			memset(w, 0, sizeof(*w));
			hipStream_t s = NULL;
			/* default priority: streams of different priorities have
			 * hung on these devices */
			if (reserve) {
				/* the size is a count of 32-bit mask words */
				if (hipExtStreamCreateWithCUMask(&s, (uint32_t)reserve, k == 0 ? mres : mrest) != hipSuccess)
					break;
			} else if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
				break;
			}
			w->id = sp.nworkers;
			w->arch = STARPU_HIP_WORKER;
			w->devid = d;
			w->node = node;
			w->subid = k;
			w->stream = s;
			w->reserved = (reserve || dedicated) && k == 0;
			if (w->reserved && dedicated) {
				hd->reserved_worker = w->id;
				fprintf(stderr, "[starpu] HIP %d: worker %d kept for routed tasks\n", d, w->id);
			} else if (w->reserved) {
				hd->reserved_worker = w->id;
				uint32_t got[16] = {0};
				int n = 0;
				if (hipExtStreamGetCUMask(s, (uint32_t)reserve, got) == hipSuccess)
					for (int q = 0; q < reserve; q++)
						n += __builtin_popcount(got[q]);
				fprintf(stderr, "[starpu] HIP %d: worker %d on %d reserved CUs\n", d, w->id, n);
			}
			snprintf(w->name, sizeof(w->name),
			         w->reserved ? (dedicated ? "HIP %d.%d (routed tasks)" : "HIP %d.%d (reserved CUs)") : "HIP %d.%d", d,
			         k);
			pthread_cond_init(&w->cond, NULL);
			w->perf_dev.type = STARPU_HIP_WORKER;
			w->perf_dev.devid = d;
			w->perf_dev.ncores = 1;
			w->perf_arch.ndevices = 1;
			w->perf_arch.devices = &w->perf_dev;
			sp.nworkers++;
			hd->nworkers++;
			(*nworkers_created)++;
		}
		/* driver threads of the device: STARPU_HIP_THREAD_PER_WORKER=1 gives
		 * every worker its own, otherwise STARPU_HIP_DRIVER_THREADS of them
		 * (default 4) share the workers round robin.  One thread spends
		 * about 16 us of host time per task (launch, two event records,
		 * completion polls); with 1024-wide BF16x3 update tiles (~40 us of
		 * GPU time each) a single thread kept only ~4 of 12 streams busy. */
		int nt = sp_env_int("STARPU_HIP_THREAD_PER_WORKER", 0) > 0 ? hd->nworkers
		                                                          : sp_env_int("STARPU_HIP_DRIVER_THREADS", 4);
		if (nt > hd->nworkers)
			nt = hd->nworkers;
		if (nt > SP_HIP_MAXTHREADS)
			nt = SP_HIP_MAXTHREADS;
		if (nt < 1)
			nt = 1;
		hd->nthreads = nt;
		/* round robin; a reserved worker gets a thread of its own when
		 * there are two or more, so its launches never wait behind
		 * another stream's */
		for (int k = 0; k < hd->nworkers; k++) {
			struct sp_worker *w = &sp.workers[hd->first_worker + k];
			if (hd->reserved_worker >= 0 && nt >= 2)
				w->thr = w->reserved ? 0 : 1 + (k - 1) % (nt - 1);
			else
				w->thr = k % nt;
		}
		for (int t = 0; t < nt; t++) {
			hd->thr[t].hd = hd;
			hd->thr[t].index = t;
			hd->thr[t].nev = 0;
			pthread_cond_init(&hd->thr[t].cond, NULL);
		}
	}
	/* the bus model: loaded from the sampling directory when a file for
	 * this host and node set exists, measured and saved otherwise, or when
	 * STARPU_BUS_CALIBRATE=1 (perfmodel_bus.c: _starpu_load_bus_performance_files) */
	if (sp.nhipdevs > 0 && (sp_env_int("STARPU_BUS_CALIBRATE", 0) == 1 || !sp_bus_load())) {
		for (int k = 0; k < sp.nhipdevs; k++)
			calibrate_bus(sp.hipdevs[k].node, sp.hipdevs[k].devid);
		if (sp_bus_save() != 0)
			fprintf(stderr, "[starpu] cannot write the bus model files\n");
	}
	return 0;
}

/* the driver thread that runs worker w's tasks */
struct sp_hipthr *sp_hip_worker_thread(const struct sp_worker *w)
{
	for (int d = 0; d < sp.nhipdevs; d++) {
		struct sp_hipdev *hd = &sp.hipdevs[d];
		if (hd->devid == w->devid && hd->nthreads > 0)
			return &hd->thr[w->thr];
	}
	return NULL;
}

/* ----------------------------------------------------------- driver -- */

/* lock held on entry and return */
static void start_job(struct sp_hipthr *th, struct sp_worker *w, struct sp_job *j)
{
	struct starpu_task *t = j->task;
	j->running = 1;
	t->status = STARPU_TASK_RUNNING;
	for (int i = 0; i < j->nbuffers; i++) {
		int e = sp_fetch_buffer(j, i, w, w->stream);
		STARPU_ASSERT_MSG(e == 0, "cannot fetch buffer %d of task %s on %s (%d)", i,
		                  t->name ? t->name : "?", w->name, e);
	}
	sp_sched_job_started(w, j);
	SP_UNLOCK();
	hipStream_t s = (hipStream_t)w->stream;
	hipEvent_t e0 = ev_get(th), e1 = ev_get(th);
	j->t_start = sp_now();
	if (e0)
		(void)hipEventRecord(e0, s);
	sp_current_worker = w;
	sp_current_job = j;
	if (t->cl) {
		starpu_hip_func_t f = t->cl->hip_funcs[j->nimpl];
		if (f)
			f(j->ifaces, t->cl_arg);
		if (!(t->cl->hip_flags[j->nimpl] & STARPU_HIP_ASYNC))
			(void)hipStreamSynchronize(s);
	}
	sp_current_job = NULL;
	if (e1)
		(void)hipEventRecord(e1, s);
	j->hip_ev_start = e0;
	j->hip_ev_end = e1;
	SP_LOCK();
	w->inflight[w->ninflight++] = j;
}

static void finish_job(struct sp_hipthr *th, struct sp_worker *w, struct sp_job *j)
{
	struct starpu_task *t = j->task;
	float ms = 0.0f;
	j->t_end = sp_now();
	if (j->hip_ev_start && j->hip_ev_end &&
	    hipEventElapsedTime(&ms, (hipEvent_t)j->hip_ev_start, (hipEvent_t)j->hip_ev_end) == hipSuccess)
		j->kernel_us = (double)ms * 1000.0;
	else
		j->kernel_us = j->t_end - j->t_start;
	ev_put(th, (hipEvent_t)j->hip_ev_start);
	ev_put(th, (hipEvent_t)j->hip_ev_end);
	j->hip_ev_start = j->hip_ev_end = NULL;
	SP_LOCK();
	/* a reserved worker runs on a subset of the CUs: its times would skew
	 * the history the other streams of the device share */
	if (!t->failed && !w->reserved)
		sp_perfmodel_feedback(j, w, j->kernel_us);
	w->ntasks++;
	w->busy_us += j->t_end - j->t_start;
	w->kernel_us += j->kernel_us;
	if (sp.profiling) {
		if (!t->profiling_info)
			t->profiling_info = calloc(1, sizeof(*t->profiling_info));
		if (t->profiling_info) {
			t->profiling_info->submit_time = j->t_submit;
			t->profiling_info->push_start_time = j->t_push;
			t->profiling_info->start_time = j->t_start;
			t->profiling_info->end_time = j->t_end;
			t->profiling_info->workerid = w->id;
			t->profiling_info->kernel_us = j->kernel_us;
			t->profiling_info->bytes_transferred = j->bytes_in;
		}
		sp_trace_add(j, w);
	}
	sp_sched_job_done(w, j);
	SP_UNLOCK();
	sp_job_terminate(j);
}

static void *hip_driver_main(void *arg)
{
	struct sp_hipthr *th = arg;
	struct sp_hipdev *hd = th->hd;
	(void)hipSetDevice(hd->devid);
	/* the idle poll below sleeps 2 us; with Linux's default timer slack of
	 * 50 us for ordinary threads every such sleep lasted ~55 us, and the
	 * completion of a task that others wait for was seen that much later */
	(void)prctl(PR_SET_TIMERSLACK, 1UL, 0UL, 0UL, 0UL);
	struct sp_job *done[64];
	struct sp_worker *donew[64];
	int idle_spins = 0;
	for (;;) {
		int started = 0, inflight = 0, want = 0;
		/* take the lock only to pop: when a stream has room and may find a
		 * task, or when nothing is in flight (then possibly to sleep) */
		for (int k = 0; k < hd->nworkers; k++) {
			const struct sp_worker *w = &sp.workers[hd->first_worker + k];
			if (w->thr != th->index)
				continue;
			inflight += w->ninflight;
			if (w->ninflight < sp.hip_pipeline && sp_sched_may_pop(w))
				want = 1;
		}
		if (want || !inflight) {
			SP_LOCK();
			for (int k = 0; k < hd->nworkers; k++) {
				struct sp_worker *w = &sp.workers[hd->first_worker + k];
				if (w->thr != th->index)
					continue;
				while (w->ninflight < sp.hip_pipeline) {
					struct sp_job *j = sp_sched_pop(w);
					if (!j)
						break;
					start_job(th, w, j);
					started++;
				}
			}
			inflight = 0;
			for (int k = 0; k < hd->nworkers; k++)
				if (sp.workers[hd->first_worker + k].thr == th->index)
					inflight += sp.workers[hd->first_worker + k].ninflight;
			if (!started && !inflight) {
				if (sp.shutting_down) {
					SP_UNLOCK();
					break;
				}
				pthread_cond_wait(&th->cond, &sp.lock);
				SP_UNLOCK();
				continue;
			}
			SP_UNLOCK();
		}
		/* poll the in-flight tasks (only this thread touches them); a
		 * stream completes in order, so each worker's poll stops at its
		 * first task still running */
		int ndone = 0;
		for (int k = 0; k < hd->nworkers; k++) {
			struct sp_worker *w = &sp.workers[hd->first_worker + k];
			if (w->thr != th->index)
				continue;
			while (w->ninflight > 0 && ndone < 64) {
				struct sp_job *j = w->inflight[0];
				hipError_t e = j->hip_ev_end ? hipEventQuery((hipEvent_t)j->hip_ev_end) : hipSuccess;
				if (e == hipErrorNotReady)
					break;
				if (e != hipSuccess)
					starpu_hip_report_error("hipEventQuery", __FILE__, __LINE__, (int)e);
				memmove(&w->inflight[0], &w->inflight[1], (size_t)(w->ninflight - 1) * sizeof(w->inflight[0]));
				w->ninflight--;
				done[ndone] = j;
				donew[ndone] = w;
				ndone++;
			}
		}
		for (int i = 0; i < ndone; i++)
			finish_job(th, donew[i], done[i]);
		if (!ndone && !started) {
			/* nothing completed yet: yield briefly, then poll again */
			if (++idle_spins > 64) {
				struct timespec ts = {0, 2000};
				nanosleep(&ts, NULL);
			} else {
				sched_yield();
			}
		} else {
			idle_spins = 0;
		}
	}
	return NULL;
}

int sp_hip_start(void)
{
	for (int d = 0; d < sp.nhipdevs; d++) {
		struct sp_hipdev *hd = &sp.hipdevs[d];
		hd->active = 1;
		for (int t = 0; t < hd->nthreads; t++) {
			int r = pthread_create(&hd->thr[t].thread, NULL, hip_driver_main, &hd->thr[t]);
			STARPU_ASSERT_MSG(r == 0, "cannot start a HIP driver thread");
		}
	}
	return 0;
}

void sp_hip_stop(void)
{
	for (int d = 0; d < sp.nhipdevs; d++) {
		struct sp_hipdev *hd = &sp.hipdevs[d];
		if (!hd->active)
			continue;
		SP_LOCK();
		for (int t = 0; t < hd->nthreads; t++)
			pthread_cond_broadcast(&hd->thr[t].cond);
		SP_UNLOCK();
		for (int t = 0; t < hd->nthreads; t++)
			pthread_join(hd->thr[t].thread, NULL);
		hd->active = 0;
	}
}

void sp_hip_shutdown(void)
{
	for (int d = 0; d < sp.nhipdevs; d++) {
		struct sp_hipdev *hd = &sp.hipdevs[d];
		(void)hipSetDevice(hd->devid);
		(void)hipDeviceSynchronize();
		for (int k = 0; k < hd->nworkers; k++) {
			struct sp_worker *w = &sp.workers[hd->first_worker + k];
			if (w->stream)
				(void)hipStreamDestroy((hipStream_t)w->stream);
			w->stream = NULL;
		}
		if (hd->xfer_stream)
			(void)hipStreamDestroy((hipStream_t)hd->xfer_stream);
		hd->xfer_stream = NULL;
		SP_LOCK();
		sp_node_release_cache(hd->node);
		SP_UNLOCK();
		for (int t = 0; t < hd->nthreads; t++) {
			struct sp_hipthr *th = &hd->thr[t];
			while (th->nev > 0)
				(void)hipEventDestroy((hipEvent_t)th->ev[--th->nev]);
			pthread_cond_destroy(&th->cond);
		}
	}
	/* memory nodes past main RAM are re-created at the next init */
	sp.nnodes = 1;
}
