/*
 * Solve server: SPD solve requests of different sizes served on one GPU
 * through the runtime (no counterpart in the reference, whose Cholesky
 * example factors one matrix per run).
 *
 * Requests arrive over time (a trace of arrival offsets).  The caller's
 * thread releases each request at its arrival time into a FIFO queue;
 * `slots` worker threads take requests from the queue as soon as a slot
 * and enough of the HBM budget are free (continuous batching: admission
 * does not wait for the requests in service to finish).  A request's
 * footprint is its FP32 factor (U and the diagonal-tile inverses,
 * chol_mixed_factor_bytes) plus its refinement workspace.  Factors stay
 * resident after their request completes, so a later request on the same
 * matrix (a re-solve with new right-hand sides) skips the factorisation;
 * when the budget is short, resident factors that no request uses are
 * evicted least recently used first.  Every reservation is also declared
 * to the runtime on the GPU's memory node (starpu_memory_allocate /
 * starpu_memory_deallocate, src/datawizard/memory_manager.c:66-141), so the
 * node's budget (STARPU_LIMIT_HIP_MEM) bounds the server's factors and the
 * runtime's own data replicates together.
 *
 * Each request is factored by the mixed-precision task DAG
 * (chol_mixed_factor) and refined to the FP64 backward error target
 * (chol_mixed_refine) on its slot's stream.  The factorisation tasks of
 * all requests in service share the runtime's HIP workers, so their
 * kernels overlap on the device; a request's arithmetic does not depend on
 * what else is in flight (the tiles see the same operations in the same
 * order), so its result is the same alone and in the mix.
 */
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <starpu.h>
#include "starpu_amd_chol.h"
#include "starpu_amd_kernels.h"

struct cache_entry {
	int matrix;                 /* caller's matrix index, -1 for a free entry */
	struct chol_mixed_factor f;
	size_t bytes;
	int pins;                   /* requests in service that use this factor */
	int ready, failed;          /* factorisation finished / failed */
	double last_use;
};

struct server {
	int hip_node;
	const struct chol_serve_config *cfg;
	struct chol_serve_request *req;
	int nreq;
	double t0;
	pthread_mutex_t mu;
	pthread_cond_t cv;
	int *queue, qhead, qtail;
	int released;
	double used, peak;
	struct cache_entry *cache;
	int ncache;
	int inflight;
	int *waited;                /* request waited for memory at the queue head */
	struct chol_serve_stats st;
};

static double now_us(void) { return starpu_timing_now(); }

static int nblocks_for(const struct server *S, int n)
{
	const int nb = S->cfg->nb > 0 ? S->cfg->nb : 1024;
	int b = n / nb;
	return b < 1 ? 1 : b;
}

static struct cache_entry *cache_find(struct server *S, int matrix)
{
	for (int i = 0; i < S->ncache; i++)
		if (S->cache[i].matrix == matrix)
			return &S->cache[i];
	return NULL;
}

static struct cache_entry *cache_new(struct server *S, int matrix)
{
	for (int i = 0; i < S->ncache; i++)
		if (S->cache[i].matrix < 0) {
			memset(&S->cache[i], 0, sizeof(S->cache[i]));
			S->cache[i].matrix = matrix;
			return &S->cache[i];
		}
	return NULL;
}

/* evict the least recently used resident factor no request uses */
static int evict_one(struct server *S)
{
	struct cache_entry *v = NULL;
	for (int i = 0; i < S->ncache; i++) {
		struct cache_entry *e = &S->cache[i];
		if (e->matrix >= 0 && e->pins == 0 && (e->ready || e->failed) && (!v || e->last_use < v->last_use))
			v = e;
	}
	if (!v)
		return 0;
	S->used -= (double)v->bytes;
	starpu_memory_deallocate((unsigned)S->hip_node, v->bytes);
	chol_mixed_factor_release(&v->f);
	v->matrix = -1;
	S->st.evictions++;
	return 1;
}

/* Admit request r (mutex held): reserve its bytes, pin or create its cache
 * entry.  Returns 0 if the budget cannot hold it now. */
static int try_admit(struct server *S, int r, struct cache_entry **entry, size_t *work_bytes, int *creator)
{
	struct chol_serve_request *q = &S->req[r];
	const int nbk = nblocks_for(S, q->n);
	struct cache_entry *e = cache_find(S, q->matrix);
	if (e && e->failed)
		e = NULL;
	const size_t wb = chol_mixed_refine_worksize(q->n, nbk, q->nrhs);
	const size_t fb = e ? 0 : chol_mixed_factor_bytes(q->n, nbk);
	const double need = (double)wb + (double)fb;
	const long long node_total = starpu_memory_get_total((unsigned)S->hip_node);
	if (need > S->cfg->budget_bytes || (node_total >= 0 && need > (double)node_total))
		return -1;          /* can never fit */
	while (S->used + need > S->cfg->budget_bytes)
		if (!evict_one(S))
			return 0;
	/* the node's budget, shared with the runtime's replicates */
	while (starpu_memory_allocate((unsigned)S->hip_node, (size_t)need, 0) != 0)
		if (!evict_one(S))
			return 0;
	if (!e) {
		/* a failed entry for this matrix is replaced */
		struct cache_entry *old = cache_find(S, q->matrix);
		if (old && old->pins == 0) {
			S->used -= (double)old->bytes;
			starpu_memory_deallocate((unsigned)S->hip_node, old->bytes);
			chol_mixed_factor_release(&old->f);
			old->matrix = -1;
		}
		e = cache_new(S, q->matrix);
		if (!e)
			return 0;
		e->bytes = fb;
		*creator = 1;
	} else {
		*creator = 0;
		S->st.cache_hits++;
	}
	e->pins++;
	S->used += need;
	if (S->used > S->peak)
		S->peak = S->used;
	*entry = e;
	*work_bytes = wb;
	return 1;
}

static void *worker(void *arg)
{
	struct server *S = arg;
	const struct chol_serve_config *cfg = S->cfg;
	hipStream_t s;
	if (sa_stream_create(&s))
		return NULL;
	/* the refinement workspace of this slot, kept between its requests and
	 * grown when a larger one comes: a device free per request would wait
	 * for the whole GPU (the requests of the other slots included) */
	void *work = NULL;
	size_t work_cap = 0;
	for (;;) {
		pthread_mutex_lock(&S->mu);
		int r = -1, creator = 0;
		struct cache_entry *e = NULL;
		size_t wb = 0;
		for (;;) {
			if (S->qhead < S->qtail) {
				const int head = S->queue[S->qhead];
				const int a = try_admit(S, head, &e, &wb, &creator);
				if (a != 0) {
					S->qhead++;
					r = head;
					if (a < 0) {              /* larger than the whole budget */
						S->req[r].status = -ENOMEM;
						e = NULL;
					}
					break;
				}
				if (!S->waited[head]) {
					S->waited[head] = 1;
					S->st.memory_waits++;
				}
			} else if (S->released == S->nreq) {
				break;
			}
			pthread_cond_wait(&S->cv, &S->mu);
		}
		if (r < 0) {
			pthread_mutex_unlock(&S->mu);
			break;
		}
		S->inflight++;
		if (S->inflight > S->st.max_inflight)
			S->st.max_inflight = S->inflight;
		pthread_mutex_unlock(&S->mu);

		struct chol_serve_request *q = &S->req[r];
		q->start_us = now_us() - S->t0;
		if (!e) {
			q->finish_us = now_us() - S->t0;
			pthread_mutex_lock(&S->mu);
			S->inflight--;
			pthread_cond_broadcast(&S->cv);
			pthread_mutex_unlock(&S->mu);
			continue;
		}
		const int nbk = nblocks_for(S, q->n);
		if (creator) {
			struct chol_mixed_factor f;
			q->factor_start_us = now_us() - S->t0;
			const int rc = chol_mixed_factor(S->hip_node, q->A, q->lda, q->n, nbk, cfg->prec, cfg->prio, s, &f,
			                                 NULL, NULL);
			q->factor_end_us = now_us() - S->t0;
			pthread_mutex_lock(&S->mu);
			if (rc == 0 && f.info == 0) {
				e->f = f;
				e->ready = 1;
			} else {
				if (rc == 0)
					chol_mixed_factor_release(&f);
				e->failed = 1;
				q->status = rc ? rc : f.info;
			}
			S->st.factorizations++;
			q->factored = 1;
			pthread_cond_broadcast(&S->cv);
			pthread_mutex_unlock(&S->mu);
		} else {
			pthread_mutex_lock(&S->mu);
			while (!e->ready && !e->failed)
				pthread_cond_wait(&S->cv, &S->mu);
			if (e->failed)
				q->status = -EDOM;
			pthread_mutex_unlock(&S->mu);
		}
		if (q->status == 0) {
			if (wb > work_cap) {
				if (work)
					sa_dev_free(work);
				work = sa_dev_malloc(wb);
				work_cap = work ? wb : 0;
			}
			if (!work) {
				q->status = -ENOMEM;
			} else {
				q->status = chol_mixed_refine(&e->f, q->A, q->lda, q->B, q->ldb, q->nrhs, q->X, q->ldx, cfg->tol,
				                              cfg->maxiter, s, work, wb, &q->iterations, &q->berr, NULL, 0);
				sa_stream_sync(s);
			}
		}
		q->finish_us = now_us() - S->t0;
		pthread_mutex_lock(&S->mu);
		S->used -= (double)wb;
		starpu_memory_deallocate((unsigned)S->hip_node, wb);
		e->pins--;
		e->last_use = q->finish_us;
		if (e->failed && e->pins == 0) {
			S->used -= (double)e->bytes;
			starpu_memory_deallocate((unsigned)S->hip_node, e->bytes);
			e->matrix = -1;
		}
		S->inflight--;
		pthread_cond_broadcast(&S->cv);
		pthread_mutex_unlock(&S->mu);
	}
	if (work)
		sa_dev_free(work);
	sa_stream_destroy(s);
	return NULL;
}

int chol_serve(int hip_node, const struct chol_serve_config *cfg, struct chol_serve_request *req, int nreq,
               struct chol_serve_stats *st)
{
	if (!cfg || !req || nreq <= 0 || cfg->slots <= 0 || cfg->budget_bytes <= 0 || cfg->prec < 0 || cfg->prec > 2)
		return -EINVAL;
	for (int i = 0; i < nreq; i++) {
		const struct chol_serve_request *q = &req[i];
		if (q->n <= 0 || q->nrhs <= 0 || q->lda < q->n || q->ldb < q->n || q->ldx < q->n || !q->A || !q->B ||
		    !q->X || q->matrix < 0 || (i > 0 && q->arrival_us < req[i - 1].arrival_us))
			return -EINVAL;
	}
	if (!starpu_is_initialized())
		return -ENODEV;
	struct server S;
	memset(&S, 0, sizeof(S));
	S.hip_node = hip_node;
	S.cfg = cfg;
	S.req = req;
	S.nreq = nreq;
	S.queue = calloc((size_t)nreq, sizeof(int));
	S.waited = calloc((size_t)nreq, sizeof(int));
	S.ncache = nreq;
	S.cache = calloc((size_t)nreq, sizeof(*S.cache));
	pthread_t *th = calloc((size_t)cfg->slots, sizeof(pthread_t));
	if (!S.queue || !S.waited || !S.cache || !th) {
		free(S.queue), free(S.waited), free(S.cache), free(th);
		return -ENOMEM;
	}
	for (int i = 0; i < nreq; i++) {
		S.cache[i].matrix = -1;
		req[i].status = 0;
		req[i].factored = 0;
		req[i].iterations = 0;
		req[i].berr = 0.0;
		req[i].factor_start_us = req[i].factor_end_us = 0.0;
	}
	pthread_mutex_init(&S.mu, NULL);
	pthread_cond_init(&S.cv, NULL);
	S.t0 = now_us();
	int nth = 0;
	for (; nth < cfg->slots; nth++)
		if (pthread_create(&th[nth], NULL, worker, &S))
			break;
	/* release the requests at their arrival times */
	for (int i = 0; i < nreq; i++) {
		const double wait = req[i].arrival_us - (now_us() - S.t0);
		if (wait > 0) {
			struct timespec ts = {(time_t)(wait / 1e6), (long)(fmod(wait, 1e6) * 1e3)};
			nanosleep(&ts, NULL);
		}
		pthread_mutex_lock(&S.mu);
		S.queue[S.qtail++] = i;
		S.released++;
		pthread_cond_broadcast(&S.cv);
		pthread_mutex_unlock(&S.mu);
	}
	for (int i = 0; i < nth; i++)
		pthread_join(th[i], NULL);
	S.st.wall_us = now_us() - S.t0;
	S.st.peak_bytes = S.peak;
	for (int i = 0; i < S.ncache; i++)
		if (S.cache[i].matrix >= 0) {
			starpu_memory_deallocate((unsigned)hip_node, S.cache[i].bytes);
			chol_mixed_factor_release(&S.cache[i].f);
		}
	pthread_mutex_destroy(&S.mu);
	pthread_cond_destroy(&S.cv);
	if (st)
		*st = S.st;
	free(S.queue);
	free(S.waited);
	free(S.cache);
	free(th);
	return nth == cfg->slots ? 0 : -EAGAIN;
}
