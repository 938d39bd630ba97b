/*
 * Task priorities and critical-chain routing of the tiled Cholesky
 * factorisation, shared by the FP64 and the mixed-precision task graphs
 * (internal header).
 */
#ifndef SA_CHOL_PRIO_H
#define SA_CHOL_PRIO_H

#include <stdint.h>
#include <string.h>

/* Task priorities of the factorisation, the four attributions of
 * cholesky_implicit.c:125-213 (priority_attribution_p), selected by the
 * `prio` argument: 1 base (2nb - 2k - ...), 2 bottom levels of the Kut
 * polyhedral tool (3nb - ...), 3 bottom levels weighted by the calibrated
 * task times (time_init, :79-110), 4 PaRSEC's.  0: every task at the
 * default priority.  The times of attribution 3 are the perfmodels'
 * expected lengths averaged in speed over the workers
 * (extract_task_duration_for_model, :50-66); while any is unknown, the base
 * attribution is used. */
struct chol_prio {
	int scheme;                  /* 0 none, 1 base, 2 Kut, 3 timings, 4 PaRSEC */
	double t_potrf, t_trsm, t_syrk, t_gemm;
};

static inline int prio_potrf(const struct chol_prio *p, int nb, int k)
{
	switch (p->scheme) {
	case 0: return 0;
	case 2: return 3 * nb - 3 * k;
	case 3: return (int)(3 * (p->t_potrf + p->t_trsm + p->t_syrk + p->t_gemm) - (p->t_potrf + p->t_trsm + p->t_gemm) * k);
	case 4: return (nb - k) * (nb - k) * (nb - k);
	default: return 2 * nb - 2 * k;
	}
}

static inline int prio_trsm(const struct chol_prio *p, int nb, int k, int m)
{
	switch (p->scheme) {
	case 0: return 0;
	case 2: return 3 * nb - (2 * k + m);
	case 3:
		return (int)(3 * (p->t_potrf + p->t_trsm + p->t_syrk + p->t_gemm) -
		             ((p->t_trsm + p->t_gemm) * k + (p->t_potrf + p->t_syrk - p->t_gemm) * m + p->t_gemm - p->t_syrk));
	case 4: return (nb - m) * (nb - m) * (nb - m) + 3 * (m - k) * (2 * nb - k - m - 1);
	default: return 2 * nb - 2 * k - m;
	}
}

static inline int prio_gemm(const struct chol_prio *p, int nb, int k, int m, int n)
{
	switch (p->scheme) {
	case 0: return 0;
	case 2: return 3 * nb - (k + n + m);
	case 3:
		return (int)(3 * (p->t_potrf + p->t_trsm + p->t_syrk + p->t_gemm) -
		             (p->t_gemm * k + p->t_trsm * n + (p->t_potrf + p->t_syrk - p->t_gemm) * m - p->t_syrk + p->t_gemm));
	case 4:
		if (n == m)
			return (nb - m) * (nb - m) * (nb - m) + 3 * (m - k);
		return (nb - m) * (nb - m) * (nb - m) + 3 * (m - n) * (2 * nb - m - n - 3) + 6 * (m - k);
	default: return (n == m) ? 2 * nb - 2 * k - n : 2 * nb - 2 * k - m - n;
	}
}

/* Workers of the critical chain.  When the GPU has a worker set aside
 * (STARPU_HIP_RESERVE_CUS: on CUs of its own, or with -1 on all of them),
 * the chain of each step -- POTRF(k), the TRSM of tile (k+1, k) and the
 * SYRK of tile (k+1, k+1) -- runs on it (STARPU_EXECUTE_ON_WORKER) and
 * every other task on the device's other workers (STARPU_TASK_WORKERIDS):
 * the chain's tasks then never wait in a stream behind update GEMMs.  (With
 * 16 streams, one set aside on all CUs, the N = 32768 BF16x3 factorisation
 * took 77-79 ms against 81.5-82 ms with 12 streams and no routing; CUs of
 * its own slowed the updates more than they sped the chain.)  `len` is 0
 * when there is no such worker on `hip_node` or routing is off: the tasks
 * then go wherever the scheduler puts them. */
struct crit_route {
	int crit;
	uint32_t others[(STARPU_NMAXWORKERS + 31) / 32];
	unsigned len;               /* 0: no routing */
};

static void crit_route_init(struct crit_route *r, int hip_node, int enable)
{
	memset(r, 0, sizeof(*r));
	r->crit = -1;
	if (!enable)
		return;
	int crit = -1;
	for (unsigned w = 0; w < starpu_worker_get_count() && crit < 0; w++)
		if (starpu_worker_get_type((int)w) == STARPU_HIP_WORKER &&
		    (int)starpu_worker_get_memory_node(w) == hip_node &&
		    starpu_hip_reserved_worker(starpu_worker_get_devid((int)w)) == (int)w)
			crit = (int)w;
	if (crit < 0)
		return;
	for (unsigned w = 0; w < starpu_worker_get_count(); w++)
		if ((int)w != crit && starpu_worker_get_type((int)w) == STARPU_HIP_WORKER &&
		    (int)starpu_worker_get_memory_node(w) == hip_node)
			r->others[w / 32] |= 1u << (w % 32);
	r->crit = crit;
	r->len = (unsigned)((starpu_worker_get_count() + 31) / 32);
}

#endif
