/*
 * Task priorities of the tiled Cholesky factorisation, shared by the FP64
 * and the mixed-precision task graphs (internal header).
 */
#ifndef SA_CHOL_PRIO_H
#define SA_CHOL_PRIO_H

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

#endif
