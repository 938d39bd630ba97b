/*
 * Runtime stress driver for the sanitizer builds (ThreadSanitizer, and
 * AddressSanitizer with UndefinedBehaviorSanitizer): the C runtime and the
 * Cholesky application compiled with instrumentation, CPU workers only,
 * running what tests/test_runtime_cpu.py runs through ctypes:
 *   - the tiled factor + solve DAG under each scheduling policy, on uneven
 *     tiles, bit-identical to the oracle's sequential tiled factor and solve;
 *   - synchronous tasks from starpu_task_create (the waiter frees them);
 *   - readers and writers on one handle under implicit dependencies;
 *   - fault-tolerant retry, tags, explicit task dependencies, callbacks.
 * Exit status 0 when every check passes; the sanitizers report (and, with
 * halt_on_error, fail the run) on any race or memory error.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <starpu.h>
#include <starpu_amd_chol.h>

#include "chol_oracle.h"

/* tests/helpers/rt_codelets.c */
int tc_sync_tasks(starpu_data_handle_t h, int ntasks);
int tc_insert_write(starpu_data_handle_t h, int id);
int tc_insert_read(starpu_data_handle_t h, int id);
void tc_log_reset(void);
int tc_log_get(int *out, int max);
int tc_ft_run(starpu_data_handle_t h, int nfail);
int tc_tags(void);
int tc_task_deps(void);

/* oracle/oracle_codelets.c */
void or_cl_potrf(void *descr[], void *arg);
void or_cl_trsm(void *descr[], void *arg);
void or_cl_syrk(void *descr[], void *arg);
void or_cl_gemm(void *descr[], void *arg);
void or_cl_trtri(void *descr[], void *arg);
void or_cl_solve_fwd(void *descr[], void *arg);
void or_cl_solve_bwd(void *descr[], void *arg);

static int failures;

#define CHECK(cond, ...)                                                   \
	do {                                                                   \
		if (!(cond)) {                                                     \
			fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);           \
			fprintf(stderr, __VA_ARGS__);                                  \
			fprintf(stderr, "\n");                                         \
			failures++;                                                    \
		}                                                                  \
	} while (0)

static unsigned long long rng_state = 88172645463325252ull;
static double urand(void)
{
	rng_state ^= rng_state << 13;
	rng_state ^= rng_state >> 7;
	rng_state ^= rng_state << 17;
	return (double)(rng_state >> 11) / 9007199254740992.0 * 2.0 - 1.0;
}

static void start(const char *sched, int ncpu)
{
	char buf[16];
	setenv("STARPU_SCHED", sched, 1);
	snprintf(buf, sizeof(buf), "%d", ncpu);
	setenv("STARPU_NCPU", buf, 1);
	setenv("STARPU_NHIP", "0", 1);
	setenv("STARPU_WATCHDOG_TIMEOUT", "0", 1);
	if (starpu_init(NULL) != 0) {
		fprintf(stderr, "starpu_init failed\n");
		exit(2);
	}
	struct chol_cpu_kernels k = {or_cl_potrf, or_cl_trsm, or_cl_syrk, or_cl_gemm,
	                             or_cl_solve_fwd, or_cl_solve_bwd, or_cl_trtri};
	chol_set_cpu_kernels(&k);
}

static void stop(void)
{
	chol_set_cpu_kernels(NULL);
	starpu_shutdown();
}

/* factor + solve of one seeded SPD matrix on the CPU workers; the factor's
 * lower triangle and X must equal the oracle's bit for bit */
static void dag_case(const char *sched, int n, int nblocks, int nrhs, int prio)
{
	double *A = malloc(sizeof(double) * n * n), *A0 = malloc(sizeof(double) * n * n);
	double *Bm = malloc(sizeof(double) * n * n);
	double *X = malloc(sizeof(double) * n * nrhs), *X0 = malloc(sizeof(double) * n * nrhs);
	for (int i = 0; i < n * n; i++)
		Bm[i] = urand();
	for (int j = 0; j < n; j++)
		for (int i = 0; i < n; i++) {
			double s = i == j ? n : 0.0;
			for (int p = 0; p < n; p++)
				s += Bm[i + p * n] * Bm[j + p * n];
			A[i + j * n] = A0[i + j * n] = s;
		}
	for (int i = 0; i < n * nrhs; i++)
		X[i] = X0[i] = urand();
	struct chol_stats st;
	int r = chol_factor_solve(STARPU_MAIN_RAM, A, n, n, nblocks, X, n, nrhs, prio, &st);
	CHECK(r == 0 && st.info == 0, "%s n=%d nb=%d: chol_factor_solve %d info %d", sched, n, nblocks, r, st.info);
	int info = or_tiled_cholesky(n, nblocks, A0, n);
	or_tiled_potrs(n, nblocks, nrhs, A0, n, X0, n);
	CHECK(info == 0, "oracle info %d", info);
	long diff = 0;
	for (int j = 0; j < n; j++)
		for (int i = j; i < n; i++)
			diff += A[i + j * n] != A0[i + j * n];
	CHECK(diff == 0, "%s n=%d nb=%d: %ld factor entries differ from the oracle", sched, n, nblocks, diff);
	CHECK(memcmp(X, X0, sizeof(double) * n * nrhs) == 0, "%s n=%d nb=%d: X differs from the oracle", sched, n,
	      nblocks);
	free(A);
	free(A0);
	free(Bm);
	free(X);
	free(X0);
}

static void task_cases(void)
{
	int v[4] = {0, 0, 0, 0};
	starpu_data_handle_t h;
	starpu_vector_data_register(&h, STARPU_MAIN_RAM, (uintptr_t)v, 4, sizeof(int));
	CHECK(tc_sync_tasks(h, 100) == 0, "synchronous tasks");
	for (int i = 1; i <= 60; i++) {
		if (i % 3 == 0)
			CHECK(tc_insert_write(h, i) == 0, "insert write");
		else
			CHECK(tc_insert_read(h, i) == 0, "insert read");
	}
	starpu_task_wait_for_all();
	starpu_data_unregister(h);
	CHECK(v[0] == 100 + 20, "element 0 is %d, want 120", v[0]);

	double x = 0.0;
	starpu_data_handle_t hx;
	starpu_variable_data_register(&hx, STARPU_MAIN_RAM, (uintptr_t)&x, sizeof(x));
	CHECK(tc_ft_run(hx, 2) == 3, "fault-tolerant retry runs");
	starpu_data_unregister(hx);
	CHECK(x == 42.0, "retry result %g", x);

	tc_log_reset();
	CHECK(tc_tags() == 0, "tags");
	tc_log_reset();
	CHECK(tc_task_deps() == 2, "task deps callbacks");
}

int main(void)
{
	const char *scheds[] = {"dmda", "eager", "prio", "random"};
	for (int s = 0; s < 4; s++) {
		start(scheds[s], 4);
		dag_case(scheds[s], 300, 7, 3, 1);
		dag_case(scheds[s], 256, 4, 16, 3);
		task_cases();
		stop();
	}
	start("dmda", 8);
	for (int prio = 0; prio <= 4; prio++)
		dag_case("dmda-prio", 192, 6, 2, prio);
	stop();
	if (failures) {
		fprintf(stderr, "rt_stress: %d check(s) failed\n", failures);
		return 1;
	}
	printf("rt_stress: all checks passed\n");
	return 0;
}
