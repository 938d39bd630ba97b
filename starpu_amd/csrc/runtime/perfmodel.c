/*
 * History-based performance models.
 *
 * Per model, per architecture combination (device type, device id, core
 * count) and per implementation, a table maps a footprint (CRC32C of the
 * buffer shapes, src/datawizard/footprint.c:23-70) to the running mean,
 * deviation and sample count of the measured execution times.
 *
 * Update rule (reference src/core/perfmodel/perfmodel_history.c:1855-2085):
 * the first measurement of a footprint is dropped; a measurement further
 * than STARPU_HISTORY_MAX_ERROR percent (default 50) from the mean counts
 * as an error instead of a sample, and once errors reach the sample count
 * the entry is flushed.  A prediction is the mean once the entry holds at
 * least STARPU_CALIBRATE_MINIMUM (10) samples, NaN before (:1740-1825);
 * a NaN prediction puts the model in calibration mode for the run.
 *
 * Files use the reference's version-45 text layout (:834-968 writer,
 * :1379 reader) at $STARPU_PERF_MODEL_DIR/codelets/45/<symbol>.<host>.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "sp_internal.h"

#define SP_PERFMODEL_VERSION 45

struct sp_hentry {
	uint32_t footprint;
	size_t size;
	double flops, mean, dev, sum, sum2;
	unsigned nsample, nerror;
};

struct sp_htable {
	struct sp_hentry *e;
	int n, cap;
};

struct sp_parch {
	int type, devid, ncores;
	int impl_set[STARPU_MAXIMPLEMENTATIONS];
	struct sp_htable impl[STARPU_MAXIMPLEMENTATIONS];
};

struct sp_mstate {
	struct sp_parch *arch;
	int narch, cap;
	int dirty;
	int registered;
};

static int history_max_error(void) { return sp_env_int("STARPU_HISTORY_MAX_ERROR", 50); }

static struct sp_mstate *mstate(struct starpu_perfmodel *m)
{
	if (!m->state) {
		m->state = calloc(1, sizeof(struct sp_mstate));
		m->is_init = 1;
	}
	return m->state;
}

static struct sp_parch *find_arch(struct sp_mstate *s, int type, int devid, int ncores, int create)
{
	for (int i = 0; i < s->narch; i++)
		if (s->arch[i].type == type && s->arch[i].devid == devid && s->arch[i].ncores == ncores)
			return &s->arch[i];
	if (!create)
		return NULL;
	if (s->narch == s->cap) {
		s->cap = s->cap ? 2 * s->cap : 4;
		s->arch = realloc(s->arch, (size_t)s->cap * sizeof(*s->arch));
	}
	struct sp_parch *a = &s->arch[s->narch++];
	memset(a, 0, sizeof(*a));
	a->type = type;
	a->devid = devid;
	a->ncores = ncores;
	return a;
}

static struct sp_parch *arch_of(struct sp_mstate *s, struct starpu_perfmodel_arch *arch, int create)
{
	STARPU_ASSERT(arch && arch->ndevices >= 1);
	return find_arch(s, arch->devices[0].type, arch->devices[0].devid, arch->devices[0].ncores, create);
}

static struct sp_hentry *find_entry(struct sp_htable *t, uint32_t fp, int create)
{
	for (int i = 0; i < t->n; i++)
		if (t->e[i].footprint == fp)
			return &t->e[i];
	if (!create)
		return NULL;
	if (t->n == t->cap) {
		t->cap = t->cap ? 2 * t->cap : 8;
		t->e = realloc(t->e, (size_t)t->cap * sizeof(*t->e));
	}
	struct sp_hentry *e = &t->e[t->n++];
	memset(e, 0, sizeof(*e));
	e->footprint = fp;
	return e;
}

/* $STARPU_PERF_MODEL_DIR, or $STARPU_HOME/.starpu/sampling (HOME, /tmp) */
static void model_dir(char *dir, size_t maxlen)
{
	const char *d = getenv("STARPU_PERF_MODEL_DIR");
	if (d && *d) {
		snprintf(dir, maxlen, "%s", d);
		return;
	}
	const char *home = getenv("STARPU_HOME");
	if (!home || !*home)
		home = getenv("HOME");
	if (!home)
		home = "/tmp";
	snprintf(dir, maxlen, "%s/.starpu/sampling", home);
}

static void host_name(char *host, size_t maxlen)
{
	const char *h = getenv("STARPU_HOSTNAME");
	if (h && *h) {
		snprintf(host, maxlen, "%s", h);
		return;
	}
	if (gethostname(host, maxlen) != 0)
		snprintf(host, maxlen, "unknown");
	host[maxlen - 1] = 0;
}

void starpu_perfmodel_get_model_path(const char *symbol, char *path, size_t maxlen)
{
	char dir[768], host[256];
	model_dir(dir, sizeof(dir));
	host_name(host, sizeof(host));
	snprintf(path, maxlen, "%s/codelets/%d/%s.%s", dir, SP_PERFMODEL_VERSION, symbol, host);
}

static void mkdir_p(const char *file)
{
	char buf[1024];
	snprintf(buf, sizeof(buf), "%s", file);
	for (char *p = buf + 1; *p; p++)
		if (*p == '/') {
			*p = 0;
			mkdir(buf, 0755);
			*p = '/';
		}
}

/* --------------------------------------------------------- v45 files -- */

/* Next non-comment, non-empty line into buf; 0 at end of file. */
static int next_line(FILE *f, char *buf, size_t n)
{
	while (fgets(buf, (int)n, f)) {
		char *p = buf;
		while (isspace((unsigned char)*p))
			p++;
		if (*p == 0 || *p == '#')
			continue;
		return 1;
	}
	return 0;
}

static int parse_file(FILE *f, struct starpu_perfmodel *m)
{
	char line[1024];
	struct sp_mstate *s = mstate(m);
	int version, ncombs;
	if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &version) != 1 || version != SP_PERFMODEL_VERSION)
		return -EINVAL;
	if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &ncombs) != 1)
		return -EINVAL;
	for (int c = 0; c < ncombs; c++) {
		int ndev, type = 0, devid = 0, ncores = 0, nimpl;
		if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &ndev) != 1 || ndev < 1)
			return -EINVAL;
		for (int d = 0; d < ndev; d++) {
			int t, id, nc;
			if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &t) != 1)
				return -EINVAL;
			if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &id) != 1)
				return -EINVAL;
			if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &nc) != 1)
				return -EINVAL;
			if (d == 0) {
				type = t;
				devid = id;
				ncores = nc;
			}
		}
		if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &nimpl) != 1)
			return -EINVAL;
		struct sp_parch *a = find_arch(s, type, devid, ncores, 1);
		for (int im = 0; im < nimpl; im++) {
			int nentries, mlr;
			if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &nentries) != 1)
				return -EINVAL;
			/* linear regression line, non-linear a b c, multiple
			 * regression flag: kept only to stay in sync */
			if (!next_line(f, line, sizeof(line)) || !next_line(f, line, sizeof(line)))
				return -EINVAL;
			if (!next_line(f, line, sizeof(line)) || sscanf(line, "%d", &mlr) != 1)
				return -EINVAL;
			if (mlr != 0)
				return -EINVAL;     /* multiple-regression models are not used here */
			for (int e = 0; e < nentries; e++) {
				unsigned fp, n;
				unsigned long size;
				double flops, mean, dev, sum, sum2;
				if (!next_line(f, line, sizeof(line)) ||
				    sscanf(line, "%x %lu %le %le %le %le %le %u", &fp, &size, &flops, &mean, &dev, &sum,
				           &sum2, &n) != 8)
					return -EINVAL;
				if (im >= STARPU_MAXIMPLEMENTATIONS)
					continue;
				a->impl_set[im] = 1;
				struct sp_hentry *he = find_entry(&a->impl[im], fp, 1);
				he->size = size;
				he->flops = flops;
				he->mean = mean;
				he->dev = dev;
				he->sum = sum;
				he->sum2 = sum2;
				he->nsample = n;
			}
		}
	}
	return 0;
}

int starpu_perfmodel_load_file(const char *filename, struct starpu_perfmodel *m)
{
	FILE *f = fopen(filename, "r");
	if (!f)
		return -errno;
	int r = parse_file(f, m);
	fclose(f);
	if (r == 0)
		m->is_loaded = 1;
	return r;
}

int starpu_perfmodel_load_symbol(const char *symbol, struct starpu_perfmodel *m)
{
	char path[1024];
	starpu_perfmodel_get_model_path(symbol, path, sizeof(path));
	if (!m->symbol)
		m->symbol = symbol;
	return starpu_perfmodel_load_file(path, m);
}

static const char *type_name(int type)
{
	switch (type) {
	case STARPU_CPU_WORKER: return "cpu";
	case STARPU_CUDA_WORKER: return "cuda";
	case STARPU_OPENCL_WORKER: return "opencl";
	case STARPU_HIP_WORKER: return "hip";
	default: return "unknown";
	}
}

int starpu_perfmodel_dump_file(struct starpu_perfmodel *m, const char *path)
{
	struct sp_mstate *s = mstate(m);
	mkdir_p(path);
	FILE *f = fopen(path, "w");
	if (!f)
		return -errno;
	fprintf(f, "##################\n# Performance Model Version\n%d\n\n", SP_PERFMODEL_VERSION);
	fprintf(f, "####################\n# COMBs\n# number of combinations\n%d\n", s->narch);
	for (int c = 0; c < s->narch; c++) {
		struct sp_parch *a = &s->arch[c];
		int nimpl = 0;
		for (int im = 0; im < STARPU_MAXIMPLEMENTATIONS; im++)
			if (a->impl_set[im])
				nimpl = im + 1;
		if (nimpl == 0)
			nimpl = 1;
		fprintf(f, "####################\n# COMB_%d\n# number of types devices\n1\n", c);
		fprintf(f, "####################\n# DEV_0\n# device type (CPU - 0, CUDA - 1, OPENCL - 2, MIC - 3)\n%d\n", a->type);
		fprintf(f, "####################\n# DEV_0\n# device id \n%d\n", a->devid);
		fprintf(f, "####################\n# DEV_0\n# number of cores \n%d\n", a->ncores);
		fprintf(f, "##########\n# number of implementations\n%d\n", nimpl);
		for (int im = 0; im < nimpl; im++) {
			struct sp_htable *t = &a->impl[im];
			fprintf(f, "#####\n# Model for %s%d_impl%d (Comb%d)\n# number of entries\n%d\n",
			        type_name(a->type), a->devid, im, c, t->n);
			fprintf(f, "# sumlnx\tsumlnx2\t\tsumlny\t\tsumlnxlny\talpha\t\tbeta\t\tn\tminx\t\tmaxx\n");
			fprintf(f, "%-15e\t%-15e\t%-15e\t%-15e\tnan            \tnan            \t0\t0              \t0              \n",
			        0.0, 0.0, 0.0, 0.0);
			fprintf(f, "# a\t\tb\t\tc\nnan            \tnan            \tnan            \n");
			fprintf(f, "# not multiple-regression-base\n0\n");
			fprintf(f, "# hash\t\tsize\t\tflops\t\tmean (us)\tdev (us)\tsum\t\tsum2\t\tn\n");
			for (int e = 0; e < t->n; e++) {
				struct sp_hentry *he = &t->e[e];
				fprintf(f, "%08x\t%-15lu\t%-15e\t%-15e\t%-15e\t%-15e\t%-15e\t%u\n", he->footprint,
				        (unsigned long)he->size, he->flops, he->mean, he->dev, he->sum, he->sum2, he->nsample);
			}
			fprintf(f, "\n");
		}
	}
	fclose(f);
	return 0;
}

int starpu_save_history_based_model(struct starpu_perfmodel *m)
{
	if (!m->symbol)
		return -EINVAL;
	char path[1024];
	starpu_perfmodel_get_model_path(m->symbol, path, sizeof(path));
	return starpu_perfmodel_dump_file(m, path);
}

static void free_state(struct starpu_perfmodel *m)
{
	struct sp_mstate *s = m->state;
	if (!s)
		return;
	for (int a = 0; a < s->narch; a++)
		for (int im = 0; im < STARPU_MAXIMPLEMENTATIONS; im++)
			free(s->arch[a].impl[im].e);
	free(s->arch);
	free(s);
	m->state = NULL;
	m->is_loaded = 0;
	m->is_init = 0;
	m->benchmarking = 0;
}

int starpu_perfmodel_unload_model(struct starpu_perfmodel *m)
{
	SP_LOCK();
	for (int i = 0; i < sp.nmodels; i++)
		if (sp.models[i] == m) {
			sp.models[i] = sp.models[--sp.nmodels];
			break;
		}
	free_state(m);
	SP_UNLOCK();
	return 0;
}

void starpu_perfmodel_init(struct starpu_perfmodel *m)
{
	SP_LOCK();
	sp_perfmodel_register(m);
	SP_UNLOCK();
}

void sp_perfmodel_register(struct starpu_perfmodel *m)
{
	struct sp_mstate *s = mstate(m);
	if (s->registered)
		return;
	s->registered = 1;
	if (m->type == STARPU_HISTORY_BASED && m->symbol && !m->is_loaded) {
		char path[1024];
		starpu_perfmodel_get_model_path(m->symbol, path, sizeof(path));
		FILE *f = fopen(path, "r");
		if (f) {
			if (parse_file(f, m) == 0)
				m->is_loaded = 1;
			else
				fprintf(stderr, "[starpu] ignoring unreadable performance model %s\n", path);
			fclose(f);
		}
	}
	if (sp.nmodels == sp.models_cap) {
		sp.models_cap = sp.models_cap ? 2 * sp.models_cap : 16;
		sp.models = realloc(sp.models, (size_t)sp.models_cap * sizeof(*sp.models));
	}
	sp.models[sp.nmodels++] = m;
}

void sp_perfmodel_save_all(void)
{
	for (int i = 0; i < sp.nmodels; i++) {
		struct starpu_perfmodel *m = sp.models[i];
		struct sp_mstate *s = m->state;
		if (s && s->dirty && m->type == STARPU_HISTORY_BASED && m->symbol &&
		    (sp.calibrate > 0 || m->benchmarking))
			starpu_save_history_based_model(m);
	}
}

void sp_perfmodel_free_all(void)
{
	for (int i = 0; i < sp.nmodels; i++)
		free_state(sp.models[i]);
	free(sp.models);
	sp.models = NULL;
	sp.nmodels = sp.models_cap = 0;
}

/* ------------------------------------------------------- footprints -- */

size_t starpu_task_data_footprint_size(struct starpu_task *t)
{
	size_t s = 0;
	int n = t->cl ? (t->cl->nbuffers == STARPU_VARIABLE_NBUFFERS ? t->nbuffers : t->cl->nbuffers) : 0;
	for (int i = 0; i < n; i++)
		if (!(STARPU_TASK_GET_MODE(t, i) & STARPU_NOFOOTPRINT))
			s += STARPU_TASK_GET_HANDLE(t, i)->size;
	return s;
}

static uint32_t task_data_footprint(struct starpu_task *t)
{
	/* fold the handle footprints of the buffers that carry one
	 * (reference src/datawizard/footprint.c:23-43) */
	uint32_t fp = 0;
	int n = t->cl ? (t->cl->nbuffers == STARPU_VARIABLE_NBUFFERS ? t->nbuffers : t->cl->nbuffers) : 0;
	for (int i = 0; i < n; i++) {
		if (STARPU_TASK_GET_MODE(t, i) & STARPU_NOFOOTPRINT)
			continue;
		fp = starpu_hash_crc32c_be(STARPU_TASK_GET_HANDLE(t, i)->footprint, fp);
	}
	return fp;
}

uint32_t starpu_task_footprint(struct starpu_perfmodel *m, struct starpu_task *t,
                               struct starpu_perfmodel_arch *arch, unsigned nimpl)
{
	(void)arch;
	if (m && m->footprint)
		return m->footprint(t);
	if (m && m->size_base) {
		size_t size = m->size_base(t, nimpl);
		return starpu_hash_crc32c_be_n(&size, sizeof(size), 0);
	}
	return task_data_footprint(t);
}

uint32_t sp_job_footprint(struct sp_job *j)
{
	if (!j->footprint_valid) {
		struct starpu_perfmodel *m = j->task->cl ? j->task->cl->model : NULL;
		j->footprint = starpu_task_footprint(m, j->task, NULL, (unsigned)j->nimpl);
		j->footprint_valid = 1;
	}
	return j->footprint;
}

/* ------------------------------------------------- predict / update -- */

double starpu_perfmodel_history_based_expected_perf(struct starpu_perfmodel *m,
                                                    struct starpu_perfmodel_arch *arch, uint32_t fp)
{
	SP_LOCK();
	struct sp_mstate *s = mstate(m);
	struct sp_parch *a = arch_of(s, arch, 0);
	double r = NAN;
	if (a) {
		struct sp_hentry *e = find_entry(&a->impl[0], fp, 0);
		if (e && e->nsample >= (unsigned)sp.calibrate_minimum)
			r = e->mean;
	}
	SP_UNLOCK();
	return r;
}

double sp_perfmodel_predict(struct starpu_perfmodel *m, struct starpu_task *t, struct sp_job *j,
                            struct starpu_perfmodel_arch *arch, int nimpl)
{
	if (!m)
		return NAN;
	switch (m->type) {
	case STARPU_COMMON:
		return m->cost_function ? m->cost_function(t, (unsigned)nimpl) : NAN;
	case STARPU_PER_ARCH:
		return m->arch_cost_function ? m->arch_cost_function(t, arch, (unsigned)nimpl) : NAN;
	case STARPU_PER_WORKER: {
		if (!m->worker_cost_function)
			return NAN;
		for (int i = 0; i < sp.nworkers; i++)
			if (&sp.workers[i].perf_arch == arch)
				return m->worker_cost_function(t, (unsigned)i, (unsigned)nimpl);
		return NAN;
	}
	case STARPU_HISTORY_BASED:
		break;
	default:
		return NAN;
	}
	sp_perfmodel_register(m);
	struct sp_mstate *s = mstate(m);
	uint32_t fp = j ? sp_job_footprint(j) : starpu_task_footprint(m, t, arch, (unsigned)nimpl);
	struct sp_parch *a = arch_of(s, arch, 0);
	double r = NAN;
	if (a) {
		struct sp_hentry *e = find_entry(&a->impl[nimpl], fp, 0);
		if (e && e->nsample >= (unsigned)sp.calibrate_minimum)
			r = e->mean;
	}
	if (isnan(r))
		m->benchmarking = 1;
	return r;
}

static void update_entry(struct starpu_perfmodel *m, struct sp_parch *a, int nimpl, uint32_t fp,
                         size_t size, double flops, double measured, unsigned number)
{
	struct sp_mstate *s = mstate(m);
	a->impl_set[nimpl] = 1;
	struct sp_hentry *e = find_entry(&a->impl[nimpl], fp, 0);
	if (!e) {
		e = find_entry(&a->impl[nimpl], fp, 1);
		e->size = size;
		e->flops = flops;
		/* the first measurement of a footprint is usually bogus */
		if (number != 1 || m->type != STARPU_HISTORY_BASED) {
			e->sum = measured * number;
			e->sum2 = measured * measured * number;
			e->nsample = number;
			e->mean = measured;
		}
	} else {
		const double dev = measured / e->mean;
		const double maxerr = history_max_error();
		if (e->nsample && (100 * dev > 100 + maxerr || 100 / dev > 100 + maxerr)) {
			e->nerror += number;
			if (e->nerror >= e->nsample) {
				e->sum = e->sum2 = e->mean = e->dev = 0.0;
				e->nsample = e->nerror = 0;
			}
		} else {
			e->sum += measured * number;
			e->sum2 += measured * measured * number;
			e->nsample += number;
			const double n = e->nsample;
			e->mean = e->sum / n;
			e->dev = sqrt(fabs(e->sum2 - e->sum * e->sum / n) / n);
		}
		if (flops != 0.0 && !isnan(e->flops)) {
			if (e->flops == 0.0)
				e->flops = flops;
			else if (fabs(e->flops - flops) / e->flops > 1e-5)
				e->flops = NAN;
		}
	}
	s->dirty = 1;
}

void sp_perfmodel_feedback(struct sp_job *j, struct sp_worker *w, double measured_us)
{
	struct starpu_task *t = j->task;
	struct starpu_perfmodel *m = t->cl ? t->cl->model : NULL;
	if (!m || m->type != STARPU_HISTORY_BASED || !(measured_us >= 0))
		return;
	sp_perfmodel_register(m);
	if (!(sp.calibrate > 0 || m->benchmarking))
		return;
	struct sp_parch *a = arch_of(mstate(m), &w->perf_arch, 1);
	update_entry(m, a, j->nimpl, sp_job_footprint(j), starpu_task_data_footprint_size(t), t->flops,
	             measured_us, 1);
}

void starpu_perfmodel_update_history_n(struct starpu_perfmodel *m, struct starpu_task *t,
                                       struct starpu_perfmodel_arch *arch, unsigned cpuid,
                                       unsigned nimpl, double average, unsigned count)
{
	(void)cpuid;
	SP_LOCK();
	sp_perfmodel_register(m);
	struct sp_parch *a = arch_of(mstate(m), arch, 1);
	uint32_t fp = starpu_task_footprint(m, t, arch, nimpl);
	update_entry(m, a, (int)nimpl, fp, starpu_task_data_footprint_size(t), t->flops, average, count);
	SP_UNLOCK();
}

void starpu_perfmodel_update_history(struct starpu_perfmodel *m, struct starpu_task *t,
                                     struct starpu_perfmodel_arch *arch, unsigned cpuid,
                                     unsigned nimpl, double measured)
{
	starpu_perfmodel_update_history_n(m, t, arch, cpuid, nimpl, measured, 1);
}

int starpu_perfmodel_history_count(struct starpu_perfmodel *m, struct starpu_perfmodel_arch *arch)
{
	SP_LOCK();
	struct sp_parch *a = m->state ? arch_of(m->state, arch, 0) : NULL;
	int n = a ? a->impl[0].n : 0;
	SP_UNLOCK();
	return n;
}

int starpu_perfmodel_history_entry(struct starpu_perfmodel *m, struct starpu_perfmodel_arch *arch, int idx,
                                   uint32_t *footprint, size_t *size, double *flops, double *mean,
                                   double *dev, unsigned *nsample)
{
	SP_LOCK();
	struct sp_parch *a = m->state ? arch_of(m->state, arch, 0) : NULL;
	if (!a || idx < 0 || idx >= a->impl[0].n) {
		SP_UNLOCK();
		return -EINVAL;
	}
	struct sp_hentry *e = &a->impl[0].e[idx];
	if (footprint) *footprint = e->footprint;
	if (size) *size = e->size;
	if (flops) *flops = e->flops;
	if (mean) *mean = e->mean;
	if (dev) *dev = e->dev;
	if (nsample) *nsample = e->nsample;
	SP_UNLOCK();
	return 0;
}

/* ------------------------------------------------------- bus model --
 * The bandwidth (bytes per microsecond, i.e. MB/s) and latency
 * (microseconds) between memory nodes, kept in
 * $STARPU_PERF_MODEL_DIR/bus/<host>.bandwidth and <host>.latency in the
 * reference's layout (src/core/perfmodel/perfmodel_bus.c:789-797 paths,
 * :1194-1300 and :1481-1590 writers, :1340-1420 reader): a comment line
 * naming the destination nodes, then STARPU_MAXNODES rows of
 * STARPU_MAXNODES tab-separated values ("%e"), 0 on the diagonal and NaN
 * for nodes that do not exist. */

void sp_bus_path(const char *type, char *path, size_t maxlen)
{
	char dir[768], host[256];
	model_dir(dir, sizeof(dir));
	host_name(host, sizeof(host));
	snprintf(path, maxlen, "%s/bus/%s.%s", dir, host, type);
}

static const char *node_prefix(int n) { return sp.nodes[n].kind == SP_NODE_CPU ? "NUMA" : "HIP"; }

static int node_index(int n)
{
	int k = 0;
	for (int i = 0; i < n; i++)
		k += sp.nodes[i].kind == sp.nodes[n].kind;
	return k;
}

static int bus_write(const char *type, double (*m)[SP_MAXNODES])
{
	char path[1024];
	sp_bus_path(type, path, sizeof(path));
	mkdir_p(path);
	FILE *f = fopen(path, "w");
	if (!f)
		return -errno;
	fprintf(f, "# ");
	for (int d = 0; d < sp.nnodes; d++)
		fprintf(f, "to %s %d\t", node_prefix(d), node_index(d));
	fprintf(f, "\n");
	for (int a = 0; a < SP_MAXNODES; a++) {
		for (int b = 0; b < SP_MAXNODES; b++) {
			const double v = (a >= sp.nnodes || b >= sp.nnodes) ? NAN : a == b ? 0.0 : m[a][b];
			if (b)
				fputc('\t', f);
			fprintf(f, "%e", v);
		}
		fprintf(f, "\n");
	}
	return fclose(f) == 0 ? 0 : -EIO;
}

/* Reads a matrix file; 1 if it holds a finite positive value for every
 * ordered pair of distinct existing nodes and NaN for the others. */
static int bus_read(const char *type, double (*m)[SP_MAXNODES])
{
	char path[1024], line[4096];
	sp_bus_path(type, path, sizeof(path));
	FILE *f = fopen(path, "r");
	if (!f)
		return 0;
	int row = 0, ok = 1;
	while (ok && row < SP_MAXNODES && fgets(line, sizeof(line), f)) {
		if (line[0] == '#' || line[0] == '\n')
			continue;
		char *p = line;
		for (int col = 0; col < SP_MAXNODES; col++) {
			char *e;
			const double v = strtod(p, &e);
			if (e == p) {
				ok = 0;
				break;
			}
			p = e;
			const int exists = row < sp.nnodes && col < sp.nnodes;
			if (exists && row != col && !(isfinite(v) && v > 0.0))
				ok = 0;
			if (!exists && !isnan(v))
				ok = 0;
			if (exists)
				m[row][col] = v;
		}
		row++;
	}
	fclose(f);
	return ok && row == SP_MAXNODES;
}

int sp_bus_save(void)
{
	int e = bus_write("bandwidth", sp.bw);
	return e ? e : bus_write("latency", sp.lat);
}

int sp_bus_load(void)
{
	double bw[SP_MAXNODES][SP_MAXNODES], lat[SP_MAXNODES][SP_MAXNODES];
	if (!bus_read("bandwidth", bw) || !bus_read("latency", lat))
		return 0;
	for (int a = 0; a < sp.nnodes; a++)
		for (int b = 0; b < sp.nnodes; b++)
			if (a != b) {
				sp.bw[a][b] = bw[a][b];
				sp.lat[a][b] = lat[a][b];
			}
	return 1;
}
