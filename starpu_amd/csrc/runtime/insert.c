/*
 * starpu_task_insert / starpu_task_build: build a task from a 0-terminated
 * variadic list of (code, value) pairs (reference
 * src/util/starpu_task_insert.c:188, starpu_task_insert_utils.c:324).
 * Data access modes (STARPU_R, STARPU_W, STARPU_RW, STARPU_SCRATCH, ...)
 * are followed by a handle; the other codes by their value.
 * STARPU_VALUE arguments are packed into cl_arg as
 *   int nargs, then per argument: size_t size, bytes
 * which starpu_codelet_unpack_args reads back.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "sp_internal.h"

struct packer {
	char *buf;
	size_t len, cap;
	int nargs;
};

static void pack_bytes(struct packer *p, const void *src, size_t n)
{
	if (p->len + n > p->cap) {
		size_t cap = p->cap ? p->cap : 64;
		while (cap < p->len + n)
			cap *= 2;
		p->buf = realloc(p->buf, cap);
		p->cap = cap;
	}
	memcpy(p->buf + p->len, src, n);
	p->len += n;
}

static void pack_value(struct packer *p, const void *ptr, size_t size)
{
	if (!p->buf) {
		int zero = 0;
		pack_bytes(p, &zero, sizeof(zero));
	}
	pack_bytes(p, &size, sizeof(size));
	pack_bytes(p, ptr, size);
	p->nargs++;
	memcpy(p->buf, &p->nargs, sizeof(p->nargs));
}

static void add_buffer(struct starpu_task *t, int *n, starpu_data_handle_t h,
                       enum starpu_data_access_mode mode)
{
	if (*n < STARPU_NMAXBUFS && !t->dyn_handles) {
		t->handles[*n] = h;
		t->modes[*n] = mode;
	} else {
		if (!t->dyn_handles) {
			t->dyn_handles = malloc(sizeof(*t->dyn_handles) * 2 * STARPU_NMAXBUFS);
			t->dyn_modes = malloc(sizeof(*t->dyn_modes) * 2 * STARPU_NMAXBUFS);
			memcpy(t->dyn_handles, t->handles, sizeof(t->handles));
			memcpy(t->dyn_modes, t->modes, sizeof(t->modes));
		} else if ((*n & (*n - 1)) == 0 && *n >= 2 * STARPU_NMAXBUFS) {
			t->dyn_handles = realloc(t->dyn_handles, sizeof(*t->dyn_handles) * 2 * (size_t)*n);
			t->dyn_modes = realloc(t->dyn_modes, sizeof(*t->dyn_modes) * 2 * (size_t)*n);
		}
		t->dyn_handles[*n] = h;
		t->dyn_modes[*n] = mode;
	}
	(*n)++;
}

static int task_vset(struct starpu_task *t, struct starpu_codelet *cl, va_list ap)
{
	struct packer pk = {0};
	int nbuf = 0;
	t->cl = cl;
	int code;
	while ((code = va_arg(ap, int)) != 0) {
		if (code < (1 << STARPU_MODE_SHIFT)) {
			/* a data access mode, possibly with flags */
			starpu_data_handle_t h = va_arg(ap, starpu_data_handle_t);
			add_buffer(t, &nbuf, h, (enum starpu_data_access_mode)code);
			continue;
		}
		switch (code) {
		case STARPU_VALUE: {
			void *ptr = va_arg(ap, void *);
			size_t size = va_arg(ap, size_t);
			pack_value(&pk, ptr, size);
			break;
		}
		case STARPU_CL_ARGS:
			t->cl_arg = va_arg(ap, void *);
			t->cl_arg_size = va_arg(ap, size_t);
			t->cl_arg_free = 1;
			break;
		case STARPU_CL_ARGS_NFREE:
			t->cl_arg = va_arg(ap, void *);
			t->cl_arg_size = va_arg(ap, size_t);
			t->cl_arg_free = 0;
			break;
		case STARPU_CALLBACK:
			t->callback_func = va_arg(ap, void (*)(void *));
			break;
		case STARPU_CALLBACK_WITH_ARG:
		case STARPU_CALLBACK_WITH_ARG_NFREE:
			t->callback_func = va_arg(ap, void (*)(void *));
			t->callback_arg = va_arg(ap, void *);
			t->callback_arg_free = code == STARPU_CALLBACK_WITH_ARG;
			break;
		case STARPU_CALLBACK_ARG:
		case STARPU_CALLBACK_ARG_NFREE:
			t->callback_arg = va_arg(ap, void *);
			t->callback_arg_free = code == STARPU_CALLBACK_ARG;
			break;
		case STARPU_PROLOGUE_CALLBACK:
			t->prologue_callback_func = va_arg(ap, void (*)(void *));
			break;
		case STARPU_PROLOGUE_CALLBACK_ARG:
		case STARPU_PROLOGUE_CALLBACK_ARG_NFREE:
			t->prologue_callback_arg = va_arg(ap, void *);
			t->prologue_callback_arg_free = code == STARPU_PROLOGUE_CALLBACK_ARG;
			break;
		case STARPU_PROLOGUE_CALLBACK_POP:
			(void)va_arg(ap, void (*)(void *));
			break;
		case STARPU_PROLOGUE_CALLBACK_POP_ARG:
		case STARPU_PROLOGUE_CALLBACK_POP_ARG_NFREE:
			(void)va_arg(ap, void *);
			break;
		case STARPU_PRIORITY:
			t->priority = va_arg(ap, int);
			break;
		case STARPU_FLOPS:
			t->flops = va_arg(ap, double);
			break;
		case STARPU_NAME:
			t->name = va_arg(ap, const char *);
			break;
		case STARPU_TAG:
			t->tag_id = va_arg(ap, starpu_tag_t);
			t->use_tag = 1;
			break;
		case STARPU_TAG_ONLY:
			t->tag_id = va_arg(ap, starpu_tag_t);
			break;
		case STARPU_EXECUTE_ON_WORKER: {
			int w = va_arg(ap, int);
			if (w >= 0) {
				t->workerid = (unsigned)w;
				t->execute_on_a_specific_worker = 1;
			}
			break;
		}
		case STARPU_TASK_WORKERIDS:
			t->workerids_len = va_arg(ap, unsigned);
			t->workerids = va_arg(ap, uint32_t *);
			break;
		case STARPU_EXECUTE_WHERE:
			t->where = (int32_t)va_arg(ap, unsigned long long);
			break;
		case STARPU_EXECUTE_ON_NODE:
		case STARPU_WORKER_ORDER:
		case STARPU_SCHED_CTX:
		case STARPU_HYPERVISOR_TAG:
		case STARPU_POSSIBLY_PARALLEL:
		case STARPU_NODE_SELECTION_POLICY:
		case STARPU_TASK_COLOR:
		case STARPU_TRANSACTION:
			(void)va_arg(ap, int);
			break;
		case STARPU_EXECUTE_ON_DATA:
			(void)va_arg(ap, starpu_data_handle_t);
			break;
		case STARPU_SEQUENTIAL_CONSISTENCY:
			t->sequential_consistency = (unsigned)va_arg(ap, int);
			break;
		case STARPU_TASK_SYNCHRONOUS:
			t->synchronous = (unsigned)va_arg(ap, int);
			break;
		case STARPU_TASK_NO_SUBMITORDER:
			t->no_submitorder = (unsigned)va_arg(ap, int);
			break;
		case STARPU_HANDLES_SEQUENTIAL_CONSISTENCY:
			t->handles_sequential_consistency = va_arg(ap, unsigned char *);
			break;
		case STARPU_TASK_PROFILING_INFO:
			(void)va_arg(ap, void *);
			break;
		case STARPU_TASK_SCHED_DATA:
			t->sched_data = va_arg(ap, void *);
			break;
		case STARPU_DATA_ARRAY: {
			starpu_data_handle_t *arr = va_arg(ap, starpu_data_handle_t *);
			int n = va_arg(ap, int);
			for (int i = 0; i < n; i++) {
				enum starpu_data_access_mode m = STARPU_CODELET_GET_MODE(cl, nbuf);
				add_buffer(t, &nbuf, arr[i], m);
			}
			break;
		}
		case STARPU_DATA_MODE_ARRAY: {
			struct starpu_data_descr *arr = va_arg(ap, struct starpu_data_descr *);
			int n = va_arg(ap, int);
			for (int i = 0; i < n; i++)
				add_buffer(t, &nbuf, arr[i].handle, arr[i].mode);
			break;
		}
		case STARPU_TASK_DEPS_ARRAY: {
			unsigned n = va_arg(ap, unsigned);
			struct starpu_task **deps = va_arg(ap, struct starpu_task **);
			starpu_task_declare_deps_array(t, n, deps);
			break;
		}
		case STARPU_TASK_END_DEPS_ARRAY:
			(void)va_arg(ap, unsigned);
			(void)va_arg(ap, void *);
			break;
		case STARPU_TASK_END_DEP:
			(void)va_arg(ap, int);
			break;
		default:
			fprintf(stderr, "[starpu] starpu_task_insert: unknown argument code %d\n", code);
			free(pk.buf);
			return -EINVAL;
		}
	}
	t->nbuffers = nbuf;
	if (pk.buf) {
		t->cl_arg = pk.buf;
		t->cl_arg_size = pk.len;
		t->cl_arg_free = 1;
	}
	if (cl && cl->nbuffers != STARPU_VARIABLE_NBUFFERS && nbuf != cl->nbuffers) {
		fprintf(stderr, "[starpu] codelet %s expects %d buffers, %d given\n",
		        cl->name ? cl->name : "?", cl->nbuffers, nbuf);
		return -EINVAL;
	}
	/* the access modes given in the call take precedence over the
	 * codelet's static modes */
	if (!t->dyn_modes && cl && cl->nbuffers != STARPU_VARIABLE_NBUFFERS) {
		int differs = 0;
		for (int i = 0; i < nbuf; i++)
			if (t->modes[i] != STARPU_CODELET_GET_MODE(cl, i))
				differs = 1;
		if (differs) {
			t->dyn_modes = malloc(sizeof(*t->dyn_modes) * (size_t)(nbuf ? nbuf : 1));
			memcpy(t->dyn_modes, t->modes, sizeof(*t->dyn_modes) * (size_t)nbuf);
		}
	}
	return 0;
}

int starpu_task_set(struct starpu_task *t, struct starpu_codelet *cl, ...)
{
	va_list ap;
	va_start(ap, cl);
	int r = task_vset(t, cl, ap);
	va_end(ap);
	return r;
}

struct starpu_task *starpu_task_build(struct starpu_codelet *cl, ...)
{
	struct starpu_task *t = starpu_task_create();
	va_list ap;
	va_start(ap, cl);
	int r = task_vset(t, cl, ap);
	va_end(ap);
	if (r) {
		starpu_task_destroy(t);
		return NULL;
	}
	return t;
}

static int vinsert(struct starpu_codelet *cl, va_list ap)
{
	struct starpu_task *t = starpu_task_create();
	int r = task_vset(t, cl, ap);
	if (r) {
		starpu_task_destroy(t);
		return r;
	}
	int sync = t->synchronous;
	if (sync)
		t->detach = 0;
	r = starpu_task_submit(t);
	if (r) {
		starpu_task_destroy(t);
		return r;
	}
	return 0;
}

int starpu_task_insert(struct starpu_codelet *cl, ...)
{
	va_list ap;
	va_start(ap, cl);
	int r = vinsert(cl, ap);
	va_end(ap);
	return r;
}

int starpu_insert_task(struct starpu_codelet *cl, ...)
{
	va_list ap;
	va_start(ap, cl);
	int r = vinsert(cl, ap);
	va_end(ap);
	return r;
}

void starpu_codelet_unpack_args(void *cl_arg, ...)
{
	if (!cl_arg)
		return;
	const char *p = cl_arg;
	int nargs;
	memcpy(&nargs, p, sizeof(nargs));
	p += sizeof(nargs);
	va_list ap;
	va_start(ap, cl_arg);
	for (int i = 0; i < nargs; i++) {
		void *dst = va_arg(ap, void *);
		if (!dst)
			break;
		size_t size;
		memcpy(&size, p, sizeof(size));
		p += sizeof(size);
		memcpy(dst, p, size);
		p += size;
	}
	va_end(ap);
}

void starpu_codelet_pack_args(void **arg_buffer, size_t *arg_buffer_size, ...)
{
	struct packer pk = {0};
	va_list ap;
	va_start(ap, arg_buffer_size);
	int code;
	while ((code = va_arg(ap, int)) != 0) {
		STARPU_ASSERT_MSG(code == STARPU_VALUE, "only STARPU_VALUE can be packed");
		void *ptr = va_arg(ap, void *);
		size_t size = va_arg(ap, size_t);
		pack_value(&pk, ptr, size);
	}
	va_end(ap);
	*arg_buffer = pk.buf;
	*arg_buffer_size = pk.len;
}
