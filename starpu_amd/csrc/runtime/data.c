/*
 * Data handles, interfaces, filters and MSI coherency.
 *
 * Every handle keeps one replicate per memory node with an MSI state
 * (OWNER / SHARED / INVALID; reference src/datawizard/coherency.h:37-42,
 * state update coherency.c:172-240).  A task's buffers are fetched onto
 * the memory node of the worker that runs it: allocate if needed (views
 * into the parent's allocation for partitioned data, compact buffers
 * otherwise, as matrix_interface.c:475-499 allocates ld = nx), copy from a
 * valid replicate when the access reads, and for writes make the local
 * replicate the single OWNER.
 *
 * Implicit dependencies (sequential consistency, reference
 * src/core/dependencies/implicit_data_deps.c:215-434): a reader depends on
 * the last writer; a writer depends on the readers since the last writer,
 * or on the last writer when there are none.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "sp_internal.h"

void sp_data_init(void) {}
void sp_data_shutdown(void) {}

size_t sp_iface_size(const union sp_iface *f)
{
	switch (f->id) {
	case STARPU_MATRIX_INTERFACE_ID:
		return f->m.nx * f->m.ny * f->m.elemsize;
	case STARPU_VECTOR_INTERFACE_ID:
		return f->v.nx * f->v.elemsize;
	case STARPU_VARIABLE_INTERFACE_ID:
		return f->var.elemsize;
	default:
		return 0;
	}
}

static uintptr_t iface_ptr(const union sp_iface *f)
{
	switch (f->id) {
	case STARPU_MATRIX_INTERFACE_ID: return f->m.ptr;
	case STARPU_VECTOR_INTERFACE_ID: return f->v.ptr;
	case STARPU_VARIABLE_INTERFACE_ID: return f->var.ptr;
	default: return 0;
	}
}

/* Compact interface for a fresh allocation at ptr (matrix ld = nx). */
static void iface_compact(union sp_iface *dst, const union sp_iface *shape, uintptr_t ptr)
{
	*dst = *shape;
	switch (shape->id) {
	case STARPU_MATRIX_INTERFACE_ID:
		dst->m.ptr = ptr;
		dst->m.dev_handle = ptr;
		dst->m.offset = 0;
		dst->m.ld = shape->m.nx;
		dst->m.allocsize = shape->m.nx * shape->m.ny * shape->m.elemsize;
		break;
	case STARPU_VECTOR_INTERFACE_ID:
		dst->v.ptr = ptr;
		dst->v.dev_handle = ptr;
		dst->v.offset = 0;
		dst->v.allocsize = shape->v.nx * shape->v.elemsize;
		break;
	case STARPU_VARIABLE_INTERFACE_ID:
		dst->var.ptr = ptr;
		dst->var.dev_handle = ptr;
		dst->var.offset = 0;
		break;
	default:
		break;
	}
}

static starpu_data_handle_t handle_new(const union sp_iface *shape, int home)
{
	starpu_data_handle_t h = calloc(1, sizeof(*h));
	STARPU_ASSERT(h);
	h->iface_id = shape->id;
	h->home_node = home;
	h->shape = *shape;
	for (int n = 0; n < SP_MAXNODES; n++) {
		h->rep[n].handle = h;
		h->rep[n].node = n;
		h->rep[n].iface = *shape;
	}
	h->seq_consistency = sp.default_seq_consistency;
	h->size = sp_iface_size(shape);
	h->registered = 1;
	uint32_t fp;
	switch (shape->id) {
	case STARPU_MATRIX_INTERFACE_ID:
		fp = starpu_hash_crc32c_be((uint32_t)shape->m.nx, (uint32_t)shape->m.ny);
		break;
	case STARPU_VECTOR_INTERFACE_ID:
		fp = starpu_hash_crc32c_be((uint32_t)shape->v.nx, 0);
		break;
	default:
		fp = starpu_hash_crc32c_be((uint32_t)shape->var.elemsize, 0);
		break;
	}
	/* src/datawizard/footprint.c:73-83: fold the interface id in. */
	h->footprint = starpu_hash_crc32c_be(fp, (uint32_t)shape->id);
	return h;
}

static void register_home(starpu_data_handle_t h, int home, const union sp_iface *f)
{
	if (home < 0)
		return;
	STARPU_ASSERT_MSG(home < sp.nnodes || !sp.initialized, "invalid home node %d", home);
	struct sp_replicate *r = &h->rep[home];
	r->iface = *f;
	r->allocated = 1;
	r->automatic = 0;
	r->state = SP_OWNER;
	r->alloc_bytes = 0;
}

void starpu_matrix_data_register(starpu_data_handle_t *handle, int home_node, uintptr_t ptr,
                                 size_t ld, size_t nx, size_t ny, size_t elemsize)
{
	union sp_iface f;
	memset(&f, 0, sizeof(f));
	f.m.id = STARPU_MATRIX_INTERFACE_ID;
	f.m.nx = nx;
	f.m.ny = ny;
	f.m.ld = ld;
	f.m.elemsize = elemsize;
	f.m.allocsize = ld * ny * elemsize;
	union sp_iface shape = f;
	shape.m.ptr = 0;
	shape.m.ld = nx;
	SP_LOCK();
	starpu_data_handle_t h = handle_new(&shape, home_node);
	f.m.ptr = ptr;
	f.m.dev_handle = ptr;
	register_home(h, home_node, &f);
	SP_UNLOCK();
	*handle = h;
}

void starpu_vector_data_register(starpu_data_handle_t *handle, int home_node, uintptr_t ptr,
                                 size_t nx, size_t elemsize)
{
	union sp_iface f;
	memset(&f, 0, sizeof(f));
	f.v.id = STARPU_VECTOR_INTERFACE_ID;
	f.v.nx = nx;
	f.v.elemsize = elemsize;
	f.v.allocsize = nx * elemsize;
	union sp_iface shape = f;
	SP_LOCK();
	starpu_data_handle_t h = handle_new(&shape, home_node);
	f.v.ptr = ptr;
	f.v.dev_handle = ptr;
	register_home(h, home_node, &f);
	SP_UNLOCK();
	*handle = h;
}

void starpu_variable_data_register(starpu_data_handle_t *handle, int home_node, uintptr_t ptr,
                                   size_t size)
{
	union sp_iface f;
	memset(&f, 0, sizeof(f));
	f.var.id = STARPU_VARIABLE_INTERFACE_ID;
	f.var.elemsize = size;
	union sp_iface shape = f;
	SP_LOCK();
	starpu_data_handle_t h = handle_new(&shape, home_node);
	f.var.ptr = ptr;
	f.var.dev_handle = ptr;
	register_home(h, home_node, &f);
	SP_UNLOCK();
	*handle = h;
}

/* ---------------------------------------------------------- accessors -- */

static const union sp_iface *local_iface(starpu_data_handle_t h)
{
	int node = sp_current_worker ? sp_current_worker->node : 0;
	if (h->rep[node].allocated)
		return &h->rep[node].iface;
	if (h->home_node >= 0)
		return &h->rep[h->home_node].iface;
	return &h->shape;
}

size_t starpu_matrix_get_nx(starpu_data_handle_t h) { return h->shape.m.nx; }
size_t starpu_matrix_get_ny(starpu_data_handle_t h) { return h->shape.m.ny; }
size_t starpu_matrix_get_elemsize(starpu_data_handle_t h) { return h->shape.m.elemsize; }
size_t starpu_matrix_get_local_ld(starpu_data_handle_t h) { return local_iface(h)->m.ld; }
uintptr_t starpu_matrix_get_local_ptr(starpu_data_handle_t h) { return local_iface(h)->m.ptr; }
size_t starpu_vector_get_nx(starpu_data_handle_t h) { return h->shape.v.nx; }
uintptr_t starpu_vector_get_local_ptr(starpu_data_handle_t h) { return local_iface(h)->v.ptr; }
uintptr_t starpu_variable_get_local_ptr(starpu_data_handle_t h) { return local_iface(h)->var.ptr; }

size_t starpu_data_get_size(starpu_data_handle_t h) { return h->size; }
uint32_t starpu_data_get_footprint(starpu_data_handle_t h) { return h->footprint; }
void *starpu_data_get_local_ptr(starpu_data_handle_t h) { return (void *)iface_ptr(local_iface(h)); }
int starpu_data_get_home_node(starpu_data_handle_t h) { return h->home_node; }

void *starpu_data_handle_to_pointer(starpu_data_handle_t h, unsigned node)
{
	if (node >= SP_MAXNODES || !h->rep[node].allocated)
		return NULL;
	return (void *)iface_ptr(&h->rep[node].iface);
}

void *starpu_data_get_interface_on_node(starpu_data_handle_t h, unsigned node)
{
	if (node >= SP_MAXNODES)
		return NULL;
	return &h->rep[node].iface;
}

unsigned starpu_data_is_on_node(starpu_data_handle_t h, unsigned node)
{
	if (node >= SP_MAXNODES)
		return 0;
	SP_LOCK();
	unsigned r = h->rep[node].state != SP_INVALID;
	SP_UNLOCK();
	return r;
}

void starpu_data_set_name(starpu_data_handle_t h, const char *name)
{
	SP_LOCK();
	free(h->name);
	h->name = name ? strdup(name) : NULL;
	SP_UNLOCK();
}

const char *starpu_data_get_name(starpu_data_handle_t h) { return h->name; }

void starpu_data_set_sequential_consistency_flag(starpu_data_handle_t h, unsigned flag)
{
	SP_LOCK();
	h->seq_consistency = flag;
	SP_UNLOCK();
}

unsigned starpu_data_get_sequential_consistency_flag(starpu_data_handle_t h) { return h->seq_consistency; }

void starpu_data_set_default_sequential_consistency_flag(unsigned flag)
{
	sp.default_seq_consistency = flag;
}

int sp_rep_valid(struct sp_replicate *r) { return r->state != SP_INVALID; }

int sp_handle_valid_node(starpu_data_handle_t h)
{
	/* prefer the OWNER, then the home node, then any SHARED copy */
	for (int n = 0; n < sp.nnodes; n++)
		if (h->rep[n].state == SP_OWNER)
			return n;
	if (h->home_node >= 0 && h->rep[h->home_node].state != SP_INVALID)
		return h->home_node;
	for (int n = 0; n < sp.nnodes; n++)
		if (h->rep[n].state != SP_INVALID)
			return n;
	return -1;
}

/* ------------------------------------------------------------ filters -- */

void starpu_filter_nparts_compute_chunk_size_and_offset(unsigned n, unsigned nparts, size_t elemsize,
                                                        unsigned id, size_t blocksize,
                                                        size_t *chunk_size, size_t *offset)
{
	/* The first n % nparts chunks are one element larger
	 * (reference src/datawizard/filters.c:1655-1679). */
	size_t base = n / nparts, rem = n % nparts;
	*chunk_size = base + (id < rem ? 1 : 0);
	if (offset)
		*offset = ((size_t)id * base + (id < rem ? id : rem)) * blocksize * elemsize;
}

static void matrix_filter(int dim, void *parent, void *child, unsigned id, unsigned nparts)
{
	struct starpu_matrix_interface *p = parent, *c = child;
	size_t n = dim == 1 ? p->nx : p->ny;
	STARPU_ASSERT_MSG(nparts <= n, "cannot split %zu elements in %u parts", n, nparts);
	size_t chunk, off;
	starpu_filter_nparts_compute_chunk_size_and_offset((unsigned)n, nparts, p->elemsize, id,
	                                                   dim == 1 ? 1 : p->ld, &chunk, &off);
	*c = *p;
	if (dim == 1)
		c->nx = chunk;
	else
		c->ny = chunk;
	if (p->ptr) {
		c->ptr = p->ptr + off;
		c->dev_handle = p->dev_handle;
		c->offset = p->offset + off;
		c->ld = p->ld;
		c->allocsize = c->ld * c->ny * c->elemsize;
	} else {
		c->ld = c->nx;
		c->allocsize = c->nx * c->ny * c->elemsize;
	}
}

void starpu_matrix_filter_block(void *parent, void *child, struct starpu_data_filter *f,
                                unsigned id, unsigned nparts)
{
	(void)f;
	matrix_filter(1, parent, child, id, nparts);
}

void starpu_matrix_filter_vertical_block(void *parent, void *child, struct starpu_data_filter *f,
                                         unsigned id, unsigned nparts)
{
	(void)f;
	matrix_filter(2, parent, child, id, nparts);
}

void starpu_vector_filter_block(void *parent, void *child, struct starpu_data_filter *f,
                                unsigned id, unsigned nparts)
{
	(void)f;
	struct starpu_vector_interface *p = parent, *c = child;
	size_t chunk, off;
	starpu_filter_nparts_compute_chunk_size_and_offset((unsigned)p->nx, nparts, p->elemsize, id, 1,
	                                                   &chunk, &off);
	*c = *p;
	c->nx = chunk;
	c->allocsize = chunk * p->elemsize;
	if (p->ptr) {
		c->ptr = p->ptr + off;
		c->offset = p->offset + off;
	}
}

static void wait_pending(starpu_data_handle_t h)
{
	while (h->pending > 0)
		pthread_cond_wait(&sp.done_cond, &sp.lock);
}

static void child_view(starpu_data_handle_t c, int n)
{
	starpu_data_handle_t p = c->parent;
	union sp_iface v;
	memset(&v, 0, sizeof(v));
	c->filter.filter_func(&p->rep[n].iface, &v, &c->filter, c->child_id, c->nparts);
	struct sp_replicate *r = &c->rep[n];
	r->iface = v;
	r->allocated = 1;
	r->automatic = 0;
	r->is_view = 1;
	r->alloc_bytes = 0;
}

void starpu_data_partition(starpu_data_handle_t h, struct starpu_data_filter *f)
{
	SP_LOCK();
	STARPU_ASSERT_MSG(!h->partitioned, "handle is already partitioned");
	wait_pending(h);
	unsigned nparts = f->get_nchildren ? f->get_nchildren(f, h) : f->nchildren;
	STARPU_ASSERT(nparts > 0);
	h->children = calloc(nparts, sizeof(*h->children));
	h->nchildren = (int)nparts;
	for (unsigned i = 0; i < nparts; i++) {
		union sp_iface shape;
		memset(&shape, 0, sizeof(shape));
		f->filter_func(&h->shape, &shape, f, i, nparts);
		starpu_data_handle_t c = handle_new(&shape, h->home_node);
		/* the home replicate is re-derived as a view below */
		c->filter = *f;
		c->child_id = i;
		c->nparts = nparts;
		c->parent = h;
		c->seq_consistency = h->seq_consistency;
		for (int n = 0; n < sp.nnodes || n == 0; n++) {
			if (!h->rep[n].allocated)
				continue;
			child_view(c, n);
			c->rep[n].state = h->rep[n].state;
		}
		h->children[i] = c;
	}
	h->partitioned = 1;
	SP_UNLOCK();
}

starpu_data_handle_t starpu_data_get_child(starpu_data_handle_t h, unsigned i)
{
	STARPU_ASSERT_MSG(h->partitioned && (int)i < h->nchildren, "no child %u", i);
	return h->children[i];
}

int starpu_data_get_nb_children(starpu_data_handle_t h) { return h->partitioned ? h->nchildren : 0; }

starpu_data_handle_t starpu_data_vget_sub_data(starpu_data_handle_t root, unsigned depth, va_list pa)
{
	starpu_data_handle_t h = root;
	for (unsigned d = 0; d < depth; d++) {
		unsigned i = va_arg(pa, unsigned);
		h = starpu_data_get_child(h, i);
	}
	return h;
}

starpu_data_handle_t starpu_data_get_sub_data(starpu_data_handle_t root, unsigned depth, ...)
{
	va_list pa;
	va_start(pa, depth);
	starpu_data_handle_t h = starpu_data_vget_sub_data(root, depth, pa);
	va_end(pa);
	return h;
}

static void map_filters_rec(starpu_data_handle_t h, unsigned nfilters, struct starpu_data_filter **fs)
{
	if (nfilters == 0)
		return;
	starpu_data_partition(h, fs[0]);
	for (int i = 0; i < h->nchildren; i++)
		map_filters_rec(h->children[i], nfilters - 1, fs + 1);
}

void starpu_data_vmap_filters(starpu_data_handle_t root, unsigned nfilters, va_list pa)
{
	struct starpu_data_filter *fs[16];
	STARPU_ASSERT(nfilters <= 16);
	for (unsigned i = 0; i < nfilters; i++)
		fs[i] = va_arg(pa, struct starpu_data_filter *);
	map_filters_rec(root, nfilters, fs);
}

void starpu_data_map_filters(starpu_data_handle_t root, unsigned nfilters, ...)
{
	va_list pa;
	va_start(pa, nfilters);
	starpu_data_vmap_filters(root, nfilters, pa);
	va_end(pa);
}

static void free_replicates(starpu_data_handle_t h)
{
	for (int n = 0; n < SP_MAXNODES; n++) {
		struct sp_replicate *r = &h->rep[n];
		if (r->in_lru) {
			struct sp_node *nd = &sp.nodes[n];
			if (r->lru_prev) r->lru_prev->lru_next = r->lru_next; else nd->lru_head = r->lru_next;
			if (r->lru_next) r->lru_next->lru_prev = r->lru_prev; else nd->lru_tail = r->lru_prev;
			r->in_lru = 0;
			r->lru_prev = r->lru_next = NULL;
		}
		if (r->ready_event) {
			sp_hip_event_destroy(r->ready_event);
			r->ready_event = NULL;
		}
		if (r->allocated && r->automatic && !r->is_view)
			sp_node_free(n, (void *)iface_ptr(&r->iface), r->alloc_bytes);
		r->allocated = 0;
		r->state = SP_INVALID;
	}
	if (h->scratch) {
		for (int w = 0; w < sp.nworkers; w++) {
			struct sp_replicate *r = &h->scratch[w];
			if (r->allocated && r->automatic)
				sp_node_free(r->node, (void *)iface_ptr(&r->iface), r->alloc_bytes);
		}
		free(h->scratch);
		h->scratch = NULL;
	}
}

static void handle_free(starpu_data_handle_t h)
{
	free(h->readers);
	free(h->children);
	free(h->name);
	h->registered = 0;
	free(h);
}

static void free_children(starpu_data_handle_t h)
{
	for (int i = 0; i < h->nchildren; i++) {
		starpu_data_handle_t c = h->children[i];
		if (c->partitioned)
			free_children(c);
		free_replicates(c);
		handle_free(c);
	}
	free(h->children);
	h->children = NULL;
	h->nchildren = 0;
	h->partitioned = 0;
}

/* Make the data of child c valid inside its region of the parent's
 * replicate on `node` (the parent's replicate there must be allocated). */
static void gather_child(starpu_data_handle_t c, int node)
{
	starpu_data_handle_t p = c->parent;
	union sp_iface region;
	memset(&region, 0, sizeof(region));
	c->filter.filter_func(&p->rep[node].iface, &region, &c->filter, c->child_id, c->nparts);
	if (c->partitioned) {
		/* c holds no data of its own: point it at its region and gather
		 * the grandchildren there */
		child_view(c, node);
		for (int i = 0; i < c->nchildren; i++)
			gather_child(c->children[i], node);
		return;
	}
	struct sp_replicate *cr = &c->rep[node];
	if (cr->allocated && cr->is_view && cr->state != SP_INVALID)
		return;     /* already valid in place */
	int s = sp_handle_valid_node(c);
	if (s < 0)
		return;     /* never written: nothing to gather */
	sp_copy_iface(&region, node, &c->rep[s].iface, s, NULL);
}

static int any_pending(starpu_data_handle_t h)
{
	for (int i = 0; i < h->nchildren; i++) {
		starpu_data_handle_t c = h->children[i];
		if (c->pending > 0 || (c->partitioned && any_pending(c)))
			return 1;
	}
	return 0;
}

void starpu_data_unpartition(starpu_data_handle_t h, unsigned node)
{
	SP_LOCK();
	STARPU_ASSERT_MSG(h->partitioned, "handle is not partitioned");
	while (any_pending(h))
		pthread_cond_wait(&sp.done_cond, &sp.lock);
	struct sp_replicate *pr = &h->rep[node];
	if (!pr->allocated) {
		size_t bytes = sp_iface_size(&h->shape);
		void *ptr = sp_node_alloc((int)node, bytes ? bytes : 1);
		STARPU_ASSERT_MSG(ptr, "cannot allocate %zu bytes on node %u to unpartition", bytes, node);
		iface_compact(&pr->iface, &h->shape, (uintptr_t)ptr);
		pr->allocated = 1;
		pr->automatic = 1;
		pr->is_view = 0;
		pr->alloc_bytes = bytes ? bytes : 1;
	}
	for (int i = 0; i < h->nchildren; i++)
		gather_child(h->children[i], (int)node);
	free_children(h);
	for (int n = 0; n < SP_MAXNODES; n++)
		if (n != (int)node)
			h->rep[n].state = SP_INVALID;
	pr->state = SP_OWNER;
	SP_UNLOCK();
}

/* ------------------------------------------------------------- fetch -- */

static int alloc_replicate(starpu_data_handle_t h, struct sp_replicate *r, int node)
{
	if (h->parent && h->parent->rep[node].allocated) {
		child_view(h, node);
		return 0;
	}
	size_t bytes = sp_iface_size(&h->shape);
	void *ptr = sp_node_alloc(node, bytes ? bytes : 1);
	if (!ptr)
		return -ENOMEM;
	iface_compact(&r->iface, &h->shape, (uintptr_t)ptr);
	r->allocated = 1;
	r->automatic = 1;
	r->is_view = 0;
	r->alloc_bytes = bytes ? bytes : 1;
	return 0;
}

void sp_lru_touch(struct sp_replicate *r)
{
	struct sp_node *nd = &sp.nodes[r->node];
	r->last_used = sp_now();
	if (r->is_view || !r->automatic || nd->kind != SP_NODE_HIP)
		return;
	if (r->in_lru) {
		if (r->lru_prev) r->lru_prev->lru_next = r->lru_next; else nd->lru_head = r->lru_next;
		if (r->lru_next) r->lru_next->lru_prev = r->lru_prev; else nd->lru_tail = r->lru_prev;
	}
	/* append at the tail (most recently used); wont_use data goes first */
	r->lru_next = NULL;
	r->lru_prev = NULL;
	if (r->wont_use && nd->lru_head) {
		r->lru_next = nd->lru_head;
		nd->lru_head->lru_prev = r;
		nd->lru_head = r;
	} else {
		r->lru_prev = nd->lru_tail;
		if (nd->lru_tail) nd->lru_tail->lru_next = r; else nd->lru_head = r;
		nd->lru_tail = r;
	}
	if (!nd->lru_head)
		nd->lru_head = r;
	r->in_lru = 1;
}

int sp_fetch_buffer(struct sp_job *j, int i, struct sp_worker *w, void *stream)
{
	struct sp_buffer *b = &j->buf[i];
	starpu_data_handle_t h = b->handle;
	const int node = w ? w->node : b->node;
	b->node = node;
	if (b->mode & STARPU_SCRATCH) {
		STARPU_ASSERT(w);
		if (!h->scratch)
			h->scratch = calloc((size_t)sp.nworkers, sizeof(*h->scratch));
		struct sp_replicate *r = &h->scratch[w->id];
		r->handle = h;
		r->node = node;
		if (!r->allocated) {
			size_t bytes = sp_iface_size(&h->shape);
			void *ptr = sp_node_alloc(node, bytes ? bytes : 1);
			if (!ptr)
				return -ENOMEM;
			iface_compact(&r->iface, &h->shape, (uintptr_t)ptr);
			r->allocated = 1;
			r->automatic = 1;
			r->alloc_bytes = bytes ? bytes : 1;
		}
		b->rep = r;
		j->ifaces[i] = &r->iface;
		return 0;
	}
	STARPU_ASSERT_MSG(!h->partitioned, "task accesses a partitioned handle");
	struct sp_replicate *r = &h->rep[node];
	if (!r->allocated) {
		int e = alloc_replicate(h, r, node);
		if (e)
			return e;
	}
	r->refcnt++;
	if ((b->mode & STARPU_R) && r->state == SP_INVALID) {
		int s = sp_handle_valid_node(h);
		if (s >= 0) {
			struct sp_replicate *src = &h->rep[s];
			if (src->ready_event && stream && src->ready_stream != stream)
				sp_hip_stream_wait_event(stream, src->ready_event);
			int e = sp_copy_iface(&r->iface, node, &src->iface, s, stream);
			if (e) {
				r->refcnt--;
				return e;
			}
			j->bytes_in += h->size;
			if (src->state == SP_OWNER)
				src->state = SP_SHARED;
			sp_lru_touch(src);
			if (stream && sp.nodes[node].kind == SP_NODE_HIP) {
				if (r->ready_event)
					sp_hip_event_destroy(r->ready_event);
				r->ready_event = sp_hip_event_record(stream);
				r->ready_stream = stream;
			}
		}
		r->state = SP_SHARED;
	} else if (r->ready_event && r->ready_stream != stream) {
		/* filled asynchronously on another stream: order after that copy */
		if (stream)
			sp_hip_stream_wait_event(stream, r->ready_event);
		else
			sp_hip_stream_sync(r->ready_stream);
	}
	if (b->mode & STARPU_W) {
		for (int n = 0; n < SP_MAXNODES; n++)
			if (n != node)
				h->rep[n].state = SP_INVALID;
		r->state = SP_OWNER;
	} else if (r->state == SP_INVALID) {
		r->state = SP_SHARED;   /* read of never-written data */
	}
	r->wont_use = 0;
	sp_lru_touch(r);
	b->rep = r;
	j->ifaces[i] = &r->iface;
	return 0;
}

/* Prefetch (lock held) at the push of a ready job to worker w: its read
 * inputs whose only valid copies are in main memory start moving to w's
 * device on the device's transfer stream, so the copies overlap the
 * kernels already queued on the worker streams (the reference's dmda
 * push calls starpu_prefetch_task_input_for,
 * deque_modeling_policy_data_aware.c:354-355).  A ready job's inputs are
 * final for it: the writers it depends on have terminated and later
 * writers wait for it.  The replicate becomes valid with a ready event
 * that the worker's stream, or any other reader's, waits for before use;
 * nothing is pinned, so it stays evictable once the copy has completed.
 * Only copies that fit the node's budget without evicting are started. */
void sp_prefetch_job(struct sp_job *j, struct sp_worker *w)
{
	if (!sp.prefetch || w->arch != STARPU_HIP_WORKER)
		return;
	void *xs = sp_hip_transfer_stream(w->devid);
	if (!xs)
		return;
	const int node = w->node;
	for (int i = 0; i < j->nbuffers; i++) {
		const struct sp_buffer *b = &j->buf[i];
		if (!(b->mode & STARPU_R) || (b->mode & STARPU_SCRATCH))
			continue;
		starpu_data_handle_t h = b->handle;
		if (h->partitioned || (h->parent && h->parent->rep[node].allocated))
			continue;
		struct sp_replicate *r = &h->rep[node];
		if (r->state != SP_INVALID)
			continue;
		const int s = sp_handle_valid_node(h);
		if (s < 0 || sp.nodes[s].kind != SP_NODE_CPU)
			continue;
		if (!r->allocated) {
			const size_t bytes = sp_iface_size(&h->shape);
			void *ptr = sp_node_try_alloc(node, bytes ? bytes : 1);
			if (!ptr)
				continue;
			iface_compact(&r->iface, &h->shape, (uintptr_t)ptr);
			r->allocated = 1;
			r->automatic = 1;
			r->is_view = 0;
			r->alloc_bytes = bytes ? bytes : 1;
		}
		struct sp_replicate *src = &h->rep[s];
		if (sp_copy_iface(&r->iface, node, &src->iface, s, xs) != 0)
			continue;
		if (r->ready_event)
			sp_hip_event_destroy(r->ready_event);
		r->ready_event = sp_hip_event_record(xs);
		r->ready_stream = xs;
		if (!r->ready_event)
			sp_hip_stream_sync(xs);   /* no event to wait on: finish the copy now */
		if (src->state == SP_OWNER)
			src->state = SP_SHARED;
		r->state = SP_SHARED;
		sp_lru_touch(r);
		sp.nodes[node].prefetches++;
	}
}

void sp_release_buffer(struct sp_job *j, int i)
{
	struct sp_buffer *b = &j->buf[i];
	if (!b->rep)
		return;
	if (!(b->mode & STARPU_SCRATCH)) {
		b->rep->refcnt--;
		STARPU_ASSERT(b->rep->refcnt >= 0);
	}
	b->rep = NULL;
}

/* ---------------------------------------------------- implicit deps -- */

static void readers_add(starpu_data_handle_t h, struct sp_job *j)
{
	for (int i = 0; i < h->nreaders; i++)
		if (h->readers[i] == j)
			return;
	if (h->nreaders == h->readers_cap) {
		h->readers_cap = h->readers_cap ? 2 * h->readers_cap : 8;
		h->readers = realloc(h->readers, (size_t)h->readers_cap * sizeof(*h->readers));
	}
	h->readers[h->nreaders++] = j;
}

/* Access mode of handle h over all the buffers of j (a handle may appear
 * more than once); returns -1 if this is not the first buffer using h. */
static int merged_mode(struct sp_job *j, int i)
{
	starpu_data_handle_t h = j->buf[i].handle;
	int mode = 0;
	for (int k = 0; k < j->nbuffers; k++) {
		if (j->buf[k].handle != h)
			continue;
		if (k < i)
			return -1;
		mode |= j->buf[k].mode;
	}
	return mode;
}

void sp_detect_implicit_deps(struct sp_job *j)
{
	struct starpu_task *t = j->task;
	for (int i = 0; i < j->nbuffers; i++) {
		int mode = merged_mode(j, i);
		if (mode < 0)
			continue;
		starpu_data_handle_t h = j->buf[i].handle;
		h->pending++;
		if (mode & STARPU_SCRATCH)
			continue;
		int seq = h->seq_consistency && (t == NULL || t->sequential_consistency);
		if (t && t->handles_sequential_consistency)
			seq = seq && t->handles_sequential_consistency[i];
		if (!seq)
			continue;
		if (!(mode & STARPU_W)) {
			if (h->last_writer)
				sp_job_add_dependency(h->last_writer, j);
			readers_add(h, j);
		} else {
			if (h->nreaders > 0) {
				for (int r = 0; r < h->nreaders; r++)
					if (h->readers[r] != j)
						sp_job_add_dependency(h->readers[r], j);
				h->nreaders = 0;
			} else if (h->last_writer && h->last_writer != j) {
				sp_job_add_dependency(h->last_writer, j);
			}
			h->last_writer = j;
		}
	}
}

void sp_remove_implicit_deps(struct sp_job *j)
{
	for (int i = 0; i < j->nbuffers; i++) {
		if (merged_mode(j, i) < 0)
			continue;
		starpu_data_handle_t h = j->buf[i].handle;
		if (h->last_writer == j)
			h->last_writer = NULL;
		for (int r = 0; r < h->nreaders; r++)
			if (h->readers[r] == j) {
				h->readers[r] = h->readers[--h->nreaders];
				break;
			}
		sp_handle_pending_dec(h);
	}
}

void sp_handle_pending_dec(starpu_data_handle_t h)
{
	h->pending--;
	STARPU_ASSERT(h->pending >= 0);
	if (h->pending == 0)
		pthread_cond_broadcast(&sp.done_cond);
}

/* --------------------------------------------------- handle lifetime -- */

void starpu_data_wont_use(starpu_data_handle_t h)
{
	SP_LOCK();
	for (int n = 1; n < sp.nnodes; n++) {
		struct sp_replicate *r = &h->rep[n];
		if (r->allocated && r->automatic && !r->is_view) {
			r->wont_use = 1;
			sp_lru_touch(r);
		}
	}
	SP_UNLOCK();
}

static void unregister_common(starpu_data_handle_t h, int coherent)
{
	STARPU_ASSERT_MSG(!h->partitioned, "unregistering a partitioned handle: unpartition it first");
	if (coherent && h->home_node >= 0)
		starpu_data_acquire_on_node(h, h->home_node, STARPU_R), starpu_data_release_on_node(h, h->home_node);
	SP_LOCK();
	wait_pending(h);
	free_replicates(h);
	handle_free(h);
	SP_UNLOCK();
}

void starpu_data_unregister(starpu_data_handle_t h) { unregister_common(h, 1); }
void starpu_data_unregister_no_coherency(starpu_data_handle_t h) { unregister_common(h, 0); }
void starpu_data_unregister_submit(starpu_data_handle_t h) { unregister_common(h, 1); }

void starpu_data_invalidate(starpu_data_handle_t h)
{
	SP_LOCK();
	wait_pending(h);
	for (int n = 0; n < SP_MAXNODES; n++)
		h->rep[n].state = SP_INVALID;
	SP_UNLOCK();
}

int starpu_data_prefetch_on_node(starpu_data_handle_t h, unsigned node, unsigned async)
{
	(void)async;
	/* Synchronous prefetch through an acquire of the replicate. */
	int r = starpu_data_acquire_on_node(h, (int)node, STARPU_R);
	if (r == 0)
		starpu_data_release_on_node(h, (int)node);
	return r;
}

int starpu_data_fetch_on_node(starpu_data_handle_t h, unsigned node, unsigned async)
{
	return starpu_data_prefetch_on_node(h, node, async);
}
