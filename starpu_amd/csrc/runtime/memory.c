/*
 * Memory nodes: allocation with a per-node budget, a cache of freed device
 * buffers, LRU eviction, and 2-D copies between nodes.
 *
 * Reference: allocation and reuse src/datawizard/memalloc.c:1573-1834,
 * reclaim/LRU :1142-1221, per-node budget src/datawizard/memory_manager.c
 * :45-235, STARPU_LIMIT_HIP_MEM src/drivers/hip/driver_hip.c:456-480,
 * copy2d src/datawizard/copy_driver.c:501.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include "sp_internal.h"

/* hipMemcpyKind values of the HIP runtime API */
enum { K_H2H = 0, K_H2D = 1, K_D2H = 2, K_D2D = 3 };

/* used_bytes of node went down (lock held): wake the memory waiters once
 * the smallest amount one of them waits for is available; those still
 * short register their amount again (memory_manager.c:129-140). */
static void mem_released(int node)
{
	struct sp_node *nd = &sp.nodes[node];
	if (nd->waiting_bytes && nd->total_bytes - nd->used_bytes >= nd->waiting_bytes) {
		nd->waiting_bytes = 0;
		pthread_cond_broadcast(&sp.mem_cond);
	}
}

int sp_memory_init(void)
{
	for (int n = 0; n < sp.nnodes; n++) {
		sp.nodes[n].used_bytes = 0;
		sp.nodes[n].waiting_bytes = 0;
		sp.nodes[n].lru_head = sp.nodes[n].lru_tail = NULL;
	}
	/* bus model defaults (bytes per microsecond, microseconds); the HIP
	 * driver overwrites host<->device entries with measured values */
	for (int a = 0; a < SP_MAXNODES; a++)
		for (int b = 0; b < SP_MAXNODES; b++) {
			sp.bw[a][b] = (a == b) ? 1e9 : 25e3;
			sp.lat[a][b] = (a == b) ? 0.0 : 10.0;
		}
	return 0;
}

static void *raw_alloc(int node, size_t bytes)
{
	struct sp_node *nd = &sp.nodes[node];
	if (nd->kind == SP_NODE_CPU) {
		void *p = NULL;
		if (posix_memalign(&p, 256, bytes))
			return NULL;
		return p;
	}
	/* reuse a cached buffer of the same size */
	struct sp_cached_buf **pp = &nd->cache;
	while (*pp) {
		if ((*pp)->bytes == bytes) {
			struct sp_cached_buf *c = *pp;
			void *p = c->ptr;
			*pp = c->next;
			nd->cached_bytes -= (long long)c->bytes;
			free(c);
			return p;
		}
		pp = &(*pp)->next;
	}
	void *p = NULL;
	if (sp_hip_malloc(nd->devid, &p, bytes) != 0)
		return NULL;
	return p;
}

void sp_node_release_cache(int node)
{
	struct sp_node *nd = &sp.nodes[node];
	while (nd->cache) {
		struct sp_cached_buf *c = nd->cache;
		nd->cache = c->next;
		sp_hip_free(nd->devid, c->ptr);
		nd->used_bytes -= (long long)c->bytes;
		free(c);
	}
	nd->cached_bytes = 0;
	mem_released(node);
}

/* Evict one unused device replicate (LRU order); 1 if memory was freed. */
static int evict_one(int node)
{
	struct sp_node *nd = &sp.nodes[node];
	for (struct sp_replicate *r = nd->lru_head; r; r = r->lru_next) {
		if (r->refcnt > 0 || !r->automatic || r->is_view || !r->allocated)
			continue;
		/* an asynchronous copy into it (a prefetch) may still be running */
		if (r->ready_event && !sp_hip_event_done(r->ready_event))
			continue;
		starpu_data_handle_t h = r->handle;
		if (h->partitioned)
			continue;     /* its memory backs the children's views */
		if (r->state == SP_OWNER) {
			/* only valid copy: write it back to the home node (or RAM) */
			int home = h->home_node >= 0 ? h->home_node : 0;
			if (home == node)
				continue;
			struct sp_replicate *hr = &h->rep[home];
			if (!hr->allocated) {
				size_t bytes = sp_iface_size(&h->shape);
				void *p = raw_alloc(home, bytes);
				if (!p)
					continue;
				hr->iface = h->shape;
				switch (h->shape.id) {
				case STARPU_MATRIX_INTERFACE_ID:
					hr->iface.m.ptr = (uintptr_t)p;
					hr->iface.m.ld = h->shape.m.nx;
					break;
				case STARPU_VECTOR_INTERFACE_ID:
					hr->iface.v.ptr = (uintptr_t)p;
					break;
				default:
					hr->iface.var.ptr = (uintptr_t)p;
					break;
				}
				hr->allocated = 1;
				hr->automatic = 1;
				hr->alloc_bytes = bytes;
				sp.nodes[home].used_bytes += (long long)bytes;
			}
			if (sp_copy_iface(&hr->iface, home, &r->iface, node, NULL) != 0)
				continue;
			hr->state = SP_OWNER;
		} else if (r->state == SP_SHARED) {
			/* another valid copy survives; if it is the last SHARED one
			 * elsewhere it becomes the owner */
			int others = 0, last = -1;
			for (int n = 0; n < SP_MAXNODES; n++)
				if (n != node && h->rep[n].state != SP_INVALID) {
					others++;
					last = n;
				}
			if (others == 0)
				continue;
			if (others == 1)
				h->rep[last].state = SP_OWNER;
		}
		/* unlink and free */
		if (r->lru_prev) r->lru_prev->lru_next = r->lru_next; else nd->lru_head = r->lru_next;
		if (r->lru_next) r->lru_next->lru_prev = r->lru_prev; else nd->lru_tail = r->lru_prev;
		r->in_lru = 0;
		r->lru_prev = r->lru_next = NULL;
		uintptr_t ptr = h->shape.id == STARPU_MATRIX_INTERFACE_ID ? r->iface.m.ptr
		              : h->shape.id == STARPU_VECTOR_INTERFACE_ID ? r->iface.v.ptr : r->iface.var.ptr;
		sp_hip_free(nd->devid, (void *)ptr);
		nd->used_bytes -= (long long)r->alloc_bytes;
		r->allocated = 0;
		r->state = SP_INVALID;
		nd->evictions++;
		mem_released(node);
		return 1;
	}
	return 0;
}

void *sp_node_alloc(int node, size_t bytes)
{
	struct sp_node *nd = &sp.nodes[node];
	if (nd->kind == SP_NODE_CPU) {
		void *p = raw_alloc(node, bytes);
		if (p)
			nd->used_bytes += (long long)bytes;
		return p;
	}
	for (;;) {
		int fits = nd->used_bytes + (long long)bytes <= nd->total_bytes;
		int cached = 0;
		for (struct sp_cached_buf *c = nd->cache; c; c = c->next)
			if (c->bytes == bytes) {
				cached = 1;
				break;
			}
		if (fits || cached) {
			void *p = raw_alloc(node, bytes);
			if (p) {
				if (!cached)
					nd->used_bytes += (long long)bytes;
				return p;
			}
		}
		/* out of budget or of device memory: drop cached buffers, then
		 * evict least recently used replicates */
		if (nd->cache) {
			sp_node_release_cache(node);
			continue;
		}
		if (!evict_one(node))
			return NULL;
	}
}

/* include/starpu_data_interfaces.h:1144,1158: bytes on a memory node,
 * counted in its budget; a device buffer freed this way goes to the node's
 * reuse cache instead of back to the driver (a hipFree waits for the whole
 * device). */
uintptr_t starpu_malloc_on_node(unsigned node, size_t size)
{
	if ((int)node < 0 || (int)node >= sp.nnodes)
		return 0;
	SP_LOCK();
	void *p = sp_node_alloc((int)node, size ? size : 1);
	SP_UNLOCK();
	return (uintptr_t)p;
}

void starpu_free_on_node(unsigned node, uintptr_t addr, size_t size)
{
	if (!addr || (int)node < 0 || (int)node >= sp.nnodes)
		return;
	SP_LOCK();
	sp_node_free((int)node, (void *)addr, size ? size : 1);
	SP_UNLOCK();
}

/* Allocation that fits the node's budget as it is (a cached buffer of the
 * size, or free budget): nothing is evicted; NULL otherwise (lock held). */
void *sp_node_try_alloc(int node, size_t bytes)
{
	struct sp_node *nd = &sp.nodes[node];
	int cached = 0;
	for (struct sp_cached_buf *c = nd->cache; c; c = c->next)
		if (c->bytes == bytes) {
			cached = 1;
			break;
		}
	if (!cached && nd->used_bytes + (long long)bytes > nd->total_bytes)
		return NULL;
	void *p = raw_alloc(node, bytes);
	if (p && !cached)
		nd->used_bytes += (long long)bytes;
	return p;
}

void sp_node_free(int node, void *ptr, size_t bytes)
{
	if (!ptr)
		return;
	struct sp_node *nd = &sp.nodes[node];
	if (nd->kind == SP_NODE_CPU) {
		free(ptr);
		nd->used_bytes -= (long long)bytes;
		mem_released(node);
		return;
	}
	/* keep it for reuse (still counted as used) */
	struct sp_cached_buf *c = malloc(sizeof(*c));
	if (!c) {
		sp_hip_free(nd->devid, ptr);
		nd->used_bytes -= (long long)bytes;
		mem_released(node);
		return;
	}
	c->ptr = ptr;
	c->bytes = bytes;
	c->next = nd->cache;
	nd->cache = c;
	nd->cached_bytes += (long long)bytes;
	/* a waiter for node memory may be satisfied by dropping the cache */
	if (nd->waiting_bytes)
		pthread_cond_broadcast(&sp.mem_cond);
}

int sp_copy_2d(int dst_node, void *dst, size_t dld, int src_node, const void *src, size_t sld,
               size_t width, size_t height, void *stream)
{
	const int dk = sp.nodes[dst_node].kind, sk = sp.nodes[src_node].kind;
	if (width == 0 || height == 0)
		return 0;
	if (dk == SP_NODE_CPU && sk == SP_NODE_CPU) {
		if (dld == width && sld == width) {
			memmove(dst, src, width * height);
		} else {
			for (size_t r = 0; r < height; r++)
				memmove((char *)dst + r * dld, (const char *)src + r * sld, width);
		}
		return 0;
	}
	int kind = (sk == SP_NODE_CPU) ? K_H2D : (dk == SP_NODE_CPU) ? K_D2H : K_D2D;
	/* copies into host memory are synchronous (stream NULL) so the host
	 * buffer is complete on return; device-bound copies are queued on the
	 * worker's stream ahead of its kernel */
	void *s = (kind == K_D2H) ? NULL : stream;
	return sp_hip_memcpy2d(dst, dld, src, sld, width, height, kind, s);
}

int sp_copy_iface(const union sp_iface *dst, int dst_node, const union sp_iface *src,
                  int src_node, void *stream)
{
	switch (dst->id) {
	case STARPU_MATRIX_INTERFACE_ID: {
		const size_t es = dst->m.elemsize;
		return sp_copy_2d(dst_node, (void *)dst->m.ptr, dst->m.ld * es, src_node,
		                  (const void *)src->m.ptr, src->m.ld * es, dst->m.nx * es, dst->m.ny, stream);
	}
	case STARPU_VECTOR_INTERFACE_ID: {
		const size_t b = dst->v.nx * dst->v.elemsize;
		return sp_copy_2d(dst_node, (void *)dst->v.ptr, b, src_node, (const void *)src->v.ptr, b, b,
		                  1, stream);
	}
	case STARPU_VARIABLE_INTERFACE_ID: {
		const size_t b = dst->var.elemsize;
		return sp_copy_2d(dst_node, (void *)dst->var.ptr, b, src_node, (const void *)src->var.ptr, b,
		                  b, 1, stream);
	}
	default:
		return -EINVAL;
	}
}

long long starpu_memory_get_total(unsigned node)
{
	if ((int)node >= sp.nnodes)
		return -1;
	return sp.nodes[node].total_bytes;
}

long long starpu_memory_get_used(unsigned node)
{
	if ((int)node >= sp.nnodes)
		return -1;
	return sp.nodes[node].used_bytes;
}

long long starpu_memory_get_available(unsigned node)
{
	if ((int)node >= sp.nnodes)
		return -1;
	struct sp_node *nd = &sp.nodes[node];
	if (nd->total_bytes == LLONG_MAX)
		return LLONG_MAX;
	return nd->total_bytes - nd->used_bytes + nd->cached_bytes;
}

/* ------------------------------------------ application reservations --
 * starpu_memory_allocate / _deallocate / _wait_available
 * (src/datawizard/memory_manager.c:66-141,207-222): an application declares
 * memory it allocates itself on a node, against the same budget as the
 * runtime's data replicates (STARPU_LIMIT_HIP_MEM or 90 % of the device's
 * free memory for a HIP node, unlimited for main RAM).  Before refusing or
 * waiting, the runtime gives back what it can reclaim on the node: its
 * cache of freed device buffers, then unused replicates (written back to
 * their home node), least recently used first. */

static int mem_fits(const struct sp_node *nd, long long bytes)
{
	return nd->total_bytes == LLONG_MAX || nd->used_bytes + bytes <= nd->total_bytes;
}

/* lock held */
static int mem_reclaim(int node, long long bytes)
{
	struct sp_node *nd = &sp.nodes[node];
	if (nd->kind == SP_NODE_CPU)
		return mem_fits(nd, bytes);
	if (!mem_fits(nd, bytes) && nd->cache)
		sp_node_release_cache(node);
	while (!mem_fits(nd, bytes))
		if (!evict_one(node))
			return 0;
	return 1;
}

/* lock held: wait until `bytes` fit on the node */
static void mem_wait(int node, long long bytes)
{
	struct sp_node *nd = &sp.nodes[node];
	int waited = 0;
	while (!mem_reclaim(node, bytes)) {
		if (!nd->waiting_bytes || bytes < nd->waiting_bytes)
			nd->waiting_bytes = bytes;
		if (!waited)
			nd->memory_waits++;
		waited = 1;
		pthread_cond_wait(&sp.mem_cond, &sp.lock);
	}
}

int starpu_memory_allocate(unsigned node, size_t size, int flags)
{
	if ((int)node < 0 || (int)node >= sp.nnodes)
		return -EINVAL;
	SP_LOCK();
	struct sp_node *nd = &sp.nodes[node];
	int ret = 0;
	if (flags & STARPU_MEMORY_WAIT) {
		if (nd->total_bytes != LLONG_MAX && (long long)size > nd->total_bytes)
			ret = -ENOMEM;          /* could never fit: waiting would not end */
		else
			mem_wait((int)node, (long long)size);
	} else if (!(flags & STARPU_MEMORY_OVERFLOW) && !mem_reclaim((int)node, (long long)size)) {
		ret = -ENOMEM;
	}
	if (ret == 0)
		nd->used_bytes += (long long)size;
	SP_UNLOCK();
	return ret;
}

void starpu_memory_deallocate(unsigned node, size_t size)
{
	if ((int)node < 0 || (int)node >= sp.nnodes)
		return;
	SP_LOCK();
	sp.nodes[node].used_bytes -= (long long)size;
	mem_released((int)node);
	SP_UNLOCK();
}

void starpu_memory_wait_available(unsigned node, size_t size)
{
	if ((int)node < 0 || (int)node >= sp.nnodes)
		return;
	SP_LOCK();
	struct sp_node *nd = &sp.nodes[node];
	if (nd->total_bytes == LLONG_MAX || (long long)size <= nd->total_bytes)
		mem_wait((int)node, (long long)size);
	SP_UNLOCK();
}

int starpu_memory_node_stats(unsigned node, long long *evictions, long long *waits, long long *prefetches)
{
	if ((int)node < 0 || (int)node >= sp.nnodes)
		return -EINVAL;
	SP_LOCK();
	if (evictions)
		*evictions = sp.nodes[node].evictions;
	if (waits)
		*waits = sp.nodes[node].memory_waits;
	if (prefetches)
		*prefetches = sp.nodes[node].prefetches;
	SP_UNLOCK();
	return 0;
}
