/*
 * starpu_stdlib.h -- host allocation helpers (reference:
 * include/starpu_stdlib.h:149 starpu_malloc_flags, :158 starpu_free_flags).
 * STARPU_MALLOC_PINNED allocates page-locked memory through the HIP runtime
 * (hipHostMalloc) when a HIP device is present.
 */
#ifndef STARPU_STDLIB_H
#define STARPU_STDLIB_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STARPU_MALLOC_PINNED ((1ULL) << 1)
#define STARPU_MALLOC_COUNT ((1ULL) << 2)
#define STARPU_MALLOC_NORECLAIM ((1ULL) << 3)
#define STARPU_MEMORY_WAIT ((1ULL) << 4)
#define STARPU_MEMORY_OVERFLOW ((1ULL) << 5)
#define STARPU_MALLOC_SIMULATION_FOLDED ((1ULL) << 6)
#define STARPU_MALLOC_SIMULATION_UNIQUE ((1ULL) << 7)

int starpu_malloc_flags(void **A, size_t dim, int flags);
int starpu_free_flags(void *A, size_t dim, int flags);
int starpu_malloc(void **A, size_t dim);
int starpu_free_noflag(void *A, size_t dim);
void starpu_malloc_set_align(size_t align);
/* Bytes the runtime may allocate on a memory node and bytes in use. */
long long starpu_memory_get_total(unsigned node);
long long starpu_memory_get_available(unsigned node);
long long starpu_memory_get_used(unsigned node);

/* Declares `size` bytes that the application allocates itself on `node`,
 * against the node's budget (include/starpu_stdlib.h:258 of the reference).
 * Without flags: 0 if they fit (after the runtime reclaims its cached
 * buffers and unused replicates on the node), -ENOMEM otherwise.
 * STARPU_MEMORY_WAIT: blocks until they fit (-ENOMEM at once if they
 * exceed the whole budget).  STARPU_MEMORY_OVERFLOW: always succeeds,
 * possibly exceeding the budget. */
int starpu_memory_allocate(unsigned node, size_t size, int flags);
/* Returns bytes declared by starpu_memory_allocate, possibly in pieces
 * (reference :272); wakes waiters for which there is now room. */
void starpu_memory_deallocate(unsigned node, size_t size);
/* Blocks until `size` bytes could be allocated on `node` (reference :285);
 * nothing is reserved. */
void starpu_memory_wait_available(unsigned node, size_t size);
/* Replicates the runtime evicted from `node` to make room, the waits of
 * memory reservations on it, and the replicates it filled ahead of their
 * task by prefetch (no reference counterpart).  NULL pointers are skipped. */
int starpu_memory_node_stats(unsigned node, long long *evictions, long long *waits, long long *prefetches);

#ifdef __cplusplus
}
#endif
#endif
