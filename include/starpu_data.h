/*
 * starpu_data.h -- data handles, access modes and application-side data
 * operations (reference: include/starpu_data.h; modes :57-120, acquire
 * :327, release :485, unregister :229, set_name :193, wont_use :629).
 */
#ifndef STARPU_DATA_H
#define STARPU_DATA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _starpu_data_state *starpu_data_handle_t;

enum starpu_data_access_mode {
	STARPU_NONE = 0,
	STARPU_R = (1 << 0),
	STARPU_W = (1 << 1),
	STARPU_RW = (STARPU_R | STARPU_W),
	STARPU_SCRATCH = (1 << 2),
	STARPU_REDUX = (1 << 3),
	STARPU_COMMUTE = (1 << 4),
	STARPU_SSEND = (1 << 5),
	STARPU_LOCALITY = (1 << 6),
	STARPU_MPI_REDUX = (1 << 7),
	STARPU_NOPLAN = (1 << 8),
	STARPU_UNMAP = (1 << 9),
	STARPU_NOFOOTPRINT = (1 << 10),
	STARPU_ACCESS_MODE_MAX = (1 << 11)
};

#define STARPU_MAIN_RAM 0
/* home_node value for data that has no home (scratch buffers). */
#define STARPU_NO_HOME_NODE (-1)

/* Wait until the handle can be accessed in `mode` from main memory
 * (implicit dependencies apply), fetch a valid copy there and keep it
 * locked until starpu_data_release.  Returns 0. */
int starpu_data_acquire(starpu_data_handle_t handle, enum starpu_data_access_mode mode);
int starpu_data_acquire_on_node(starpu_data_handle_t handle, int node, enum starpu_data_access_mode mode);
void starpu_data_release(starpu_data_handle_t handle);
void starpu_data_release_on_node(starpu_data_handle_t handle, int node);

/* Waits for all tasks using the handle, writes the valid copy back to the
 * home node and frees every replicate the runtime allocated. */
void starpu_data_unregister(starpu_data_handle_t handle);
/* Same, without writing back. */
void starpu_data_unregister_no_coherency(starpu_data_handle_t handle);
void starpu_data_unregister_submit(starpu_data_handle_t handle);

void starpu_data_set_name(starpu_data_handle_t handle, const char *name);
const char *starpu_data_get_name(starpu_data_handle_t handle);
/* Hint that the handle will not be used again soon: its device replicates
 * become the first eviction candidates. */
void starpu_data_wont_use(starpu_data_handle_t handle);
int starpu_data_prefetch_on_node(starpu_data_handle_t handle, unsigned node, unsigned async);
int starpu_data_fetch_on_node(starpu_data_handle_t handle, unsigned node, unsigned async);
void starpu_data_invalidate(starpu_data_handle_t handle);
void starpu_data_set_sequential_consistency_flag(starpu_data_handle_t handle, unsigned flag);
unsigned starpu_data_get_sequential_consistency_flag(starpu_data_handle_t handle);
void starpu_data_set_default_sequential_consistency_flag(unsigned flag);
size_t starpu_data_get_size(starpu_data_handle_t handle);
uint32_t starpu_data_get_footprint(starpu_data_handle_t handle);
void *starpu_data_get_local_ptr(starpu_data_handle_t handle);
void *starpu_data_handle_to_pointer(starpu_data_handle_t handle, unsigned node);
/* 1 if `node` holds a valid (OWNER or SHARED) copy. */
unsigned starpu_data_is_on_node(starpu_data_handle_t handle, unsigned node);
/* Interface of the replicate on `node` (struct starpu_*_interface *). */
void *starpu_data_get_interface_on_node(starpu_data_handle_t handle, unsigned node);
int starpu_data_get_home_node(starpu_data_handle_t handle);

#ifdef __cplusplus
}
#endif
#endif
