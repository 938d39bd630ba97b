/*
 * starpu_data_filters.h -- partitioning of a handle into sub-handles
 * (reference: include/starpu_data_filters.h; struct :40-120, partition
 * :145, get_sub_data :176, map_filters :191, matrix block filters
 * :448,:472; chunk rule src/datawizard/filters.c:1655-1679).
 */
#ifndef STARPU_DATA_FILTERS_H
#define STARPU_DATA_FILTERS_H

#include <stdarg.h>

#include "starpu_data.h"

#ifdef __cplusplus
extern "C" {
#endif

struct starpu_data_filter {
	/* Fill child_interface (the id-th of nparts) from parent_interface. */
	void (*filter_func)(void *parent_interface, void *child_interface,
	                    struct starpu_data_filter *, unsigned id, unsigned nparts);
	unsigned nchildren;
	unsigned (*get_nchildren)(struct starpu_data_filter *, starpu_data_handle_t initial_handle);
	void *filter_arg;
	void *filter_arg_ptr;
};

void starpu_data_partition(starpu_data_handle_t initial_handle, struct starpu_data_filter *f);
/* Gathers the children's valid data into the parent's replicate on
 * gathering_node and makes the parent usable again. */
void starpu_data_unpartition(starpu_data_handle_t root_data, unsigned gathering_node);
starpu_data_handle_t starpu_data_get_child(starpu_data_handle_t handle, unsigned i);
int starpu_data_get_nb_children(starpu_data_handle_t handle);
starpu_data_handle_t starpu_data_get_sub_data(starpu_data_handle_t root_data, unsigned depth, ...);
starpu_data_handle_t starpu_data_vget_sub_data(starpu_data_handle_t root_data, unsigned depth, va_list pa);
void starpu_data_map_filters(starpu_data_handle_t root_data, unsigned nfilters, ...);
void starpu_data_vmap_filters(starpu_data_handle_t root_data, unsigned nfilters, va_list pa);

/* Split the rows (nx) into nparts blocks. */
void starpu_matrix_filter_block(void *parent_interface, void *child_interface,
                                struct starpu_data_filter *f, unsigned id, unsigned nparts);
/* Split the columns (ny) into nparts blocks. */
void starpu_matrix_filter_vertical_block(void *parent_interface, void *child_interface,
                                         struct starpu_data_filter *f, unsigned id, unsigned nparts);
/* Split a vector into nparts contiguous blocks. */
void starpu_vector_filter_block(void *parent_interface, void *child_interface,
                                struct starpu_data_filter *f, unsigned id, unsigned nparts);

void starpu_filter_nparts_compute_chunk_size_and_offset(unsigned n, unsigned nparts, size_t elemsize,
                                                        unsigned id, size_t blocksize,
                                                        size_t *chunk_size, size_t *offset);

#ifdef __cplusplus
}
#endif
#endif
