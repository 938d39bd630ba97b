/*
 * starpu_data_interfaces.h -- matrix, vector and variable data interfaces
 * (reference: include/starpu_data_interfaces.h; matrix :1213-1227,
 * register :1245, accessors :1370-1413; variable :2332-2401).
 * The structure fields are the codelet ABI: a codelet receives
 * `void *buffers[]` whose entries point to these structures.
 */
#ifndef STARPU_DATA_INTERFACES_H
#define STARPU_DATA_INTERFACES_H

#include <stddef.h>
#include <stdint.h>

#include "starpu_data.h"

#ifdef __cplusplus
extern "C" {
#endif

enum starpu_data_interface_id {
	STARPU_UNKNOWN_INTERFACE_ID = -1,
	STARPU_MATRIX_INTERFACE_ID = 0,
	STARPU_BLOCK_INTERFACE_ID = 1,
	STARPU_VECTOR_INTERFACE_ID = 2,
	STARPU_CSR_INTERFACE_ID = 3,
	STARPU_BCSR_INTERFACE_ID = 4,
	STARPU_VARIABLE_INTERFACE_ID = 5,
	STARPU_VOID_INTERFACE_ID = 6,
	STARPU_MULTIFORMAT_INTERFACE_ID = 7,
	STARPU_COO_INTERFACE_ID = 8,
	STARPU_TENSOR_INTERFACE_ID = 9,
	STARPU_NDIM_INTERFACE_ID = 10,
	STARPU_MAX_INTERFACE_ID = 11
};

/* Dense column-major matrix: nx rows (contiguous), ny columns, ld elements
 * between the starts of consecutive columns. */
struct starpu_matrix_interface {
	enum starpu_data_interface_id id;
	uintptr_t ptr;
	uintptr_t dev_handle;
	size_t offset;
	size_t nx;
	size_t ny;
	size_t ld;
	size_t elemsize;
	size_t allocsize;
};

/* include/starpu_data_interfaces.h:1144,1158 of the reference: size bytes
 * on memory node dst_node (0 on failure), and their release. */
uintptr_t starpu_malloc_on_node(unsigned dst_node, size_t size);
void starpu_free_on_node(unsigned dst_node, uintptr_t addr, size_t size);

void starpu_matrix_data_register(starpu_data_handle_t *handle, int home_node, uintptr_t ptr,
                                 size_t ld, size_t nx, size_t ny, size_t elemsize);
size_t starpu_matrix_get_nx(starpu_data_handle_t handle);
size_t starpu_matrix_get_ny(starpu_data_handle_t handle);
size_t starpu_matrix_get_local_ld(starpu_data_handle_t handle);
uintptr_t starpu_matrix_get_local_ptr(starpu_data_handle_t handle);
size_t starpu_matrix_get_elemsize(starpu_data_handle_t handle);

#define STARPU_MATRIX_GET_PTR(interface) (((struct starpu_matrix_interface *)(interface))->ptr)
#define STARPU_MATRIX_GET_DEV_HANDLE(interface) (((struct starpu_matrix_interface *)(interface))->dev_handle)
#define STARPU_MATRIX_GET_OFFSET(interface) (((struct starpu_matrix_interface *)(interface))->offset)
#define STARPU_MATRIX_GET_NX(interface) (((struct starpu_matrix_interface *)(interface))->nx)
#define STARPU_MATRIX_GET_NY(interface) (((struct starpu_matrix_interface *)(interface))->ny)
#define STARPU_MATRIX_GET_LD(interface) (((struct starpu_matrix_interface *)(interface))->ld)
#define STARPU_MATRIX_GET_ELEMSIZE(interface) (((struct starpu_matrix_interface *)(interface))->elemsize)
#define STARPU_MATRIX_GET_ALLOCSIZE(interface) (((struct starpu_matrix_interface *)(interface))->allocsize)

/* Contiguous vector of nx elements. */
struct starpu_vector_interface {
	enum starpu_data_interface_id id;
	uintptr_t ptr;
	uintptr_t dev_handle;
	size_t offset;
	size_t nx;
	size_t elemsize;
	size_t slice_base;
	size_t allocsize;
};

void starpu_vector_data_register(starpu_data_handle_t *handle, int home_node, uintptr_t ptr,
                                 size_t nx, size_t elemsize);
size_t starpu_vector_get_nx(starpu_data_handle_t handle);
uintptr_t starpu_vector_get_local_ptr(starpu_data_handle_t handle);

#define STARPU_VECTOR_GET_PTR(interface) (((struct starpu_vector_interface *)(interface))->ptr)
#define STARPU_VECTOR_GET_DEV_HANDLE(interface) (((struct starpu_vector_interface *)(interface))->dev_handle)
#define STARPU_VECTOR_GET_OFFSET(interface) (((struct starpu_vector_interface *)(interface))->offset)
#define STARPU_VECTOR_GET_NX(interface) (((struct starpu_vector_interface *)(interface))->nx)
#define STARPU_VECTOR_GET_ELEMSIZE(interface) (((struct starpu_vector_interface *)(interface))->elemsize)
#define STARPU_VECTOR_GET_ALLOCSIZE(interface) (((struct starpu_vector_interface *)(interface))->allocsize)

/* One element of elemsize bytes (scalars, workspaces, info flags). */
struct starpu_variable_interface {
	enum starpu_data_interface_id id;
	uintptr_t ptr;
	uintptr_t dev_handle;
	size_t offset;
	size_t elemsize;
};

void starpu_variable_data_register(starpu_data_handle_t *handle, int home_node, uintptr_t ptr,
                                   size_t size);
uintptr_t starpu_variable_get_local_ptr(starpu_data_handle_t handle);

#define STARPU_VARIABLE_GET_PTR(interface) (((struct starpu_variable_interface *)(interface))->ptr)
#define STARPU_VARIABLE_GET_OFFSET(interface) (((struct starpu_variable_interface *)(interface))->offset)
#define STARPU_VARIABLE_GET_ELEMSIZE(interface) (((struct starpu_variable_interface *)(interface))->elemsize)
#define STARPU_VARIABLE_GET_DEV_HANDLE(interface) (((struct starpu_variable_interface *)(interface))->dev_handle)

#ifdef __cplusplus
}
#endif
#endif
