/*
 * starpu_task_dep.h -- tag dependencies (reference:
 * include/starpu_task_dep.h:160 starpu_tag_declare_deps, :188
 * starpu_tag_wait; src/core/dependencies/tags.c).
 */
#ifndef STARPU_TASK_DEP_H
#define STARPU_TASK_DEP_H

#include "starpu_task.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The task carrying tag `id` (task->use_tag = 1) starts only after the
 * tasks carrying the ndeps listed tags have completed. */
void starpu_tag_declare_deps(starpu_tag_t id, unsigned ndeps, ...);
void starpu_tag_declare_deps_array(starpu_tag_t id, unsigned ndeps, starpu_tag_t *array);
int starpu_tag_wait(starpu_tag_t id);
int starpu_tag_wait_array(unsigned ntags, starpu_tag_t *id);
void starpu_tag_notify_from_apps(starpu_tag_t id);
void starpu_tag_restart(starpu_tag_t id);
void starpu_tag_remove(starpu_tag_t id);

#ifdef __cplusplus
}
#endif
#endif
