/*
 * starpu_perfmodel.h -- performance models feeding the dmda scheduler
 * (reference: include/starpu_perfmodel.h; history-based model
 * src/core/perfmodel/perfmodel_history.c:1740-2085; v45 file format
 * :834-968,1216,1379; footprints src/datawizard/footprint.c:23-116).
 */
#ifndef STARPU_PERFMODEL_H
#define STARPU_PERFMODEL_H

#include <stdint.h>
#include <stdio.h>

#include "starpu_worker.h"

#ifdef __cplusplus
extern "C" {
#endif

struct starpu_task;

enum starpu_perfmodel_type {
	STARPU_PERFMODEL_INVALID = 0,
	STARPU_PER_WORKER,
	STARPU_PER_ARCH,
	STARPU_COMMON,
	STARPU_HISTORY_BASED,
	STARPU_REGRESSION_BASED,
	STARPU_NL_REGRESSION_BASED,
	STARPU_MULTIPLE_REGRESSION_BASED
};

struct starpu_perfmodel_device {
	enum starpu_worker_archtype type;
	int devid;
	int ncores;
};

struct starpu_perfmodel_arch {
	int ndevices;
	struct starpu_perfmodel_device *devices;
};

struct starpu_perfmodel_history_table;

struct starpu_perfmodel {
	enum starpu_perfmodel_type type;
	/* STARPU_COMMON: cost in microseconds for any worker. */
	double (*cost_function)(struct starpu_task *, unsigned nimpl);
	/* STARPU_PER_ARCH: cost in microseconds for one architecture. */
	double (*arch_cost_function)(struct starpu_task *, struct starpu_perfmodel_arch *arch, unsigned nimpl);
	/* STARPU_PER_WORKER: cost in microseconds on one worker. */
	double (*worker_cost_function)(struct starpu_task *, unsigned workerid, unsigned nimpl);
	size_t (*size_base)(struct starpu_task *, unsigned nimpl);
	uint32_t (*footprint)(struct starpu_task *);
	const char *symbol;
	unsigned is_loaded;
	unsigned benchmarking;
	unsigned is_init;
	/* runtime-owned state (history tables per architecture combination) */
	void *state;
};

/* Expected length in microseconds of `task` on `arch`, NaN when the model
 * is not calibrated there (fewer than STARPU_CALIBRATE_MINIMUM samples). */
double starpu_perfmodel_history_based_expected_perf(struct starpu_perfmodel *model,
                                                    struct starpu_perfmodel_arch *arch,
                                                    uint32_t footprint);
/* Feeds one measurement (microseconds) into the history of `model` for
 * the footprint of `task` on `arch` (reference tests/perfmodels/feed.c). */
void starpu_perfmodel_update_history(struct starpu_perfmodel *model, struct starpu_task *task,
                                     struct starpu_perfmodel_arch *arch, unsigned cpuid,
                                     unsigned nimpl, double measured);
void starpu_perfmodel_update_history_n(struct starpu_perfmodel *model, struct starpu_task *task,
                                       struct starpu_perfmodel_arch *arch, unsigned cpuid,
                                       unsigned nimpl, double average, unsigned count);
void starpu_perfmodel_init(struct starpu_perfmodel *model);
/* Loads $STARPU_PERF_MODEL_DIR/codelets/45/<symbol>.<hostname> if present. */
int starpu_perfmodel_load_symbol(const char *symbol, struct starpu_perfmodel *model);
int starpu_perfmodel_load_file(const char *filename, struct starpu_perfmodel *model);
int starpu_save_history_based_model(struct starpu_perfmodel *model);
int starpu_perfmodel_dump_file(struct starpu_perfmodel *model, const char *path);
int starpu_perfmodel_unload_model(struct starpu_perfmodel *model);
void starpu_perfmodel_get_model_path(const char *symbol, char *path, size_t maxlen);
/* Number of history entries and their (footprint, size, flops, mean, dev,
 * n) values for `arch`, for inspection in tests and tools. */
int starpu_perfmodel_history_count(struct starpu_perfmodel *model, struct starpu_perfmodel_arch *arch);
int starpu_perfmodel_history_entry(struct starpu_perfmodel *model, struct starpu_perfmodel_arch *arch,
                                   int idx, uint32_t *footprint, size_t *size, double *flops,
                                   double *mean, double *dev, unsigned *nsample);

struct starpu_perfmodel_arch *starpu_worker_get_perf_archtype(int workerid, unsigned sched_ctx_id);
uint32_t starpu_task_footprint(struct starpu_perfmodel *model, struct starpu_task *task,
                               struct starpu_perfmodel_arch *arch, unsigned nimpl);
size_t starpu_task_data_footprint_size(struct starpu_task *task);

#ifdef __cplusplus
}
#endif
#endif
