/*
 * Internal state of the starpu_amd host runtime.
 *
 * Concurrency model: one process-wide mutex (sp.lock) guards every piece of
 * scheduling state -- jobs and their dependency counts, handles and their
 * MSI replicate states, worker queues, performance-model tables.  Codelets
 * run outside the lock.  Condition variables wake CPU workers, the HIP
 * driver thread and application threads blocked in waits.
 */
#ifndef SP_INTERNAL_H
#define SP_INTERNAL_H

#include <pthread.h>
#include <stdint.h>
#include <stdio.h>

#include <starpu.h>

#define SP_MAXNODES STARPU_MAXNODES

/* MSI replicate states (reference src/datawizard/coherency.h:37-42). */
enum sp_rep_state { SP_INVALID = 0, SP_SHARED = 1, SP_OWNER = 2 };

enum sp_node_kind { SP_NODE_CPU = 0, SP_NODE_HIP = 1 };

struct sp_job;
struct sp_worker;

union sp_iface {
	enum starpu_data_interface_id id;
	struct starpu_matrix_interface m;
	struct starpu_vector_interface v;
	struct starpu_variable_interface var;
};

struct sp_replicate {
	struct _starpu_data_state *handle;
	int node;
	int state;          /* enum sp_rep_state */
	int allocated;      /* memory exists on this node */
	int automatic;      /* the runtime allocated it (and frees it) */
	int is_view;        /* points into the parent's allocation */
	int refcnt;         /* running jobs using it: not evictable */
	int wont_use;
	double last_used;
	size_t alloc_bytes;
	union sp_iface iface;
	struct sp_replicate *lru_prev, *lru_next;
	int in_lru;
	/* set when an asynchronous copy filled this replicate on `ready_stream`:
	 * work on any other stream waits on the event before using the data */
	void *ready_event;
	void *ready_stream;
};

struct _starpu_data_state {
	enum starpu_data_interface_id iface_id;
	int home_node;
	union sp_iface shape;           /* interface with ptr = 0: sizes only */
	struct sp_replicate rep[SP_MAXNODES];
	/* per-worker scratch replicates for STARPU_SCRATCH accesses */
	struct sp_replicate *scratch;   /* array [nworkers], lazily allocated */
	/* implicit (sequential-consistency) dependencies */
	unsigned seq_consistency;
	struct sp_job *last_writer;
	struct sp_job **readers;
	int nreaders, readers_cap;
	int pending;                    /* submitted, not terminated jobs using it */
	/* partitioning */
	struct starpu_data_filter filter;  /* filter that produced this child */
	unsigned child_id, nparts;
	struct _starpu_data_state *parent;
	struct _starpu_data_state **children;
	int nchildren;
	int partitioned;
	char *name;
	size_t size;                    /* payload bytes */
	uint32_t footprint;             /* crc32c of the shape (matrix: nx, ny) */
	int lazy_unregister;
	int registered;
};

struct sp_buffer {
	starpu_data_handle_t handle;
	enum starpu_data_access_mode mode;
	struct sp_replicate *rep;       /* replicate used for this execution */
	int node;
};

struct sp_job {
	struct starpu_task *task;
	int nbuffers;
	struct sp_buffer *buf;          /* nbuffers entries */
	struct sp_buffer buf_inline[STARPU_NMAXBUFS];
	void **ifaces;                  /* passed to the codelet */
	void *ifaces_inline[STARPU_NMAXBUFS];
	int ndeps;                      /* unmet dependencies */
	struct sp_job **succ;
	int nsucc, succ_cap;
	int submitted, terminated, running;
	int is_acquire;                 /* application data acquisition */
	int acquire_ready;
	int workerid;
	int nimpl;
	double predicted, predicted_transfer;
	double t_submit, t_push, t_start, t_end;
	double kernel_us;
	size_t bytes_in;
	uint32_t footprint;
	int footprint_valid;
	pthread_cond_t cond;            /* task_wait / acquire waiters */
	struct sp_job *qnext;           /* worker queue link */
	void *hip_ev_start, *hip_ev_end;
	int hip_slot;
	int tag_pending;                /* waiting on its tag's dependencies */
	struct sp_job *explicit_deps_pending_next;
};

struct sp_worker {
	int id;
	enum starpu_worker_archtype arch;
	int devid;                      /* HIP device, or CPU index */
	int node;                       /* memory node */
	int subid;                      /* stream index within the HIP device */
	char name[48];
	/* queue sorted by decreasing priority, FIFO among equals */
	struct sp_job *qhead;
	int qlen;
	double exp_start, exp_len, exp_end;
	pthread_t thread;
	pthread_cond_t cond;            /* CPU worker sleeps here */
	/* HIP */
	void *stream;                   /* hipStream_t */
	int thr;                        /* index of its driver thread in the device */
	int reserved;                   /* runs only the tasks placed on it explicitly */
	int ninflight;
	struct sp_job *inflight[8];
	/* stats */
	unsigned long ntasks;
	double busy_us, kernel_us;
	struct starpu_perfmodel_arch perf_arch;
	struct starpu_perfmodel_device perf_dev;
};

struct sp_node {
	enum sp_node_kind kind;
	int devid;
	long long total_bytes;          /* budget */
	long long used_bytes;
	struct sp_replicate *lru_head, *lru_tail;   /* head = least recently used */
	/* freed device buffers kept for reuse, keyed by size */
	struct sp_cached_buf *cache;
	long long cached_bytes;
	/* smallest amount a starpu_memory_allocate(STARPU_MEMORY_WAIT) or
	 * starpu_memory_wait_available caller waits for (0: none) */
	long long waiting_bytes;
	long long evictions;            /* replicates evicted to make room */
	long long memory_waits;         /* waits of memory reservations */
	long long prefetches;           /* replicates filled ahead of their task */
};

struct sp_cached_buf {
	void *ptr;
	size_t bytes;
	struct sp_cached_buf *next;
};

#define SP_HIP_MAXTHREADS 16
#define SP_EVPOOL 64

/* One driver thread of a device: it drives the device's workers (streams)
 * whose `thr` is its index -- launching their tasks and polling their
 * completion events -- and keeps its own pool of events. */
struct sp_hipthr {
	struct sp_hipdev *hd;
	int index;
	pthread_t thread;
	pthread_cond_t cond;
	void *ev[SP_EVPOOL];            /* hipEvent_t */
	int nev;
};

struct sp_hipdev {
	int devid;
	int node;
	void *xfer_stream;              /* hipStream_t of prefetch copies */
	int first_worker, nworkers;
	int reserved_worker;            /* worker on the reserved CUs, or -1 */
	int nthreads;
	struct sp_hipthr thr[SP_HIP_MAXTHREADS];
	int active;
};

struct sp_tag;

struct sp_runtime {
	pthread_mutex_t lock;
	pthread_cond_t done_cond;       /* wait_for_all, acquire, tag waits */
	pthread_cond_t mem_cond;        /* waiters for node memory */
	int initialized;
	int shutting_down;
	int paused;
	struct sp_worker workers[STARPU_NMAXWORKERS];
	int nworkers, ncpu, nhip;
	struct sp_node nodes[SP_MAXNODES];
	int nnodes;
	struct sp_hipdev hipdevs[STARPU_MAXHIPDEVS];
	int nhipdevs;
	int hip_pipeline;
	int prefetch;                   /* STARPU_PREFETCH: dmda prefetches task inputs */
	long nsubmitted;                /* submitted, not terminated */
	long nready;
	const char *policy;             /* "dmda", "eager", "prio", "random" */
	double alpha, beta;
	int calibrate;
	int calibrate_minimum;
	int profiling;
	/* bus model: bandwidth (bytes/us) and latency (us) between nodes */
	double bw[SP_MAXNODES][SP_MAXNODES];
	double lat[SP_MAXNODES][SP_MAXNODES];
	/* registered perfmodels (to save at shutdown) */
	struct starpu_perfmodel **models;
	int nmodels, models_cap;
	/* tags */
	struct sp_tag **tags;
	int tags_cap, ntags;
	/* trace */
	struct starpu_trace_record *trace;
	size_t ntrace, trace_cap;
	/* watchdog */
	int watchdog_us;
	double last_completion;
	pthread_t watchdog_thread;
	int watchdog_running;
	unsigned default_seq_consistency;
};

extern struct sp_runtime sp;
extern __thread struct sp_worker *sp_current_worker;
extern __thread struct sp_job *sp_current_job;

#define SP_LOCK() pthread_mutex_lock(&sp.lock)
#define SP_UNLOCK() pthread_mutex_unlock(&sp.lock)

/* core.c */
double sp_now(void);
int sp_env_int(const char *name, int def);
double sp_env_double(const char *name, double def);
void sp_wake_all_workers(void);

/* task.c */
struct sp_job *sp_job_of(struct starpu_task *task);
void sp_job_ready(struct sp_job *j);                 /* lock held */
void sp_job_terminate(struct sp_job *j);              /* lock NOT held */
void sp_job_add_dependency(struct sp_job *pred, struct sp_job *succ); /* lock held */
void sp_handle_pending_dec(starpu_data_handle_t h);   /* lock held */

/* data.c */
void sp_data_init(void);
void sp_data_shutdown(void);
int sp_fetch_buffer(struct sp_job *j, int i, struct sp_worker *w, void *stream); /* lock held */
void sp_release_buffer(struct sp_job *j, int i);      /* lock held */
size_t sp_iface_size(const union sp_iface *f);
void sp_detect_implicit_deps(struct sp_job *j);       /* lock held */
void sp_remove_implicit_deps(struct sp_job *j);       /* lock held */
int sp_rep_valid(struct sp_replicate *r);
int sp_handle_valid_node(starpu_data_handle_t h);     /* lock held */
void sp_lru_touch(struct sp_replicate *r);
void sp_prefetch_job(struct sp_job *j, struct sp_worker *w);
void *sp_node_try_alloc(int node, size_t bytes);

/* memory.c */
void *sp_node_alloc(int node, size_t bytes);          /* lock held; may evict */
void sp_node_free(int node, void *ptr, size_t bytes); /* lock held */
void sp_node_release_cache(int node);
int sp_copy_2d(int dst_node, void *dst, size_t dld_bytes, int src_node, const void *src,
               size_t sld_bytes, size_t width_bytes, size_t height, void *stream);
int sp_copy_iface(const union sp_iface *dst, int dst_node, const union sp_iface *src,
                  int src_node, void *stream);
int sp_memory_init(void);

/* sched.c */
void sp_sched_push(struct sp_job *j);                 /* lock held */
struct sp_job *sp_sched_pop(struct sp_worker *w);     /* lock held */
int sp_sched_may_pop(const struct sp_worker *w);      /* no lock: a hint */
void sp_sched_job_started(struct sp_worker *w, struct sp_job *j);
void sp_sched_job_done(struct sp_worker *w, struct sp_job *j);
int sp_worker_can_execute(struct sp_worker *w, struct starpu_task *t, int *nimpl);
double sp_predict_length(struct starpu_task *t, struct sp_worker *w, int nimpl, struct sp_job *j);
void sp_sched_reset(void);

/* perfmodel.c */
void sp_perfmodel_register(struct starpu_perfmodel *m);       /* lock held */
double sp_perfmodel_predict(struct starpu_perfmodel *m, struct starpu_task *t,
                            struct sp_job *j, struct starpu_perfmodel_arch *arch, int nimpl);
void sp_perfmodel_feedback(struct sp_job *j, struct sp_worker *w, double measured_us);
uint32_t sp_job_footprint(struct sp_job *j);
void sp_perfmodel_save_all(void);
void sp_perfmodel_free_all(void);

/* drivers */
int sp_cpu_workers_start(void);
int sp_hip_init(int nhip, int nworker_per_hip, int *nworkers_created);
/* bus model files ($STARPU_PERF_MODEL_DIR/bus/<host>.<type>) */
void sp_bus_path(const char *type, char *path, size_t maxlen);
int sp_bus_save(void);
int sp_bus_load(void);
struct sp_hipthr *sp_hip_worker_thread(const struct sp_worker *w);
int sp_hip_start(void);
void sp_hip_stop(void);
void sp_hip_shutdown(void);
void sp_execute_on_cpu(struct sp_worker *w, struct sp_job *j);

/* hip runtime wrappers used by the runtime (driver_hip.c) */
int sp_hip_malloc(int devid, void **p, size_t bytes);
int sp_hip_free(int devid, void *p);
int sp_hip_memcpy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t w,
                    size_t h, int kind, void *stream);
int sp_hip_stream_sync(void *stream);
void *sp_hip_event_record(void *stream);              /* NULL on failure */
int sp_hip_stream_wait_event(void *stream, void *event);
void sp_hip_event_destroy(void *event);
int sp_hip_event_done(void *event);                   /* 1 once the event completed */
void *sp_hip_transfer_stream(int devid);
long long sp_hip_total_mem(int devid);
int sp_hip_host_alloc(void **p, size_t bytes);
int sp_hip_host_free(void *p);

/* tags.c */
void sp_tag_job_submitted(struct sp_job *j);          /* lock held */
void sp_tag_job_done(struct sp_job *j);               /* lock held */
void sp_tags_free(void);

/* trace */
void sp_trace_add(struct sp_job *j, struct sp_worker *w);

#endif
