/*
 * msw_batch.h -- batched and served sweeps (libminisweep.so, this library).
 *
 * The reference runs one problem per process (Runner_run_case,
 * src/4_driver/runner.c:51-182, calling Sweeper_sweep once per iteration);
 * these entry points have no counterpart there.  They run many independent
 * sweeps of uneven shape together on one GPU:
 *
 *   MswBatch   a fixed set of problems swept together; the whole run (matrix
 *              staging, boundary fill and every tile launch of every problem)
 *              is captured once into a hipGraph and replayed with one
 *              hipGraphLaunch per run.  An autotuner picks the angles per wave
 *              of each (nm, na) class and the schedule heuristic by timing.
 *   MswServer  a request loop: problems are admitted by HBM footprint into a
 *              device arena while others are in flight (continuous batching),
 *              their state stays resident until they finish, and completion
 *              times give p50/p99 latency.
 *
 * Every array argument is a DEVICE pointer unless named *_host; streams are
 * hipStream_t passed as void*.  Return 0 or an MSW_ERR_* code
 * (msw_last_error() for the text).  A problem's result is the same bits as a
 * single msw_plan_sweep of it with the same precision and angles per wave.
 */
#ifndef MSW_BATCH_H
#define MSW_BATCH_H

#include <stddef.h>

#include "msw_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- batched sweeps -------------------------------------------------------- */

typedef struct MswBatch MswBatch;

typedef struct {
  int precision;     /* MSW_PREC_* for every problem of the batch (default FP64;
                        MSW_PREC_MIXED_FP32 and _BF16 are lossy, see msw_plan.h) */
  int use_graph;     /* 1 (default): run = one replay of a captured hipGraph */
  int schedule_kind; /* tile schedule heuristic: 0 most-remaining octant first, 1 the
                        reference's octant order */
  int stagger;       /* 1: start problems at offsets that even out the work of
                        each combined launch; 0: all start together */
  int row_groups;    /* rows of every tile in groups of 4: 1, 2 or 4 (0: the
                        plan default of the precision) */
  int stream_target; /* z-stream steps per workgroup a tick aims for: energy
                        groups streamed = target / ncell_z (0: 128) */
  int xcd_map;       /* 1 (default): the unknowns of a tile get workgroup ids on
                        one XCD, as in a single sweep (each problem's launch
                        padded to whole rounds of the 8 XCDs); 0: ids in order */
} MswBatchOptions;

void msw_batch_default_options(MswBatchOptions* opt);

/* nprob problems with shapes dims[0..nprob-1] (nm 4, 8 or 16; nu any). */
int msw_batch_create(int nprob, const MswDims* dims, const MswBatchOptions* opt,
                     MswBatch** batch);

/* Buffers of problem i (reference layouts, as msw_plan_sweep).  The matrices
 * are read at every run. */
int msw_batch_set_problem(MswBatch* batch, int i, const double* vi, double* vo,
                          const double* a_from_m, const double* m_from_a);

/* vo_i <- S(vi_i) for every problem, enqueued on stream.  The first run after
 * a change of problems or tuning builds (and captures) the launch sequence. */
int msw_batch_run(MswBatch* batch, void* stream);

/* Autotune: time each candidate (angles per wave of every (nm, na) class in
 * turn, then the tile height in row groups, the energy groups streamed per
 * workgroup, the classes' angles per wave once more, the schedule heuristic
 * and staggering) over reps runs and keep the fastest.  Runs the batch many
 * times; the problems' vo are overwritten.  Returns MSW_ERR_UNSUPPORTED unless
 * the environment sets MSW_BATCH_AUTOTUNE=1 (a GPU run of it ended with a host
 * fault of unknown cause); the batch then keeps its default configuration. */
int msw_batch_autotune(MswBatch* batch, void* stream, int reps);

/* Angles per wave used for problems with this (nm, na) (0 = default). */
int msw_batch_set_class_nac(MswBatch* batch, int nm, int na, int angles_per_wave);

/* Kernel launches in one run; device bytes held; chosen configuration as a
 * JSON text (truncated to len). */
int msw_batch_num_kernel_launches(const MswBatch* batch);
size_t msw_batch_workspace_bytes(const MswBatch* batch);
int msw_batch_describe(const MswBatch* batch, char* buf, int len);

int msw_batch_destroy(MswBatch* batch);

/* ---- served sweeps -------------------------------------------------------- */

typedef struct MswServer MswServer;

typedef struct {
  int precision;            /* MSW_PREC_* of every request (default FP64) */
  size_t hbm_budget_bytes;  /* device arena for admitted problems */
  int max_active;           /* problems in flight at once (1 = sequential) */
  int keep_outputs;         /* 1: keep vo in HBM until msw_server_fetch (a tag
                               completed again before its fetch keeps only the
                               newer result) */
  int schedule_kind;        /* tile schedule heuristic, as MswPlanOptions (0 most
                               remaining octant first, 1 the reference's octant
                               order); msw_server_tune_from_batch replaces it */
  int xcd_map;              /* as MswBatchOptions.xcd_map (default 1) */
} MswServerOptions;

typedef struct {
  long long tag;
  double normsq, normsqdiff; /* FP64, on the device, of (vi, vo) */
  double t_submit, t_admit, t_done; /* seconds on the server's clock */
  int status;                /* MSW_OK, or why the request failed */
} MswResult;

void msw_server_default_options(MswServerOptions* opt);

/* Starts the server's host thread (it owns a HIP stream on the current
 * device). */
int msw_server_create(const MswServerOptions* opt, MswServer** server);

/* Queue one sweep of the manufactured problem of shape d (vi_host NULL:
 * Quantities_init_state on the device; else size_state doubles copied from
 * host memory at submit). */
int msw_server_submit(MswServer* server, const MswDims* d, long long tag,
                      const double* vi_host);

/* Serve with the kernel configuration an autotuned batch of the same
 * precision settled on (angles per wave per (nm, na) class, tile rows, stream
 * length, schedule heuristic): tune once on representative problems, then
 * serve.  Applies to requests admitted after the call.  A request's result is
 * then the same bits as its msw_plan_sweep with those angles per wave and
 * tile rows. */
int msw_server_tune_from_batch(MswServer* server, const MswBatch* batch);

/* Completed requests (up to max); returns the count or a negated error. */
int msw_server_poll(MswServer* server, MswResult* out, int max);

/* With keep_outputs: copy request tag's vo to host memory and release it. */
int msw_server_fetch(MswServer* server, long long tag, double* vo_host);

/* Block until nothing is queued or in flight (timeout_s <= 0: no limit);
 * returns MSW_ERR_INVALID on timeout. */
int msw_server_drain(MswServer* server, double timeout_s);

/* The server's clock (seconds) and counters: out[0] requests done, out[1]
 * kernel launches, out[2] combined ticks, out[3] peak arena bytes used. */
double msw_server_now(const MswServer* server);
int msw_server_stats(const MswServer* server, double* out4);

int msw_server_destroy(MswServer* server);

#ifdef __cplusplus
}
#endif

#endif
