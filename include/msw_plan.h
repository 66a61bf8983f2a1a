/*
 * msw_plan.h -- flat C-ABI of the MI355X sweep library (libminisweep.so).
 *
 * Plain pointers and sizes; every array argument is a DEVICE pointer (HBM) and
 * every stream is a hipStream_t passed as void* (NULL = the null stream).
 * All functions return 0 on success and a nonzero MSW_ERR_* code otherwise;
 * msw_last_error() gives the message.  These are the entry points the drop-in
 * Sweeper_* API (sweeper.h) is built on, and what a binding from another
 * language calls.
 *
 * Array layouts are the reference's (src/2_sweeper_base/array_accessors_kernels.h):
 *   state      v[iz][ie][iy][ix][iu][im]                       (:82-88)
 *   a_from_m   A[octant][im][ia], im stride = nm               (:346-349)
 *   m_from_a   M[octant][ia][im], im stride = nm               (:410-413)
 */
#ifndef MSW_PLAN_H
#define MSW_PLAN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MSW_OK = 0,
  MSW_ERR_INVALID = 1,   /* bad dimensions / arguments */
  MSW_ERR_HIP = 2,       /* a HIP runtime call failed */
  MSW_ERR_UNSUPPORTED = 3,
  MSW_ERR_NOMEM = 4
};

/* Arithmetic of a sweep. */
enum {
  MSW_PREC_FP64 = 0,        /* plain FP64, parity path */
  MSW_PREC_MIXED_FP32 = 1,  /* FP32 contractions + FP32 face recurrence and octant
                               sum; FP64 state, last-visit sum and norms */
  MSW_PREC_MIXED_BF16 = 2   /* as MIXED_FP32 with BF16 contractions (FP32
                               accumulate): exact only on BF16-exact inputs */
};

/* Problem geometry of one sweep (reference Dimensions + runtime nu). */
typedef struct {
  int ncell_x, ncell_y, ncell_z; /* grid cells */
  int ne;                        /* energy groups */
  int nm;                        /* moments */
  int na;                        /* angles per octant */
  int nu;                        /* unknowns per cell/moment/group */
} MswDims;

typedef struct MswPlan MswPlan;

/* Last error message of the calling thread ("" if none). */
const char* msw_last_error(void);

/* Number of doubles of a state vector: ncx*ncy*ncz*ne*nm*nu. */
size_t msw_size_state(const MswDims* d);

/* Flop count of niterations sweeps, as the reference counts it
 * (src/4_driver/runner.c:162-166). */
double msw_flops(const MswDims* d, int niterations);

/* Kernel family of a plan. */
enum {
  MSW_ENGINE_VALU = 0, /* contractions as per-lane FMAs, matrices in scalar registers */
  MSW_ENGINE_MFMA = 1  /* contractions on the matrix cores (default) */
};

/* Plan options (this library). */
typedef struct {
  int precision;       /* MSW_PREC_* */
  int engine;          /* MSW_ENGINE_* */
  int angles_per_wave; /* 0: chosen from na; MFMA: 16, 32 or 64 (64 in FP64
                          spills registers: correct, slow); VALU: 8, 16 or 32 */
  int groups_per_workgroup; /* MFMA: energy groups streamed through one tile
                               pipeline (0: chosen per launch from ncell_z) */
  int tile_rows;            /* MFMA: cell rows per tile, 4, 8 or 16 (0: chosen) */
  int z_block;              /* MFMA, FP64: sweep the grid z_block planes at a time
                               (must divide ncell_z; 0: all planes at once).  The
                               reference's --nblock_z; needed by msw_plan_sweep_host
                               for states larger than HBM */
  int schedule_kind;        /* order of the (octant, tile) visits: 0 (default) most
                               remaining visits first; 1 the reference's KBA octant
                               order {0,4,2,6,3,7,1,5}
                               (src/3_sweeper/stepscheduler_kba.c:155) */
  int energy_offset;        /* global index of the plan's first energy group (0): the
                               plan sweeps groups [energy_offset, energy_offset + ne)
                               of a larger problem, whose per-group scale factors
                               (src/3_sweeper/quantities_testing_kernels.h:52-70 of the
                               reference: they depend on the global group index)
                               it applies.  Energy groups do not couple, so the
                               groups of one problem can be split over plans (one
                               GPU's share each) and their results are those of the
                               whole problem's groups; state arrays hold the plan's
                               ne groups in the usual layout */
} MswPlanOptions;

void msw_plan_default_options(MswPlanOptions* opt);

/* Create a sweep plan on the current device: schedule, face buffers and
 * the per-octant matrix staging.  Replaces the allocation half of
 * Sweeper_create (src/3_sweeper/sweeper_kba_c.h:44-287). */
int msw_plan_create(const MswDims* d, int precision, MswPlan** plan);

/* Same with explicit options (kernel engine, angles per wave). */
int msw_plan_create_ex(const MswDims* d, const MswPlanOptions* opt, MswPlan** plan);

/* Device bytes the plan holds. */
size_t msw_plan_workspace_bytes(const MswPlan* plan);

/* Number of kernel launches one sweep takes. */
int msw_plan_num_launches(const MswPlan* plan);

/* One sweep vo <- S(vi), all arrays resident in HBM, enqueued on stream.
 * Replaces Sweeper_sweep (src/3_sweeper/sweeper_kba_c.h:588-845).
 * vo need not be zeroed; vi is not modified.  Sweeps of one plan on
 * different streams must not overlap: they share its face planes. */
int msw_plan_sweep(MswPlan* plan, const double* vi, double* vo,
                   const double* a_from_m, const double* m_from_a,
                   void* stream);

/* msw_plan_sweep that also writes each octant's outflow faces at the
 * domain's downstream boundary, in the reference's Faces layouts with all
 * eight octants (src/3_sweeper/array_accessors_kernels.h:481-486, 545-550,
 * 609-614; the reference's Faces of faces_kba.h:29-142 with
 * noctant_per_block 8), FP64 device arrays:
 *   facexy[8][ne][ny][nx][nu][na]  the z faces: each (x, y) column's last cell
 *   facexz[8][ne][nz][nx][nu][na]  the y faces: each (x, z) row's last cell
 *   faceyz[8][ne][nz][ny][nu][na]  the x faces: each (y, z) row's last cell
 * "last" in the octant's direction; each value is the solved angular value
 * times (1+octant), as the reference stores its faces
 * (src/3_sweeper/quantities_testing_kernels.h:437-443).  MFMA engine, plans
 * without z blocks (else MSW_ERR_UNSUPPORTED). */
int msw_plan_sweep_faces(MswPlan* plan, const double* vi, double* vo,
                         const double* a_from_m, const double* m_from_a, double* facexy,
                         double* facexz, double* faceyz, void* stream);

/* Out-of-core sweep (this library; the reference's z-block streaming,
 * src/3_sweeper/sweeper_kba_c.h:719-781): vi_host and vo_host are HOST arrays
 * in the state layout, pinned (hipHostMalloc) for the copies to overlap the
 * sweep.  The state goes through HBM one z block at a time, each block's vi
 * (and, in the second pass, its partial vo) copied in while the previous
 * block is swept, its vo copied back after: two passes of four octants.
 * With n z blocks, 5 - 3/n state volumes cross PCIe per sweep (vi twice, vo
 * back twice, vo in once, less the block at the turn between the passes,
 * which stays on the device); copies in and copies out overlap.  Device
 * memory: two slabs of vi and vo of one block each, plus the plan.
 * a_from_m, m_from_a are device arrays.  Enqueued on stream; synchronize it
 * before reading vo_host.  Calls on one plan must not overlap (they share the
 * slabs and face planes). */
int msw_plan_sweep_host(MswPlan* plan, const double* vi_host, double* vo_host,
                        const double* a_from_m, const double* m_from_a, void* stream);

/* Per-launch timing (this library): when enabled, msw_plan_sweep brackets
 * each sweep-kernel launch with hipEvents on the sweep's stream. */
int msw_plan_set_timing(MswPlan* plan, int enable);

/* Milliseconds of each sweep-kernel launch of the last timed sweep; waits for
 * it to finish.  Fills up to n entries and returns how many, or the negated
 * MSW_ERR_* code on failure. */
int msw_plan_last_timings(MswPlan* plan, float* ms, int n);

/* Release the plan.  Replaces Sweeper_destroy (sweeper_kba_c.h:292-329). */
int msw_plan_destroy(MswPlan* plan);

/* normsq = sum vo^2, normsqdiff = sum (vi-vo)^2 in FP64 on the device,
 * written to out2[0], out2[1] (device pointer).  Deterministic order.  Uses a
 * 16 KB device scratch per calling thread: one thread's calls on different
 * streams must not overlap.
 * Replaces get_state_norms (src/2_sweeper_base/array_operations.c:69-110). */
int msw_state_norms(const MswDims* d, const double* vi, const double* vo,
                    double* out2, void* stream);

/* msw_state_norms without any allocation in the call (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under HIP graph capture or a loop that forbids hipMalloc.  Same kernels,
 * same result bits as msw_state_norms.  The block partials (16 KiB) live in
 * a static device array of the library, one per device, held until the
 * process ends; calls that overlap on one device race on it.  Any valid
 * state size fits (the partials' count does not depend on it). */
int msw_state_norms_noalloc(const MswDims* d, const double* vi, const double* vo,
                            double* out2, void* stream);

/* Fill v with the manufactured input state (Quantities_init_state,
 * src/3_sweeper/quantities_testing.h:60-90) on the device. */
int msw_init_state(const MswDims* d, double* v, void* stream);

/* Host: fill a_from_m, m_from_a (each 8*nm*na doubles, host memory) as
 * Quantities_init_am_matrices_ (src/3_sweeper/quantities_testing_c.h:40-201). */
int msw_init_matrices_host(int nm, int na, double* a_from_m, double* m_from_a);

#ifdef __cplusplus
}
#endif

#endif
