/*
 * sweeper.c -- Quantities and the drop-in Sweeper_* entry points.
 *
 *   Quantities_create   src/3_sweeper/quantities_testing_c.h:28-35, 206-333
 *                       (one process: ix_base = iy_base = 0, global = local)
 *   Sweeper_create      src/3_sweeper/sweeper_kba_c.h:44-287 (same flags,
 *                       same Insist checks; the plan replaces scheduler,
 *                       faces and scratch)
 *   Sweeper_sweep       src/3_sweeper/sweeper_kba_c.h:588-845
 *   Sweeper_destroy     src/3_sweeper/sweeper_kba_c.h:292-329
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "sweeper.h"

#define MSW_INSIST_RC(rc)                                                    \
  do {                                                                       \
    if ((rc) != MSW_OK) {                                                    \
      char buf_[512];                                                        \
      snprintf(buf_, sizeof(buf_), "sweep library: %s", msw_last_error());   \
      insist_(buf_, __FILE__, __LINE__);                                     \
    }                                                                        \
  } while (0)

/* ---- Quantities ---------------------------------------------------------- */

void Quantities_create(Quantities* quan, const Dimensions dims, Env* env) {
  const size_t n = (size_t)dims.nm * dims.na * NOCTANT;
  Pointer_create(&quan->a_from_m, n, Env_cuda_is_using_device(env));
  Pointer_create(&quan->m_from_a, n, Env_cuda_is_using_device(env));
  Pointer_allocate(&quan->a_from_m);
  Pointer_allocate(&quan->m_from_a);
  MSW_INSIST_RC(msw_init_matrices_host(dims.nm, dims.na, Pointer_h(&quan->a_from_m),
                                       Pointer_h(&quan->m_from_a)));
  Pointer_update_d(&quan->a_from_m);
  Pointer_update_d(&quan->m_from_a);

  quan->ix_base_vals = (int*)malloc(2 * sizeof(int));
  quan->iy_base_vals = (int*)malloc(2 * sizeof(int));
  quan->ix_base_vals[0] = 0;
  quan->ix_base_vals[1] = dims.ncell_x;
  quan->iy_base_vals[0] = 0;
  quan->iy_base_vals[1] = dims.ncell_y;
  quan->ix_base = 0;
  quan->iy_base = 0;
  quan->ncell_x_g = dims.ncell_x;
  quan->ncell_y_g = dims.ncell_y;
  quan->ncell_z_g = dims.ncell_z;
}

void Quantities_destroy(Quantities* quan) {
  Pointer_destroy(&quan->a_from_m);
  Pointer_destroy(&quan->m_from_a);
  free(quan->ix_base_vals);
  free(quan->iy_base_vals);
  quan->ix_base_vals = NULL;
  quan->iy_base_vals = NULL;
}

/* src/3_sweeper/quantities_testing_c.h:356-359 */
double Quantities_flops_per_solve(const Dimensions dims) {
  (void)dims;
  return 3. + 3. * NDIM;
}

/* ---- Sweeper ------------------------------------------------------------- */

Sweeper Sweeper_null(void) {
  Sweeper s;
  memset(&s, 0, sizeof(s));
  return s;
}

static int is_pow2_upto8(int v) { return v > 0 && v <= NOCTANT && (v & (v - 1)) == 0; }

void Sweeper_create(Sweeper* sweeper, Dimensions dims, const Quantities* quan,
                    Env* env, Arguments* args) {
  int nsemiblock_default;
  int dims_b_ncell_z;
  sweeper->is_face_comm_async =
      Arguments_consume_int_or_default(args, "--is_face_comm_async", Bool_true);

  Insist(dims.ncell_x > 0 ? "Currently required that all spatial blocks be nonempty" : 0);
  Insist(dims.ncell_y > 0 ? "Currently required that all spatial blocks be nonempty" : 0);
  Insist(dims.ncell_z > 0 ? "Currently required that all spatial blocks be nonempty" : 0);

  sweeper->nblock_z = Arguments_consume_int_or_default(args, "--nblock_z", 1);
  Insist(sweeper->nblock_z > 0 ? "Invalid z blocking factor supplied" : 0);
  Insist(dims.ncell_z % sweeper->nblock_z == 0
             ? "Currently require all blocks have same z dimension" : 0);
  dims_b_ncell_z = dims.ncell_z / sweeper->nblock_z;

  /* The device runs every octant concurrently, so thread counts are
   * accepted for any value the reference's device build accepts. */
  sweeper->nthread_octant = Arguments_consume_int_or_default(args, "--nthread_octant", 1);
  Insist(is_pow2_upto8(sweeper->nthread_octant) ? "Invalid thread count supplied" : 0);
  sweeper->noctant_per_block = sweeper->nthread_octant;
  sweeper->nblock_octant = NOCTANT / sweeper->noctant_per_block;

  nsemiblock_default = sweeper->nthread_octant == 8 && sweeper->nblock_z % 2 == 0
                           ? 4 : sweeper->nthread_octant;
  sweeper->nsemiblock = Arguments_consume_int_or_default(args, "--nsemiblock",
                                                         nsemiblock_default);
  Insist(is_pow2_upto8(sweeper->nsemiblock) ? "Invalid semiblock count supplied" : 0);
  Insist((sweeper->nsemiblock >= sweeper->nthread_octant ||
          (sweeper->nthread_octant == 8 && sweeper->nblock_z % 2 == 0 &&
           sweeper->nsemiblock == 4))
             ? "Incomplete set of semiblock steps requires atomic vo update" : 0);

  sweeper->ncell_x_per_subblock = Arguments_consume_int_or_default(
      args, "--ncell_x_per_subblock",
      sweeper->nsemiblock >= 2 ? (dims.ncell_x + 1) / 2 : dims.ncell_x);
  Insist(sweeper->ncell_x_per_subblock > 0 ? "Invalid subblock size supplied" : 0);
  sweeper->ncell_y_per_subblock = Arguments_consume_int_or_default(
      args, "--ncell_y_per_subblock",
      sweeper->nsemiblock >= 4 ? (dims.ncell_y + 1) / 2 : dims.ncell_y);
  Insist(sweeper->ncell_y_per_subblock > 0 ? "Invalid subblock size supplied" : 0);
  sweeper->ncell_z_per_subblock = Arguments_consume_int_or_default(
      args, "--ncell_z_per_subblock",
      sweeper->nsemiblock >= 8 ? (dims_b_ncell_z + 1) / 2 : dims_b_ncell_z);
  Insist(sweeper->ncell_z_per_subblock > 0 ? "Invalid subblock size supplied" : 0);

  sweeper->dims = dims;
  sweeper->dims_b = dims;
  sweeper->dims_b.ncell_z = dims_b_ncell_z;
  sweeper->dims_g = dims;
  sweeper->dims_g.ncell_x = quan->ncell_x_g;
  sweeper->dims_g.ncell_y = quan->ncell_y_g;

  sweeper->nthread_e = Arguments_consume_int_or_default(args, "--nthread_e", 1);
  Insist(sweeper->nthread_e > 0 ? "Invalid thread count supplied." : 0);
  sweeper->nthread_x = 1;
  sweeper->nthread_y = Arguments_consume_int_or_default(args, "--nthread_y", 1);
  Insist(sweeper->nthread_y > 0 ? "Invalid thread count supplied." : 0);
  sweeper->nthread_z = Arguments_consume_int_or_default(args, "--nthread_z", 1);
  Insist(sweeper->nthread_z > 0 ? "Invalid thread count supplied." : 0);

  sweeper->precision = env->precision_;
  /* z blocks are built into the FP64 sweep only (the mixed forms keep their
     FP32 running sum for the whole grid); refused rather than ignored */
  Insist((sweeper->nblock_z == 1 || sweeper->precision == MSW_PREC_FP64)
             ? "z blocking (--nblock_z > 1) is supported in FP64 only" : 0);
  {
    MswDims d;
    d.ncell_x = dims.ncell_x;
    d.ncell_y = dims.ncell_y;
    d.ncell_z = dims.ncell_z;
    d.ne = dims.ne;
    d.nm = dims.nm;
    d.na = dims.na;
    d.nu = NU;
    sweeper->plan = NULL;
    /* --nblock_z: the FP64 sweep runs z block by z block, as the reference's
       (sweeper_kba_c.h:719-781); the mixed forms sweep the grid at once */
    MswPlanOptions opt;
    msw_plan_default_options(&opt);
    opt.precision = sweeper->precision;
    if (sweeper->nblock_z > 1 && sweeper->precision == MSW_PREC_FP64)
      opt.z_block = dims.ncell_z / sweeper->nblock_z;
    MSW_INSIST_RC(msw_plan_create_ex(&d, &opt, &sweeper->plan));
  }
}

void Sweeper_destroy(Sweeper* sweeper, Env* env) {
  (void)env;
  if (sweeper->plan) msw_plan_destroy(sweeper->plan);
  sweeper->plan = NULL;
}

void Sweeper_sweep(Sweeper* sweeper, Pointer* vo, Pointer* vi,
                   const Quantities* quan, Env* env) {
  Insist(sweeper->plan != NULL);
  Insist(vi->d_ != NULL && vo->d_ != NULL ? "Sweep state must have device storage" : 0);
  if (!env->is_state_resident_) Pointer_update_d_stream(vi, Env_stream(env));
  MSW_INSIST_RC(msw_plan_sweep(sweeper->plan, Pointer_const_d(vi), Pointer_d(vo),
                               Pointer_const_d(&quan->a_from_m),
                               Pointer_const_d(&quan->m_from_a), Env_stream(env)));
  if (!env->is_state_resident_) {
    Pointer_update_h_stream(vo, Env_stream(env));
    if (hipStreamSynchronize((hipStream_t)Env_stream(env)) != hipSuccess)
      insist_("device sweep failed", __FILE__, __LINE__);
  }
}
