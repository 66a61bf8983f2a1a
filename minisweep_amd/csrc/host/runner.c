/*
 * runner.c -- Runner_run_case, compare_runs and the host array operations.
 *
 *   Runner_run_case   src/4_driver/runner.c:51-182
 *   compare_runs      src/4_driver/runner.c:187-241
 *   initialize_state, initialize_state_zero, get_state_norms, copy_vector
 *                     src/2_sweeper_base/array_operations.c:26-126
 *
 * Differences, by design: the state is initialised on the host, copied to
 * HBM once, and stays resident through the timed iterations (the reference's
 * device path streamed faces and state over PCIe inside every sweep); the
 * norms are reduced in FP64 on the device.  The timed region is the same:
 * the iteration loop only, after setup.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "runner.h"

Runner Runner_null(void) {
  Runner r;
  memset(&r, 0, sizeof(r));
  return r;
}
void Runner_create(Runner* runner) { (void)runner; }
void Runner_destroy(Runner* runner) { (void)runner; }

static size_t ind_state(const Dimensions d, int nu, int ix, int iy, int iz, int ie,
                        int im, int iu) {
  return (size_t)im + d.nm * ((size_t)iu + nu * ((size_t)ix + d.ncell_x *
         ((size_t)iy + d.ncell_y * ((size_t)ie + (size_t)d.ne * iz))));
}

void initialize_state(P* v, const Dimensions dims, int nu, const Quantities* quan) {
  int ix, iy, iz, ie, im, iu;
  for (iz = 0; iz < dims.ncell_z; ++iz)
    for (iy = 0; iy < dims.ncell_y; ++iy)
      for (ix = 0; ix < dims.ncell_x; ++ix)
        for (ie = 0; ie < dims.ne; ++ie)
          for (iu = 0; iu < nu; ++iu)
            for (im = 0; im < dims.nm; ++im)
              v[ind_state(dims, nu, ix, iy, iz, ie, im, iu)] =
                  Quantities_init_state(quan, ix, iy, iz, ie, im, iu, dims);
}

void initialize_state_zero(P* v, const Dimensions dims, int nu) {
  memset(v, 0, Dimensions_size_state(dims, nu) * sizeof(P));
}

void get_state_norms(const P* vi, const P* vo, const Dimensions dims, int nu,
                     P* normsqp, P* normsqdiffp, Env* env) {
  int ix, iy, iz, ie, im, iu;
  P n = 0, nd = 0;
  for (iz = 0; iz < dims.ncell_z; ++iz)
    for (iy = 0; iy < dims.ncell_y; ++iy)
      for (ix = 0; ix < dims.ncell_x; ++ix)
        for (ie = 0; ie < dims.ne; ++ie)
          for (im = 0; im < dims.nm; ++im)
            for (iu = 0; iu < nu; ++iu) {
              const size_t i = ind_state(dims, nu, ix, iy, iz, ie, im, iu);
              const P df = vi[i] - vo[i];
              n += vo[i] * vo[i];
              nd += df * df;
            }
  *normsqp = Env_sum_P(env, n);
  *normsqdiffp = Env_sum_P(env, nd);
}

void copy_vector(P* vo, const P* vi, size_t n) { memcpy(vo, vi, n * sizeof(P)); }

void Runner_run_case(Runner* runner, Arguments* args, Env* env) {
  Dimensions dims;
  Quantities quan;
  Sweeper sweeper = Sweeper_null();
  Pointer vi = Pointer_null(), vo = Pointer_null();
  int iteration, niterations;
  Timer t1, t2;
  double* norms_d = NULL;
  double norms_h[2];

  runner->normsq = 0;
  runner->normsqdiff = 0;
  runner->time = 0;
  runner->flops = 0;
  runner->floprate = 0;

  dims.ncell_x = Arguments_consume_int_or_default(args, "--ncell_x", 5);
  dims.ncell_y = Arguments_consume_int_or_default(args, "--ncell_y", 5);
  dims.ncell_z = Arguments_consume_int_or_default(args, "--ncell_z", 5);
  dims.ne = Arguments_consume_int_or_default(args, "--ne", 30);
  dims.na = Arguments_consume_int_or_default(args, "--na", 33);
  niterations = Arguments_consume_int_or_default(args, "--niterations", 1);
  /* runtime moment count: this library's extension of the compile-time NM */
  dims.nm = Arguments_consume_int_or_default(args, "--nm", NM);

  Insist(dims.ncell_x > 0 ? "Invalid ncell_x supplied." : 0);
  Insist(dims.ncell_y > 0 ? "Invalid ncell_y supplied." : 0);
  Insist(dims.ncell_z > 0 ? "Invalid ncell_z supplied." : 0);
  Insist(dims.ne > 0 ? "Invalid ne supplied." : 0);
  Insist(dims.nm > 0 ? "Invalid nm supplied." : 0);
  Insist(dims.na > 0 ? "Invalid na supplied." : 0);
  Insist(niterations >= 0 ? "Invalid iteration count supplied." : 0);

  Quantities_create(&quan, dims, env);

  Pointer_create(&vi, Dimensions_size_state(dims, NU), Env_cuda_is_using_device(env));
  Pointer_set_pinned(&vi, Bool_true);
  Pointer_allocate(&vi);
  Pointer_create(&vo, Dimensions_size_state(dims, NU), Env_cuda_is_using_device(env));
  Pointer_set_pinned(&vo, Bool_true);
  Pointer_allocate(&vo);

  initialize_state(Pointer_h(&vi), dims, NU, &quan);
  initialize_state_zero(Pointer_h(&vo), dims, NU);

  Sweeper_create(&sweeper, dims, &quan, env, args);

  Insist(Arguments_are_all_consumed(args) ? "Invalid argument detected." : 0);

  /* state resident in HBM for the iterations */
  Pointer_update_d(&vi);
  Pointer_update_d(&vo);
  Env_set_state_resident(env, Bool_true);

  t1 = Env_get_synced_time(env);
  for (iteration = 0; iteration < niterations; ++iteration)
    Sweeper_sweep(&sweeper, iteration % 2 == 0 ? &vo : &vi,
                  iteration % 2 == 0 ? &vi : &vo, &quan, env);
  t2 = Env_get_synced_time(env);
  runner->time = t2 - t1;
  Env_set_state_resident(env, Bool_false);

  runner->flops = Env_sum_d(env, niterations *
      (Dimensions_size_state(dims, NU) * NOCTANT * 2. * dims.na +
       Dimensions_size_state_angles(dims, NU) * Quantities_flops_per_solve(dims) +
       Dimensions_size_state(dims, NU) * NOCTANT * 2. * dims.na));
  runner->floprate = runner->time <= (Timer)0 ? 0 : runner->flops / runner->time / 1e9;

  {
    MswDims d;
    d.ncell_x = dims.ncell_x;
    d.ncell_y = dims.ncell_y;
    d.ncell_z = dims.ncell_z;
    d.ne = dims.ne;
    d.nm = dims.nm;
    d.na = dims.na;
    d.nu = NU;
    Insist(hipMalloc((void**)&norms_d, 2 * sizeof(double)) == hipSuccess);
    Insist(msw_state_norms(&d, Pointer_const_d(&vi), Pointer_const_d(&vo), norms_d,
                           Env_stream(env)) == MSW_OK);
    Insist(hipMemcpyAsync(norms_h, norms_d, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          (hipStream_t)Env_stream(env)) == hipSuccess);
    Insist(hipStreamSynchronize((hipStream_t)Env_stream(env)) == hipSuccess);
    (void)hipFree(norms_d);
  }
  runner->normsq = Env_sum_P(env, norms_h[0]);
  runner->normsqdiff = Env_sum_P(env, norms_h[1]);

  /* leave the host mirrors current, as the reference does */
  Pointer_update_h(&vi);
  Pointer_update_h(&vo);

  Pointer_destroy(&vi);
  Pointer_destroy(&vo);
  Sweeper_destroy(&sweeper, env);
  Quantities_destroy(&quan);
}

Bool_t compare_runs(const char* argstring1, const char* argstring2, Env* env) {
  Arguments args1 = Arguments_null(), args2 = Arguments_null();
  Runner r1 = Runner_null(), r2 = Runner_null();
  Bool_t pass;

  Arguments_create_from_string(&args1, argstring1);
  Env_set_values(env, &args1);
  printf("%s // ", argstring1);
  Runner_run_case(&r1, &args1, env);

  Arguments_create_from_string(&args2, argstring2);
  Env_set_values(env, &args2);
  printf("%s // ", argstring2);
  Runner_run_case(&r2, &args2, env);

  pass = r1.normsqdiff == P_zero() && r2.normsqdiff == P_zero() &&
         r1.normsq == r2.normsq;
  printf("%e %e %e %e // %i %i %i // %s\n", r1.normsqdiff, r2.normsqdiff, r1.normsq,
         r2.normsq, r1.normsq == r2.normsq, r1.normsqdiff == P_zero(),
         r2.normsqdiff == P_zero(), pass ? "PASS" : "FAIL");
  Arguments_destroy(&args1);
  Arguments_destroy(&args2);
  return pass;
}
