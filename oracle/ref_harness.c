/*
 * ref_harness.c -- drives the compiled reference sweeper on chosen inputs.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference's
 * own sources (serial KBA build, src/ under the reference tree) into
 * oracle/_ref/; it runs only in the development container, to write the
 * golden vectors under tests/golden/.
 *
 * It repeats the setup of Runner_run_case (src/4_driver/runner.c:51-182) but
 * lets the input state be perturbed by a seeded generator so that the sweep is
 * checked on non-affine data, then dumps the final vi and vo buffers.
 *
 * usage: ref_harness "<argstring>" <seed> <perturb> <out.bin>
 *   seed    : 64-bit generator seed
 *   perturb : 0 = manufactured vi, 1 = vi*(1+u/4), 2 = vi = u,
 *             3 = vi*(1+u*2^-8), u in [-1,1)
 * out.bin   : int32 header [ncx ncy ncz ne nm na nu niter] then
 *             double normsq, normsqdiff, then vi0 (input), vi, vo (final).
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>

#include "arguments.h"
#include "env.h"
#include "definitions.h"
#include "dimensions.h"
#include "pointer.h"
#include "quantities.h"
#include "array_operations.h"
#include "sweeper.h"

static uint64_t rng_state;
static double rng_uniform(void) {
  /* splitmix64 */
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s \"<args>\" seed perturb out.bin\n", argv[0]);
    return 2;
  }
  const char* argstring = argv[1];
  rng_state = strtoull(argv[2], NULL, 10);
  const int perturb = atoi(argv[3]);
  const char* outpath = argv[4];

  Env env = Env_null();
  Env_initialize(&env, 1, argv);
  Arguments args = Arguments_null();
  Arguments_create_from_string(&args, argstring);
  Env_set_values(&env, &args);

  Dimensions dims;
  dims.ncell_x = Arguments_consume_int_or_default(&args, "--ncell_x", 5);
  dims.ncell_y = Arguments_consume_int_or_default(&args, "--ncell_y", 5);
  dims.ncell_z = Arguments_consume_int_or_default(&args, "--ncell_z", 5);
  dims.ne = Arguments_consume_int_or_default(&args, "--ne", 30);
  dims.na = Arguments_consume_int_or_default(&args, "--na", 33);
  const int niterations =
      Arguments_consume_int_or_default(&args, "--niterations", 1);
  dims.nm = NM;

  Quantities quan;
  Quantities_create(&quan, dims, &env);

  Pointer vi = Pointer_null();
  Pointer vo = Pointer_null();
  Pointer_create(&vi, Dimensions_size_state(dims, NU), Bool_false);
  Pointer_allocate(&vi);
  Pointer_create(&vo, Dimensions_size_state(dims, NU), Bool_false);
  Pointer_allocate(&vo);

  initialize_state(Pointer_h(&vi), dims, NU, &quan);
  initialize_state_zero(Pointer_h(&vo), dims, NU);
  const size_t n = Dimensions_size_state(dims, NU);
  size_t i;
  for (i = 0; i < n; ++i) {
    if (perturb == 1) Pointer_h(&vi)[i] *= 1.0 + 0.25 * rng_uniform();
    if (perturb == 2) Pointer_h(&vi)[i] = rng_uniform();
    if (perturb == 3) Pointer_h(&vi)[i] *= 1.0 + rng_uniform() / 256.0;
  }
  double* vi0 = (double*)malloc(n * sizeof(double));
  for (i = 0; i < n; ++i) vi0[i] = Pointer_h(&vi)[i];

  Sweeper sweeper = Sweeper_null();
  Sweeper_create(&sweeper, dims, &quan, &env, &args);
  if (!Arguments_are_all_consumed(&args)) {
    fprintf(stderr, "unconsumed arguments\n");
    return 3;
  }
  int it;
  for (it = 0; it < niterations; ++it)
    Sweeper_sweep(&sweeper, it % 2 == 0 ? &vo : &vi, it % 2 == 0 ? &vi : &vo,
                  &quan, &env);

  P normsq = 0, normsqdiff = 0;
  get_state_norms(Pointer_h(&vi), Pointer_h(&vo), dims, NU, &normsq,
                  &normsqdiff, &env);

  FILE* f = fopen(outpath, "wb");
  if (!f) return 4;
  int32_t hdr[8] = {dims.ncell_x, dims.ncell_y, dims.ncell_z, dims.ne,
                    dims.nm,      dims.na,      NU,           niterations};
  fwrite(hdr, sizeof(int32_t), 8, f);
  double norms[2] = {normsq, normsqdiff};
  fwrite(norms, sizeof(double), 2, f);
  fwrite(vi0, sizeof(double), n, f);
  fwrite(Pointer_h(&vi), sizeof(double), n, f);
  fwrite(Pointer_h(&vo), sizeof(double), n, f);
  fclose(f);

  printf("Normsq result: %.8e  diff: %.3e\n", (double)normsq, (double)normsqdiff);

  Sweeper_destroy(&sweeper, &env);
  Pointer_destroy(&vi);
  Pointer_destroy(&vo);
  Quantities_destroy(&quan);
  Arguments_destroy(&args);
  Env_finalize(&env);
  free(vi0);
  return 0;
}
