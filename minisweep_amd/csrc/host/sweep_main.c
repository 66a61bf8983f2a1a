/*
 * sweep_main.c -- the `sweep` command line driver (src/4_driver/sweep.c:27-77).
 *
 * Same flags, same result line:
 *   Normsq result: %.8e  diff: %.3e  PASS|FAIL  time: %.3f  GF/s: %.3f
 * and the TESTS banner when run without arguments.
 */
#include <stdio.h>

#include "runner.h"

int main(int argc, char** argv) {
  Env env = Env_null();
  Arguments args = Arguments_null();
  Runner runner = Runner_null();

  Env_initialize(&env, argc, argv);
  Arguments_create(&args, argc, argv);
  Runner_create(&runner);
  Env_set_values(&env, &args);

  Runner_run_case(&runner, &args, &env);

  printf("Normsq result: %.8e  diff: %.3e  %s  time: %.3f  GF/s: %.3f\n",
         (double)runner.normsq, (double)runner.normsqdiff,
         runner.normsqdiff == P_zero() ? "PASS" : "FAIL", (double)runner.time,
         runner.floprate);
  if (argc == 1) {
    const int ntest = 1;
    const int ntest_passed = runner.normsqdiff == P_zero() ? 1 : 0;
    printf("TESTS %i    PASSED %i    FAILED %i\n", ntest, ntest_passed,
           ntest - ntest_passed);
  }
  Runner_destroy(&runner);
  Arguments_destroy(&args);
  Env_finalize(&env);
  return 0;
}
