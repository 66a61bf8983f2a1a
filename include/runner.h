/*
 * runner.h -- driver API of the MI355X sweep (drop-in for the reference's
 * src/4_driver/runner.h:26-58 and src/2_sweeper_base/array_operations.h:27-55).
 */
#ifndef MSW_RUNNER_H
#define MSW_RUNNER_H

#include "sweeper.h"
#include "array_operations.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/4_driver/runner.h:26-40 */
typedef struct {
  P normsq;
  P normsqdiff;
  double flops;
  double floprate;
  Timer time;
} Runner;

Runner Runner_null(void);
void Runner_create(Runner* runner);
void Runner_destroy(Runner* runner);

/* Runner_run_case (src/4_driver/runner.c:51-182): consumes --ncell_x/y/z,
 * --ne, --na, --niterations, plus this library's --nm (moments, default NM),
 * creates the problem, runs the ping-pong iterations with the state resident
 * in HBM, times them, and computes the FP64 norms on the device. */
void Runner_run_case(Runner* runner, Arguments* args, Env* env);

/* compare_runs (src/4_driver/runner.c:187-241): two runs must both give
 * normsqdiff == 0 and equal normsq, compared exactly. */
Bool_t compare_runs(const char* argstring1, const char* argstring2, Env* env);

#ifdef __cplusplus
}
#endif

#endif
