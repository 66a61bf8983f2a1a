/*
 * env.h -- reference header name for callers written against the
 * reference (it replaces src/1_base/env.h).  The declarations it stands for,
 * Env_*, Insist and the base types (declared in sweeper.h), live in sweeper.h; this file only forwards there, so that the
 * reference's unchanged drivers (src/4_driver/sweep.c:13-22, runner.c:13-22,
 * tester.c:13-22) compile against this library's include directory.
 */
#ifndef MSW_FWD_ENV_H
#define MSW_FWD_ENV_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sweeper.h"

#endif
