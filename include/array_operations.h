/*
 * array_operations.h -- host state-array helpers, drop-in for the
 * reference's src/2_sweeper_base/array_operations.h:27-55
 * (array_operations.c:26-126).  The state layout is the reference's
 * v[iz][ie][iy][ix][iu][im].
 */
#ifndef MSW_ARRAY_OPERATIONS_H
#define MSW_ARRAY_OPERATIONS_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sweeper.h"

#ifdef __cplusplus
extern "C" {
#endif

/* vi <- Quantities_init_state of every element (array_operations.c:26-50). */
void initialize_state(P* v, const Dimensions dims, int nu, const Quantities* quan);
/* v <- 0 (array_operations.c:55-64). */
void initialize_state_zero(P* v, const Dimensions dims, int nu);
/* normsq = sum vo^2, normsqdiff = sum (vi - vo)^2 (array_operations.c:69-110). */
void get_state_norms(const P* vi, const P* vo, const Dimensions dims, int nu,
                     P* normsqp, P* normsqdiffp, Env* env);
/* vo <- vi, n elements (array_operations.c:115-126). */
void copy_vector(P* vo, const P* vi, size_t n);

#ifdef __cplusplus
}
#endif

#endif
