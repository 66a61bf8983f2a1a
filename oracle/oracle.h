/*
 * oracle.h -- CPU restatement of the minisweep sweep algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP sweep:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (libminisweep.so) never links or calls it.
 *
 * Every function restates the reference algorithm from its sources
 * (citations are relative to the reference tree):
 *   - scale factors / boundary faces / state init:
 *       src/3_sweeper/quantities_testing_kernels.h:52-332,
 *       src/3_sweeper/quantities_testing.h:60-90
 *   - a_from_m / m_from_a construction: src/3_sweeper/quantities_testing_c.h:40-201
 *   - the sweep (octant -> face init -> ie -> upwind cells -> m2a -> solve -> a2m):
 *       src/3_sweeper/sweeper_simple_c.h:83-277
 *   - the cell solve: src/3_sweeper/quantities_testing_kernels.h:337-447
 *   - norms: src/2_sweeper_base/array_operations.c:69-110
 *
 * Layouts (src/2_sweeper_base/array_accessors_kernels.h):
 *   state     v[iz][ie][iy][ix][iu][im]              (:82-88)
 *   a_from_m  A[octant][im][ia]   (stride nm)        (:346-349)
 *   m_from_a  M[octant][ia][im]   (stride nm)        (:410-413)
 *
 * Pinning: see tests/test_oracle.py -- closed-form normsq of the manufactured
 * problem, vo == vi exactly, golden vectors from the compiled reference on
 * non-affine inputs, brute-force dense-operator checks on tiny grids.
 */
#ifndef MSW_ORACLE_H
#define MSW_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int ncell_x, ncell_y, ncell_z, ne, nm, na, nu;
} OracleDims;

size_t oracle_size_state(OracleDims d);

/* a_from_m and m_from_a, each 8*nm*na doubles. */
void oracle_init_matrices(OracleDims d, double* a_from_m, double* m_from_a);

/* Manufactured input state (Quantities_init_state). */
void oracle_init_state(OracleDims d, double* v);

/* One sweep vo <- S(vi).  nthreads<=0 means all OpenMP threads. */
void oracle_sweep(OracleDims d, const double* a_from_m, const double* m_from_a,
                  const double* vi, double* vo, int nthreads);

/* The same sweep in the reference's tiled-octant order
 * (sweeper_tileoctants_c.h:127-189): a second CPU ordering, equal to
 * oracle_sweep up to rounding. */
void oracle_sweep_tileoctants(OracleDims d, const double* a_from_m, const double* m_from_a,
                              const double* vi, double* vo, int nthreads);

/* The same sweep evaluated in long double (64-bit significand), rounded to
 * double at the end: the reference point for ill-conditioned inputs, where
 * any FP64 evaluation (the reference's included) is 1e-12-level off. */
void oracle_sweep_ld(OracleDims d, const double* a_from_m, const double* m_from_a,
                     const double* vi, double* vo, int nthreads);

/* The sweep and each octant's outflow faces at the domain's downstream
 * boundary, in the reference's Faces layouts with all eight octants
 * (array_accessors_kernels.h:481-486, 545-550, 609-614):
 * out_xy[8][ne][ny][nx][nu][na], out_xz[8][ne][nz][nx][nu][na],
 * out_yz[8][ne][nz][ny][nu][na]; values scaled by (1+octant) as the
 * reference stores them (quantities_testing_kernels.h:437-443). */
void oracle_sweep_faces(OracleDims d, const double* a_from_m, const double* m_from_a,
                        const double* vi, double* vo, double* out_xy, double* out_xz,
                        double* out_yz, int nthreads);

/* Same sweep restricted to the energy groups [ie_begin, ie_end): used to time
 * a bounded sample of a large problem. */
void oracle_sweep_range(OracleDims d, const double* a_from_m,
                        const double* m_from_a, const double* vi, double* vo,
                        int ie_begin, int ie_end, int nthreads);

/* normsq = sum vo^2, normsqdiff = sum (vi-vo)^2, summed in the reference's
 * loop order (sequential). */
void oracle_norms(OracleDims d, const double* vi, const double* vo,
                  double* normsq, double* normsqdiff);

/* Boundary face value (Quantities_init_facexy/xz/yz share one formula). */
double oracle_face_value(OracleDims d, int ix_g, int iy_g, int iz_g, int ie,
                         int ia, int iu, int octant);

/* Flop count of niterations sweeps (src/4_driver/runner.c:162-166). */
double oracle_flops(OracleDims d, int niterations);

#ifdef __cplusplus
}
#endif

#endif
