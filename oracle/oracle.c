/*
 * oracle.c -- CPU restatement of the minisweep sweep (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file; the product library never links it.  See oracle.h for the list
 * of reference functions restated here.
 *
 * Compiled with -ffp-contract=off: the reference is plain ANSI C built for
 * x86-64 without FMA, and its manufactured answer (vo == vi exactly) depends
 * on the separate roundings of the solve (quantities_testing_kernels.h:421-434).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { NOCT = 8 };

/* Octant bit b set => sweep runs downward along axis b
 * (src/2_sweeper_base/definitions_kernels.h:46-60). */
static int dir_of(int octant, int axis) { return (octant >> axis) & 1 ? -1 : +1; }

/* quantities_testing_kernels.h:52-69 */
static int sf_energy(int ie) {
  int r = (ie * 1366 + 150889) % 714025;
  return 1 << (r & 3);
}

/* quantities_testing_kernels.h:75-91 */
static int sf_unknown(int iu) {
  int r = (iu * 741 + 66037) % 312500;
  return 1 << (r & 3);
}

/* quantities_testing_kernels.h:97-126: only the last assignment survives, so
 * the factor is 2^((ix+3iy+7iz+2) & 3), evaluated in two's-complement int. */
static int sf_space(int ix, int iy, int iz) {
  int r = ix + 3 * iy + 7 * iz + 2;
  return 1 << (r & 3);
}

/* quantities_testing_kernels.h:132-141 */
static int sf_angle(int ia) { return 1 << (ia & 7); }

/* quantities_testing_kernels.h:203-214 */
static int sf_octant(int octant) { return 1 + octant; }

size_t oracle_size_state(OracleDims d) {
  return (size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.nm * d.nu;
}

double oracle_flops(OracleDims d, int niterations) {
  const double ss = (double)oracle_size_state(d);
  const double ssa = (double)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.na *
                     d.nu * NOCT;
  /* quantities_testing_c.h:356-359: 3 + 3*NDIM flops per solve */
  return niterations * (ss * NOCT * 2. * d.na + ssa * 12. + ss * NOCT * 2. * d.na);
}

/* state index (array_accessors_kernels.h:82-88), iz slowest */
static size_t ind_state(const OracleDims* d, int ix, int iy, int iz, int ie,
                        int im, int iu) {
  return (size_t)im + d->nm * ((size_t)iu + d->nu * ((size_t)ix + d->ncell_x *
         ((size_t)iy + d->ncell_y * ((size_t)ie + (size_t)d->ne * iz))));
}

/* Quantities_init_am_matrices_ (quantities_testing_c.h:40-201).  Same
 * integer construction; our storage strides by the runtime nm. */
void oracle_init_matrices(OracleDims d, double* A, double* M) {
  const int nm = d.nm, na = d.na;
  int oct, im, ia, i;
  memset(A, 0, sizeof(double) * NOCT * nm * na);
  memset(M, 0, sizeof(double) * NOCT * nm * na);
#define AFM(o, m, a) A[(a) + na * ((m) + nm * (o))]
#define MFA(o, m, a) M[(m) + nm * ((a) + na * (o))]
  for (oct = 0; oct < NOCT; ++oct)
    for (i = 0; i < na; ++i) {
      const int quot = (i + 1) / nm, rem = (i + 1) % nm;
      AFM(oct, nm - 1, i) += quot;
      if (rem != 0) {
        AFM(oct, 0, i) += -1.0;
        AFM(oct, rem, i) += 1.0;
      }
    }
  for (oct = 0; oct < NOCT; ++oct)
    for (im = 0; im < nm - 2; ++im)
      for (ia = 0; ia < na; ++ia) {
        const int rv = 21 + (im + nm * ia) % 17;
        AFM(oct, im + 0, ia) += -1.0 * rv;
        AFM(oct, im + 1, ia) += 2.0 * rv;
        AFM(oct, im + 2, ia) += -1.0 * rv;
      }
  for (oct = 0; oct < NOCT; ++oct)
    for (i = 0; i < nm; ++i) {
      const int quot = (i + 1) / na, rem = (i + 1) % na;
      MFA(oct, i, na - 1) += quot;
      if (rem != 0) {
        MFA(oct, i, 0) += -1.0;
        MFA(oct, i, rem) += 1.0;
      }
    }
  for (oct = 0; oct < NOCT; ++oct)
    for (im = 0; im < nm; ++im)
      for (ia = 0; ia < na - 2; ++ia) {
        const int rv = 37 + (im + nm * ia) % 19;
        MFA(oct, im, ia + 0) += -1.0 * rv;
        MFA(oct, im, ia + 1) += 2.0 * rv;
        MFA(oct, im, ia + 2) += -1.0 * rv;
      }
  for (oct = 0; oct < NOCT; ++oct)
    for (im = 0; im < nm; ++im)
      for (ia = 0; ia < na; ++ia) {
        MFA(oct, im, ia) /= NOCT;
        MFA(oct, im, ia) /= sf_angle(ia);
      }
#undef AFM
#undef MFA
}

/* Quantities_init_state (quantities_testing.h:60-90) looped as
 * initialize_state (array_operations.c:26-48). */
void oracle_init_state(OracleDims d, double* v) {
  int ix, iy, iz, ie, im, iu;
  for (iz = 0; iz < d.ncell_z; ++iz)
    for (iy = 0; iy < d.ncell_y; ++iy)
      for (ix = 0; ix < d.ncell_x; ++ix)
        for (ie = 0; ie < d.ne; ++ie)
          for (im = 0; im < d.nm; ++im)
            for (iu = 0; iu < d.nu; ++iu)
              v[ind_state(&d, ix, iy, iz, ie, im, iu)] =
                  ((double)(1 + im)) * (double)sf_space(ix, iy, iz) *
                  (double)sf_energy(ie) * (double)sf_unknown(iu);
}

/* Quantities_init_facexy/xz/yz (quantities_testing_kernels.h:219-332): one
 * formula, evaluated at the out-of-domain coordinate of the face. */
double oracle_face_value(OracleDims d, int ix_g, int iy_g, int iz_g, int ie,
                         int ia, int iu, int octant) {
  (void)d;
  return ((double)(1 + ia)) * (double)sf_angle(ia) *
         (double)sf_space(ix_g, iy_g, iz_g) * (double)sf_energy(ie) *
         (double)sf_unknown(iu) * (double)sf_octant(octant);
}

/* face layouts as ref_facexy/xz/yz with one octant and one ie:
 * fxy[iy][ix][iu][ia], fxz[iz][ix][iu][ia], fyz[iz][iy][iu][ia] */
#define FXY(x, y, a, u) fxy[(a) + na * ((u) + nu * ((x) + nx * (y)))]
#define FXZ(x, z, a, u) fxz[(a) + na * ((u) + nu * ((x) + nx * (z)))]
#define FYZ(y, z, a, u) fyz[(a) + na * ((u) + nu * ((y) + ny * (z)))]

/* Faces of one (octant, ie) set "one cell outside" the domain
 * (sweeper_simple_c.h:124-186). */
static void init_faces(const OracleDims* d, int octant, int ie, double* fxy, double* fxz,
                       double* fyz) {
  const int nx = d->ncell_x, ny = d->ncell_y, nz = d->ncell_z;
  const int na = d->na, nu = d->nu;
  const int dx = dir_of(octant, 0), dy = dir_of(octant, 1), dz = dir_of(octant, 2);
  int ix, iy, iz, ia, iu;
  iz = dz == +1 ? -1 : nz;
  for (iy = 0; iy < ny; ++iy)
    for (ix = 0; ix < nx; ++ix)
      for (iu = 0; iu < nu; ++iu)
        for (ia = 0; ia < na; ++ia)
          FXY(ix, iy, ia, iu) = oracle_face_value(*d, ix, iy, iz, ie, ia, iu, octant);
  iy = dy == +1 ? -1 : ny;
  for (iz = 0; iz < nz; ++iz)
    for (ix = 0; ix < nx; ++ix)
      for (iu = 0; iu < nu; ++iu)
        for (ia = 0; ia < na; ++ia)
          FXZ(ix, iz, ia, iu) = oracle_face_value(*d, ix, iy, iz, ie, ia, iu, octant);
  ix = dx == +1 ? -1 : nx;
  for (iz = 0; iz < nz; ++iz)
    for (iy = 0; iy < ny; ++iy)
      for (iu = 0; iu < nu; ++iu)
        for (ia = 0; ia < na; ++ia)
          FYZ(iy, iz, ia, iu) = oracle_face_value(*d, ix, iy, iz, ie, ia, iu, octant);
}

/* One (octant, ie) over the cells [x0, x1) x [y0, y1) x [z0, z1), upwind
 * first; the faces carry the values "one cell upwind" in and out, so a box
 * continues where the previous boxes of the octant left them.  Body follows
 * sweeper_simple_c.h:187-270 with the solve of
 * quantities_testing_kernels.h:380-446. */
static void sweep_box(const OracleDims* d, const double* A, const double* M,
                      const double* vi, double* vo_part, int octant, int ie,
                      double* fxy, double* fxz, double* fyz, double* vs, int x0, int x1,
                      int y0, int y1, int z0, int z1) {
  const int nx = d->ncell_x, ny = d->ncell_y;
  const int na = d->na, nm = d->nm, nu = d->nu;
  const int dx = dir_of(octant, 0), dy = dir_of(octant, 1), dz = dir_of(octant, 2);
  int ix, iy, iz, ia, iu, im;
  {
    const int ixb = dx == +1 ? x0 : x1 - 1, iyb = dy == +1 ? y0 : y1 - 1,
              izb = dz == +1 ? z0 : z1 - 1;
    const double so = (double)sf_octant(octant);
    const double so_r = 1.0 / so;
    int cz, cy, cx;
    for (cz = 0, iz = izb; cz < z1 - z0; ++cz, iz += dz)
      for (cy = 0, iy = iyb; cy < y1 - y0; ++cy, iy += dy)
        for (cx = 0, ix = ixb; cx < x1 - x0; ++cx, ix += dx) {
          const size_t base = ind_state(d, ix, iy, iz, ie, 0, 0);
          const double s = (double)sf_space(ix, iy, iz);
          const double s_r = 1.0 / s;
          const double sx_r = 1.0 / (double)sf_space(ix - dx, iy, iz);
          const double sy_r = 1.0 / (double)sf_space(ix, iy - dy, iz);
          const double sz_r = 1.0 / (double)sf_space(ix, iy, iz - dz);
          /* m2a: vs[iu][ia] = sum_im A[oct][im][ia] * vi[iu][im] (:219-232) */
          for (iu = 0; iu < nu; ++iu)
            for (ia = 0; ia < na; ++ia) {
              double r = 0.0;
              for (im = 0; im < nm; ++im)
                r += A[ia + na * (im + nm * octant)] * vi[base + im + (size_t)nm * iu];
              vs[ia + na * iu] = r;
            }
          /* solve (quantities_testing_kernels.h:421-443) */
          for (ia = 0; ia < na; ++ia) {
            const double wx = 1.0 / 2.0, wy = 1.0 / 4.0;
            const double wz = 1.0 / 4.0 - 1.0 / (double)sf_angle(ia);
            for (iu = 0; iu < nu; ++iu) {
              const double result =
                  (vs[ia + na * iu] * s_r +
                   (FXY(ix, iy, ia, iu) * wx * sz_r +
                    FXZ(ix, iz, ia, iu) * wy * sy_r +
                    FYZ(iy, iz, ia, iu) * wz * sx_r) * so_r) * s;
              const double scaled = result * so;
              vs[ia + na * iu] = result;
              FXY(ix, iy, ia, iu) = scaled;
              FXZ(ix, iz, ia, iu) = scaled;
              FYZ(iy, iz, ia, iu) = scaled;
            }
          }
          /* a2m: vo[iu][im] += sum_ia M[oct][ia][im] * vs[iu][ia] (:256-269) */
          for (iu = 0; iu < nu; ++iu)
            for (im = 0; im < nm; ++im) {
              double r = 0.0;
              for (ia = 0; ia < na; ++ia)
                r += M[im + nm * (ia + na * octant)] * vs[ia + na * iu];
              vo_part[base + im + (size_t)nm * iu] += r;
            }
        }
  }
}
#undef FXY
#undef FXZ
#undef FYZ

/* Per-(octant, ie) sweep over the whole grid; faces are private to the call. */
static void sweep_one(const OracleDims* d, const double* A, const double* M,
                      const double* vi, double* vo_part, int octant, int ie,
                      double* fxy, double* fxz, double* fyz, double* vs) {
  init_faces(d, octant, ie, fxy, fxz, fyz);
  sweep_box(d, A, M, vi, vo_part, octant, ie, fxy, fxz, fyz, vs, 0, d->ncell_x, 0,
            d->ncell_y, 0, d->ncell_z);
}

static void sweep_range_faces(OracleDims d, const double* A, const double* M,
                              const double* vi, double* vo, int ie_begin, int ie_end,
                              int nthreads, double* out_xy, double* out_xz, double* out_yz);

void oracle_sweep_range(OracleDims d, const double* A, const double* M,
                        const double* vi, double* vo, int ie_begin, int ie_end,
                        int nthreads) {
  sweep_range_faces(d, A, M, vi, vo, ie_begin, ie_end, nthreads, NULL, NULL, NULL);
}

/* The sweep, and each octant's outflow faces at the domain's downstream
 * boundary in the reference's Faces layouts with all eight octants
 * (array_accessors_kernels.h:481-486, 545-550, 609-614):
 *   out_xy[oct][ie][iy][ix][iu][ia], out_xz[oct][ie][iz][ix][iu][ia],
 *   out_yz[oct][ie][iz][iy][iu][ia].
 * After an octant's sweep the face arrays hold the last cell's value of each
 * line, scaled by (1+octant) as the reference stores it
 * (quantities_testing_kernels.h:437-443). */
void oracle_sweep_faces(OracleDims d, const double* A, const double* M, const double* vi,
                        double* vo, double* out_xy, double* out_xz, double* out_yz,
                        int nthreads) {
  sweep_range_faces(d, A, M, vi, vo, 0, d.ne, nthreads, out_xy, out_xz, out_yz);
}

static void sweep_range_faces(OracleDims d, const double* A, const double* M,
                              const double* vi, double* vo, int ie_begin, int ie_end,
                              int nthreads, double* out_xy, double* out_xz, double* out_yz) {
  const int nx = d.ncell_x, ny = d.ncell_y, nz = d.ncell_z;
  const size_t nfxy = (size_t)nx * ny * d.na * d.nu;
  const size_t nfxz = (size_t)nx * nz * d.na * d.nu;
  const size_t nfyz = (size_t)ny * nz * d.na * d.nu;
  int ie;
  /* vo = 0 (sweeper_simple_c.h:106) for the groups swept */
  {
    int iz;
    for (iz = 0; iz < nz; ++iz)
      for (ie = ie_begin; ie < ie_end; ++ie)
        memset(&vo[ind_state(&d, 0, 0, iz, ie, 0, 0)], 0,
               sizeof(double) * (size_t)nx * ny * d.nm * d.nu);
  }
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (ie = ie_begin; ie < ie_end; ++ie) {
    /* Energy groups are independent; octants are accumulated in the
     * reference order 0..7 (sweeper_simple_c.h:110), so the result does not
     * depend on the thread count. */
    double* fxy = (double*)malloc(sizeof(double) * nfxy);
    double* fxz = (double*)malloc(sizeof(double) * nfxz);
    double* fyz = (double*)malloc(sizeof(double) * nfyz);
    double* vs = (double*)malloc(sizeof(double) * (size_t)d.na * d.nu);
    int octant;
    for (octant = 0; octant < NOCT; ++octant) {
      sweep_one(&d, A, M, vi, vo, octant, ie, fxy, fxz, fyz, vs);
      if (out_xy) {
        const size_t k = (size_t)octant * d.ne + ie;
        memcpy(out_xy + k * nfxy, fxy, sizeof(double) * nfxy);
        memcpy(out_xz + k * nfxz, fxz, sizeof(double) * nfxz);
        memcpy(out_yz + k * nfyz, fyz, sizeof(double) * nfyz);
      }
    }
    free(fxy);
    free(fxz);
    free(fyz);
    free(vs);
  }
  (void)nthreads;
}

void oracle_sweep(OracleDims d, const double* A, const double* M,
                  const double* vi, double* vo, int nthreads) {
  oracle_sweep_range(d, A, M, vi, vo, 0, d.ne, nthreads);
}

/* The sweep of oracle_sweep evaluated in long double (x87 extended: 64-bit
 * significand), every value rounded to double only at the end.  Same loop
 * order, boundary faces and solve (sweeper_simple_c.h:110-270,
 * quantities_testing_kernels.h:380-446); its rounding is 2^-11 of FP64's,
 * so it stands for the exact sweep when judging how far an FP64 evaluation
 * is from it on ill-conditioned inputs. */
typedef long double LD;
void oracle_sweep_ld(OracleDims d, const double* A, const double* M, const double* vi,
                     double* vo, int nthreads) {
  const int nx = d.ncell_x, ny = d.ncell_y, nz = d.ncell_z, na = d.na, nm = d.nm, nu = d.nu;
  const size_t nfxy = (size_t)nx * ny * na * nu, nfxz = (size_t)nx * nz * na * nu;
  const size_t nfyz = (size_t)ny * nz * na * nu;
  const size_t group = (size_t)nx * ny * nm * nu;
  int ie;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (ie = 0; ie < d.ne; ++ie) {
    LD* fxy = (LD*)malloc(sizeof(LD) * nfxy);
    LD* fxz = (LD*)malloc(sizeof(LD) * nfxz);
    LD* fyz = (LD*)malloc(sizeof(LD) * nfyz);
    LD* vs = (LD*)malloc(sizeof(LD) * (size_t)na * nu);
    LD* acc = (LD*)calloc(group * nz, sizeof(LD)); /* this group's vo, [iz][iy][ix][iu][im] */
    int octant, ix, iy, iz, ia, iu, im;
    for (octant = 0; octant < NOCT; ++octant) {
      const int dx = dir_of(octant, 0), dy = dir_of(octant, 1), dz = dir_of(octant, 2);
      const LD so = (LD)sf_octant(octant), so_r = 1.0L / so;
      /* boundary faces one cell outside the domain (sweeper_simple_c.h:124-186) */
      iz = dz == +1 ? -1 : nz;
      for (iy = 0; iy < ny; ++iy)
        for (ix = 0; ix < nx; ++ix)
          for (iu = 0; iu < nu; ++iu)
            for (ia = 0; ia < na; ++ia)
              fxy[ia + na * (iu + nu * (ix + nx * iy))] =
                  (LD)oracle_face_value(d, ix, iy, iz, ie, ia, iu, octant);
      iy = dy == +1 ? -1 : ny;
      for (iz = 0; iz < nz; ++iz)
        for (ix = 0; ix < nx; ++ix)
          for (iu = 0; iu < nu; ++iu)
            for (ia = 0; ia < na; ++ia)
              fxz[ia + na * (iu + nu * (ix + nx * iz))] =
                  (LD)oracle_face_value(d, ix, iy, iz, ie, ia, iu, octant);
      ix = dx == +1 ? -1 : nx;
      for (iz = 0; iz < nz; ++iz)
        for (iy = 0; iy < ny; ++iy)
          for (iu = 0; iu < nu; ++iu)
            for (ia = 0; ia < na; ++ia)
              fyz[ia + na * (iu + nu * (iy + ny * iz))] =
                  (LD)oracle_face_value(d, ix, iy, iz, ie, ia, iu, octant);
      int cx, cy, cz;
      for (cz = 0, iz = dz == +1 ? 0 : nz - 1; cz < nz; ++cz, iz += dz)
        for (cy = 0, iy = dy == +1 ? 0 : ny - 1; cy < ny; ++cy, iy += dy)
          for (cx = 0, ix = dx == +1 ? 0 : nx - 1; cx < nx; ++cx, ix += dx) {
            const size_t base = ind_state(&d, ix, iy, iz, ie, 0, 0);
            LD* out = acc + (size_t)iz * group + ((size_t)iy * nx + ix) * nu * nm;
            const LD s = (LD)sf_space(ix, iy, iz), s_r = 1.0L / s;
            const LD sx_r = 1.0L / (LD)sf_space(ix - dx, iy, iz);
            const LD sy_r = 1.0L / (LD)sf_space(ix, iy - dy, iz);
            const LD sz_r = 1.0L / (LD)sf_space(ix, iy, iz - dz);
            for (iu = 0; iu < nu; ++iu)
              for (ia = 0; ia < na; ++ia) {
                LD r = 0.0L;
                for (im = 0; im < nm; ++im)
                  r += (LD)A[ia + na * (im + nm * octant)] * (LD)vi[base + im + (size_t)nm * iu];
                vs[ia + na * iu] = r;
              }
            for (ia = 0; ia < na; ++ia) {
              const LD wz = 0.25L - 1.0L / (LD)sf_angle(ia);
              for (iu = 0; iu < nu; ++iu) {
                LD* f1 = &fxy[ia + na * (iu + nu * (ix + nx * iy))];
                LD* f2 = &fxz[ia + na * (iu + nu * (ix + nx * iz))];
                LD* f3 = &fyz[ia + na * (iu + nu * (iy + ny * iz))];
                const LD result = (vs[ia + na * iu] * s_r +
                                   (*f1 * 0.5L * sz_r + *f2 * 0.25L * sy_r + *f3 * wz * sx_r) *
                                       so_r) * s;
                vs[ia + na * iu] = result;
                *f1 = *f2 = *f3 = result * so;
              }
            }
            for (iu = 0; iu < nu; ++iu)
              for (im = 0; im < nm; ++im) {
                LD r = 0.0L;
                for (ia = 0; ia < na; ++ia)
                  r += (LD)M[im + nm * (ia + na * octant)] * vs[ia + na * iu];
                out[im + (size_t)nm * iu] += r;
              }
          }
    }
    for (iz = 0; iz < nz; ++iz) {
      const LD* src = acc + (size_t)iz * group;
      double* dst = &vo[ind_state(&d, 0, 0, iz, ie, 0, 0)];
      size_t k;
      for (k = 0; k < group; ++k) dst[k] = (double)src[k];
    }
    free(fxy);
    free(fxz);
    free(fyz);
    free(vs);
    free(acc);
  }
  (void)nthreads;
}

/* The reference's tiled-octant ordering (sweeper_tileoctants_c.h:127-189):
 * 8 tile steps; at step t octant o sweeps the half-domain box whose half
 * along axis b is the lower one when (bit b of t is 0) equals (o runs upward
 * along b), so the 8 octants of a step work on disjoint boxes and each
 * octant visits its boxes upwind first.  Each octant keeps its faces between
 * steps.  The cells receive the octants' contributions in another order than
 * oracle_sweep's 0..7, so the two agree to rounding. */
void oracle_sweep_tileoctants(OracleDims d, const double* A, const double* M,
                              const double* vi, double* vo, int nthreads) {
  const int nx = d.ncell_x, ny = d.ncell_y, nz = d.ncell_z;
  const size_t nfxy = (size_t)nx * ny * d.na * d.nu;
  const size_t nfxz = (size_t)nx * nz * d.na * d.nu;
  const size_t nfyz = (size_t)ny * nz * d.na * d.nu;
  const size_t nf = nfxy + nfxz + nfyz;
  int ie;
  memset(vo, 0, sizeof(double) * oracle_size_state(d));
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (ie = 0; ie < d.ne; ++ie) {
    double* faces = (double*)malloc(sizeof(double) * nf * NOCT);
    double* vs = (double*)malloc(sizeof(double) * (size_t)d.na * d.nu);
    int step, octant;
    for (octant = 0; octant < NOCT; ++octant) {
      double* f = faces + nf * octant;
      init_faces(&d, octant, ie, f, f + nfxy, f + nfxy + nfxz);
    }
    for (step = 0; step < NOCT; ++step)
      for (octant = 0; octant < NOCT; ++octant) {
        double* f = faces + nf * octant;
        int lo[3], hi[3], ext[3] = {nx, ny, nz}, b;
        for (b = 0; b < 3; ++b) {
          const int lower = (((step >> b) & 1) == 0) == (dir_of(octant, b) == +1);
          lo[b] = lower ? 0 : ext[b] / 2;
          hi[b] = lower ? ext[b] / 2 : ext[b];
        }
        sweep_box(&d, A, M, vi, vo, octant, ie, f, f + nfxy, f + nfxy + nfxz, vs, lo[0], hi[0],
                  lo[1], hi[1], lo[2], hi[2]);
      }
    free(faces);
    free(vs);
  }
  (void)nthreads;
}

/* get_state_norms (array_operations.c:69-110), same loop order */
void oracle_norms(OracleDims d, const double* vi, const double* vo,
                  double* normsq, double* normsqdiff) {
  int ix, iy, iz, ie, im, iu;
  double n = 0.0, nd = 0.0;
  for (iz = 0; iz < d.ncell_z; ++iz)
    for (iy = 0; iy < d.ncell_y; ++iy)
      for (ix = 0; ix < d.ncell_x; ++ix)
        for (ie = 0; ie < d.ne; ++ie)
          for (im = 0; im < d.nm; ++im)
            for (iu = 0; iu < d.nu; ++iu) {
              const size_t i = ind_state(&d, ix, iy, iz, ie, im, iu);
              const double df = vi[i] - vo[i];
              n += vo[i] * vo[i];
              nd += df * df;
            }
  *normsq = n;
  *normsqdiff = nd;
}
