// msw_device.h -- device-side helpers shared by the HIP sweep kernels.
//
// The manufactured-problem formulas follow
// src/3_sweeper/quantities_testing_kernels.h:52-214 of the reference; the
// boundary-face value follows :219-332 and the solve :380-446.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msw {

constexpr int kNoct = 8;
constexpr int kTile = 8;      // xy tile edge: one wave64 = 8x8 cells
constexpr int kSkew = 2 * (kTile - 1);

// One (octant, tile) visit in a launch.  Tile coordinates are in the
// octant's own frame (upwind corner at 0,0).  flags: kVisitFirst, this visit
// is the first of the sweep to reach the tile, so it writes vo instead of
// adding; kVisitLast, no later visit reaches it (the mixed forms then write
// the FP64 vo from their FP32 running sum).
constexpr int kVisitFirst = 1;
constexpr int kVisitLast = 2;
struct TileItem {
  int octant;
  int flags;
  int tx;
  int ty;
};

__host__ __device__ __forceinline__ int sf_space_exp(int x, int y, int z) {
  return (x + 3 * y + 7 * z + 2) & 3;
}
__host__ __device__ __forceinline__ int sf_energy_exp(int ie) {
  return ((ie * 1366 + 150889) % 714025) & 3;
}
__host__ __device__ __forceinline__ int sf_unknown_exp(int iu) {
  return ((iu * 741 + 66037) % 312500) & 3;
}
__host__ __device__ __forceinline__ int sf_energy(int ie) { return 1 << sf_energy_exp(ie); }
__host__ __device__ __forceinline__ int sf_unknown(int iu) { return 1 << sf_unknown_exp(iu); }

// The cell solve.  Contraction is off: the reference rounds each product and
// sum separately, and its exact manufactured answer depends on that.
//   cz = xfluxweight * Sz^-1, cy = yfluxweight * Sy^-1, cx = zfluxweight(ia) * Sx^-1
// are exact power-of-two rescalings of the reference's coefficients, so the
// roundings are the reference's ones.
__device__ __forceinline__ double solve_fp64(double vs, double fz, double fy,
                                             double fx, double s_r, double cz,
                                             double cy, double cx, double oct_r,
                                             double s) {
#pragma clang fp contract(off)
  return (vs * s_r + ((fz * cz + fy * cy) + fx * cx) * oct_r) * s;
}

__device__ __forceinline__ float solve_fp32(float vs, float fz, float fy,
                                            float fx, float s_r, float cz,
                                            float cy, float cx, float oct_r,
                                            float s) {
#pragma clang fp contract(off)
  return (vs * s_r + ((fz * cz + fy * cy) + fx * cx) * oct_r) * s;
}

// Read-only data that is uniform across a wave is loaded through the
// constant address space, which the backend turns into scalar loads.
typedef const __attribute__((address_space(4))) double* cdptr;

// Arguments of the sweep kernels (all launches of one sweep).
struct SweepParams {
  const double* vi;
  double* vo;
  const double* at;  // a_from_m transposed: [8][na][nm]
  const double* mf;  // m_from_a:            [8][na][nm]
  void* fyz;         // x-direction faces: VALU [8][ne][nu][ny][nz][na] (grid frame),
                     //   MFMA [8][nu][ny][ne][nz][na] (octant-frame z)
  void* fxz;         // y-direction faces: the same with nx for ny
  float* vo32;       // mixed forms: FP32 running sum of vo over the octants (state layout)
  // Operand copies of vi made once per sweep by the mixed forms, state
  // layout, each value divided by its cell's scale factor K = S E U (spatial,
  // group and unknown factors; a power of two, so the division is exact):
  //   FP32 form: vhi = the FP32 rounding, vlo = the BF16 rounding of the rest
  //   BF16 form: vib = the BF16 rounding (through FP32)
  const float* vhi;
  const unsigned short* vlo;
  const unsigned short* vib;
  const TileItem* items;
  int nx, ny, nz, ne, na, nu;
  int na_pad;  // na rounded up to whole angle chunks (matrix/face stride)
  int ntx, nty;
  int e_off = 0;  // global index of group 0 (scale factors use e_off + ie)
  // z blocks (MFMA engine, FP64): the sweep runs one block of nz planes at a
  // time, vi/vo pointing at the block.  z_off is the block's first plane in
  // the grid and nz_grid the grid's extent (0: not blocked, nz is the grid).
  // fz holds the z faces between blocks, [4][nu][ne][ny][nx][na_pad] indexed
  // by octant & 3, angles in lane order; z_in: this block takes its z inflow
  // from fz (it is not the first block in the octant's direction), z_out: it
  // writes its last plane's outflow there (not the last block).
  int z_off = 0, nz_grid = 0;
  void* fz = nullptr;
  int z_in = 0, z_out = 0;
  // Outflow faces at the domain's downstream boundary, written in the
  // reference's Faces layouts with all eight octants when not null
  // (msw_plan_sweep_faces; whole-grid sweeps): oxy[8][ne][ny][nx][nu][na],
  // oxz[8][ne][nz][nx][nu][na], oyz[8][ne][nz][ny][nu][na], FP64, scaled by
  // (1+octant) as the reference stores them.
  double* oxy = nullptr;
  double* oxz = nullptr;
  double* oyz = nullptr;
};

// A launch that runs segments of several sweeps: segment k covers
// workgroups [start[k], start[k+1]) (the last one up to the grid size) and is
// launch-local work of problem prob[k], whose items start at item_base[k].
// epw_ng[k] packs the energy groups streamed through one workgroup (low
// kEpwBits bits), a flag (bit kEpwBits) and the segment's (item, group chunk)
// count above it.  With the flag the segment's workgroups are that count
// padded to whole rounds of the XCDs (xcd_round_ids) and mapped to them as
// a single sweep's are; without it they are count * nu, in id order.
// Passed by value as a kernel argument.
constexpr int kMaxSegments = 128;
constexpr int kEpwBits = 6;
constexpr int kEpwMask = (1 << kEpwBits) - 1;
constexpr int kXcdFlag = 1 << kEpwBits;
struct MultiLaunch {
  int nseg;
  int start[kMaxSegments];
  int prob[kMaxSegments];
  int item_base[kMaxSegments];
  int epw_ng[kMaxSegments];
};

// Workgroups of a launch of ngroups (item, group chunk) pairs of nu unknowns
// each, padded so that all unknowns of a pair share one of the 8 XCDs.
__host__ __device__ __forceinline__ long long xcd_round_ids(long long ngroups, int nu) {
  return (ngroups + 7) / 8 * 8 * nu;
}

}  // namespace msw
