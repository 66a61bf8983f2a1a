// msw_internal.h -- host-side declarations shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "msw_device.h"
#include "../../include/msw_plan.h"

namespace msw {

// One kernel launch of a sweep: which items, grid and block shape.
struct LaunchSpec {
  int item_base;     // first TileItem of this launch
  unsigned grid;     // items * ne * nu workgroups
  int nwa;           // waves per workgroup (angle chunks)
  int nac;           // angles per wave
  int nm;            // moments
  int epw;           // energy groups streamed through one workgroup (MFMA)
  int ny;            // row groups of 4 rows per tile (MFMA)
  size_t lds_bytes;  // dynamic LDS per workgroup
};

// Sweep with m2a/a2m on MFMA (sweep_mfma.hip), in the plan's precision
// (MSW_PREC_*).  nac is 16, 32 or 64; nm 4, 8 or 16; (nwa, ny) a shape for
// which mfma_shape_supported is true.
void launch_mfma(int precision, const LaunchSpec& ls, const SweepParams& sp, hipStream_t st);
bool mfma_shape_supported(int nwa, int ny);
size_t mfma_lds_bytes(int precision, int nac, int nwa, int ny);
// Writes the domain-boundary inflow into the MFMA kernels' face planes; runs
// before the launches of every sweep.
void launch_mfma_fill(int precision, const SweepParams& sp, hipStream_t st);
// Several sweeps' launches in one kernel launch (one kernel variant: all
// segments' problems share nm, nac, nwa and the precision).  probs is a
// device array of the problems' parameters.
void launch_mfma_multi(int precision, int nm, int nac, int nwa, int ny, unsigned grid,
                       const SweepParams* probs, const MultiLaunch& ml, hipStream_t st);
void launch_mfma_fill_multi(int precision, unsigned grid, const SweepParams* probs,
                            const MultiLaunch& ml, hipStream_t st);
// Workgroups of one sweep's boundary fill.
int mfma_fill_rows(const SweepParams& sp);
// Bytes of one face value of the MFMA kernels in a precision.
size_t mfma_face_elem_bytes(int precision);
// MFMA tile: mfma_tile_x() cells along x, ny row groups of mfma_tile_rows()
// rows along y.
// The sweep kernel addresses one (octant, unknown) face plane with 32-bit
// offsets: max(ncell_x, ncell_y) * ne * ncell_z * na_pad < 2^31.
bool mfma_face_offsets_fit(const MswDims& d, int na_pad);
int mfma_tile_x();
int mfma_tile_rows();

// Sets the calling thread's msw_last_error() text; returns code.
int set_error(int code, const std::string& msg);
// Default angles per wave of the MFMA kernels for a precision and na.
int default_mfma_nac(int precision, int na);
// Energy groups per workgroup of the MFMA kernels: about 128 stream steps
// per workgroup, then fewer while a launch of `waves1` waves at one group per
// workgroup would drop below 4 waves per SIMD.
int default_mfma_epw(int nz, int ne);
int launch_mfma_epw(int preferred, int ne, long long waves1, int precision);
// Default row groups per MFMA tile.
int default_mfma_ny(int precision, int nwa, int ncell_y);

// Host list schedule of (octant, tile) visits for ntx x nty tiles (kind 0:
// most-remaining octant first; 1: the reference's octant order).
// Launches of (octant, tile) visits: every octant in oct_mask visits every
// tile once, upwind tiles first; one visit per physical tile per launch.
// Flags mark a tile's first visit (touched count 0) and last (count 7), the
// counts starting at touched0 (visits made by earlier passes).
std::vector<std::vector<TileItem>> build_schedule(int ntx, int nty, int kind,
                                                  unsigned oct_mask = 0xFF, int touched0 = 0);

// Small kernels of sweep_plan.hip.
void launch_pack(const double* afm, const double* mfa, double* at, double* mf, int nm, int na,
                 int na_pad, hipStream_t st);
void launch_init_state(const MswDims& d, double* v, hipStream_t st);
size_t norms_scratch_doubles();
// dst <- BF16 bits of (float)src, 4 values per thread; n a multiple of 4.
void launch_to_bf16(const double* src, unsigned short* dst, size_t n, hipStream_t st);
void launch_norms(const double* vi, const double* vo, size_t n, double* part, double* out2,
                  hipStream_t st);

}  // namespace msw
