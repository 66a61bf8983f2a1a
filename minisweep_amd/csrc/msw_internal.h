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
  int ngroups;       // (item, group chunk) pairs (MFMA: grid = xcd_round_ids(ngroups, nu))
};

// Sweep with m2a/a2m on MFMA (sweep_mfma.hip), in the plan's precision
// (MSW_PREC_*).  nac is 16, 32 or 64; nm 4, 8 or 16; (nwa, ny) a shape for
// which mfma_shape_supported is true.  The mixed forms read the operand
// copies of vi (launch_operand_copy) made for the sweep.
void launch_mfma(int precision, const LaunchSpec& ls, const SweepParams& sp, hipStream_t st);
bool mfma_shape_supported(int nwa, int ny);
size_t mfma_lds_bytes(int precision, int nac, int nwa, int ny);
// Several sweeps' launches in one kernel launch (one kernel variant: all
// segments' problems share nm, nac, nwa and the precision).  probs is a
// device array of the problems' parameters.
void launch_mfma_multi(int precision, int nm, int nac, int nwa, int ny, unsigned grid,
                       const SweepParams* probs, const MultiLaunch& ml, hipStream_t st);
// The same per (precision, nm): sweep_mfma_inst.hip compiled once for each.
#define MSW_LAUNCH_NAME_(P, NM) launch_mfma_##P##_nm##NM
#define MSW_LAUNCH_NAME(P, NM) MSW_LAUNCH_NAME_(P, NM)
#define MSW_LAUNCH_MULTI_NAME_(P, NM) launch_mfma_multi_##P##_nm##NM
#define MSW_LAUNCH_MULTI_NAME(P, NM) MSW_LAUNCH_MULTI_NAME_(P, NM)
#define MSW_DECLARE_LAUNCH(P, NM)                                                          \
  void MSW_LAUNCH_NAME(P, NM)(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st); \
  void MSW_LAUNCH_MULTI_NAME(P, NM)(int nac, int nwa, int ny, unsigned grid, size_t lds,   \
                                    const SweepParams* probs, const MultiLaunch& ml,       \
                                    hipStream_t st);
MSW_DECLARE_LAUNCH(F64, 4)
MSW_DECLARE_LAUNCH(F64, 8)
MSW_DECLARE_LAUNCH(F64, 16)
MSW_DECLARE_LAUNCH(F32, 4)
MSW_DECLARE_LAUNCH(F32, 8)
MSW_DECLARE_LAUNCH(F32, 16)
MSW_DECLARE_LAUNCH(BF16, 4)
MSW_DECLARE_LAUNCH(BF16, 8)
MSW_DECLARE_LAUNCH(BF16, 16)
#undef MSW_DECLARE_LAUNCH
// The mixed forms' m2a operand copies of vi (each value divided by its
// cell's scale K = S E U): FP32 form hi (FP32) and lo (BF16 of the rest), BF16
// form lo = the BF16 value (hi unused); e_off is the global index of the
// state's group 0.  Bytes of device memory they take for a state of n values
// (0 in FP64).
size_t operand_copy_bytes(int precision, size_t n);
void launch_operand_copy(int precision, const MswDims& d, const double* vi, float* hi,
                         unsigned short* lo, int e_off, hipStream_t st);
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
// Most energy groups one tile pipeline streams (the kernel keeps the groups'
// scale exponents in one 64-bit word).
constexpr int kMaxStreamGroups = 32;
// Default angles per wave of the MFMA kernels for a precision and problem;
// alone: the problem's launches run by themselves (a plan), not combined with
// other problems' (batch and server ticks, which keep one choice per class so
// that a request's bits do not depend on what else is in flight).
int default_mfma_nac(int precision, const MswDims& d, bool alone);
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
void launch_norms(const double* vi, const double* vo, size_t n, double* part, double* out2,
                  hipStream_t st);

}  // namespace msw
