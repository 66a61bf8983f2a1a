// msw_internal.h -- host-side declarations shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "msw_device.h"

namespace msw {

// One kernel launch of a sweep: which items, grid and block shape.
struct LaunchSpec {
  int item_base;     // first TileItem of this launch
  unsigned grid;     // items * ne * nu workgroups
  int nwa;           // waves per workgroup (angle chunks)
  int nac;           // angles per wave
  int nm;            // moments
  size_t lds_bytes;  // dynamic LDS per workgroup
};

// Sweep with m2a/a2m on MFMA (sweep_mfma.hip), in the plan's precision
// (MSW_PREC_*).  nac is 16, 32 or 64; nwa 1..4; nm 4, 8 or 16.
void launch_mfma(int precision, const LaunchSpec& ls, const SweepParams& sp, hipStream_t st);
size_t mfma_lds_bytes(int precision, int nac, int nwa);
// Writes the domain-boundary inflow into the MFMA kernels' face planes; runs
// before the launches of every sweep.
void launch_mfma_fill(int precision, const SweepParams& sp, hipStream_t st);
// Bytes of one face value of the MFMA kernels in a precision.
size_t mfma_face_elem_bytes(int precision);
// Tile of one wave of the MFMA kernels (cells along x and y).
int mfma_tile_x();
int mfma_tile_y();

}  // namespace msw
