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

// FP64 sweep with m2a/a2m on MFMA (sweep_mfma.hip).  nac must be 16 or 32.
void launch_mfma_f64(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st);
size_t mfma_f64_lds_bytes(int nac, int nwa);
// Tile of one wave of the MFMA kernels (cells along x and y).
int mfma_tile_x();
int mfma_tile_y();

}  // namespace msw
