// sweep_mfma_f32.hip -- the tile sweep kernels (sweep_tile.h) of the mixed FP32
// form, every workgroup shape, moment count and angles per wave; one file per
// precision so that the three build in parallel.
#include "sweep_tile.h"

namespace msw {

void launch_mfma_f32(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  launch_prec<PrecF32>(ls, sp, st);
}

void launch_mfma_multi_f32(int nm, int nac, int nwa, int ny, unsigned grid, size_t lds,
                          const SweepParams* probs, const MultiLaunch& ml, hipStream_t st) {
  launch_multi_prec<PrecF32>(nm, nac, nwa, ny, grid, lds, probs, ml, st);
}

}  // namespace msw
