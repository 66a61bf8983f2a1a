// sweep_mfma_bf16.hip -- the tile sweep kernels (sweep_tile.h) of the mixed BF16
// form, every workgroup shape, moment count and angles per wave; one file per
// precision so that the three build in parallel.
#include "sweep_tile.h"

namespace msw {

void launch_mfma_bf16(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  launch_prec<PrecBF16>(ls, sp, st);
}

void launch_mfma_multi_bf16(int nm, int nac, int nwa, int ny, unsigned grid, size_t lds,
                          const SweepParams* probs, const MultiLaunch& ml, hipStream_t st) {
  launch_multi_prec<PrecBF16>(nm, nac, nwa, ny, grid, lds, probs, ml, st);
}

}  // namespace msw
