// sweep_mfma_inst.hip -- instantiates the tile sweep kernels (sweep_tile.h) of
// one precision and one moment count, every workgroup shape and angles per
// wave.  The Makefile compiles this file once per (precision, nm) with
// -DMSW_INST_PREC=F64|F32|BF16 -DMSW_INST_NM=4|8|16, so that the nine objects
// build in parallel.
#include "sweep_tile.h"

#define MSW_PREC_TYPE_F64 PrecF64
#define MSW_PREC_TYPE_F32 PrecF32
#define MSW_PREC_TYPE_BF16 PrecBF16
#define MSW_CAT2(a, b) a##b
#define MSW_CAT(a, b) MSW_CAT2(a, b)
#define MSW_PREC_T MSW_CAT(MSW_PREC_TYPE_, MSW_INST_PREC)

namespace msw {

void MSW_LAUNCH_NAME(MSW_INST_PREC, MSW_INST_NM)(const LaunchSpec& ls, const SweepParams& sp,
                                               hipStream_t st) {
  launch_nm<MSW_PREC_T, MSW_INST_NM>(ls, sp, st);
}

void MSW_LAUNCH_MULTI_NAME(MSW_INST_PREC, MSW_INST_NM)(int nac, int nwa, int ny, unsigned grid,
                                                     size_t lds, const SweepParams* probs,
                                                     const MultiLaunch& ml, hipStream_t st) {
  launch_multi_nm<MSW_PREC_T, MSW_INST_NM>(nac, nwa, ny, grid, lds, probs, ml, st);
}

}  // namespace msw
