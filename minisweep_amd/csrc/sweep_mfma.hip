// sweep_mfma.hip -- dispatch of the tile sweep kernels (sweep_tile.h) by
// precision, workgroup shapes, and the mixed forms' operand copies of vi.
// The kernels themselves are instantiated by sweep_mfma_inst.hip, compiled once
// per (precision, nm).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "msw_device.h"
#include "msw_internal.h"
#include "../../include/msw_plan.h"

namespace msw {

namespace {
// The mixed forms' m2a operand, made once per sweep: each value of vi
// divided by its cell's scale K = 2^((x + 3y + 7z + 2) & 3) E(group) U(unknown)
// (a power of two: exact), then
//   split = 1: hi = its FP32 rounding, lo = the BF16 rounding of what FP32
//              leaves (together vi to about 2^-33 relative);
//   split = 0: bf = its BF16 rounding through FP32 (the two roundings the
//              BF16 form's contraction applies to its operand).
// A thread handles runs of 4 consecutive values: 4 moments of one (cell,
// unknown), nm being a power of two from 4 (2^nm_log2).  Block b copies part
// `b % chunks` of the (z, group) plane b / chunks, so that the plane's z and
// group (and the group's scale) are fixed per block and a run needs only its
// cell's index within the plane.  I: the in-plane index type, 32 bits
// unless a plane holds 2^32 values or more.
template <class I>
__global__ __launch_bounds__(256) void k_operand_copy(const double* __restrict__ vi,
                                                      float* __restrict__ hi,
                                                      unsigned short* __restrict__ lo, int nx,
                                                      int ny, int ne, int nm_log2, int nu,
                                                      int chunks, int e_off) {
  const int plane = (int)blockIdx.x / chunks;  // z * ne + group
  const int part = (int)blockIdx.x - plane * chunks;
  const int ie = plane % ne, z = plane / ne;
  const I cellv = (I)nu << nm_log2;            // values of one cell
  const I nvals = (I)ny * (I)nx * cellv;       // values of the plane
  const int ez = sf_energy_exp(e_off + ie);
  const size_t base = (size_t)plane * nvals;
  for (I off = ((I)part * blockDim.x + threadIdx.x) * 4; off < nvals;
       off += (I)chunks * blockDim.x * 4) {
    const I cell = off / cellv;                // y * nx + x
    const int u = (int)((off >> nm_log2) - cell * (I)nu);
    const int y = (int)(cell / (I)nx);
    const int x = (int)(cell - (I)y * (I)nx);
    const size_t i = base + off;
    const int e = -(sf_space_exp(x, y, z) + ez + sf_unknown_exp(u));
    const double2 a = *reinterpret_cast<const double2*>(vi + i);
    const double2 b = *reinterpret_cast<const double2*>(vi + i + 2);
    const double v[4] = {__builtin_ldexp(a.x, e), __builtin_ldexp(a.y, e),
                         __builtin_ldexp(b.x, e), __builtin_ldexp(b.y, e)};
    unsigned short w[4];
    if (hi) {
      float h[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        h[k] = (float)v[k];
        w[k] = __builtin_bit_cast(unsigned short, (__bf16)(float)(v[k] - (double)h[k]));
      }
      *reinterpret_cast<float4*>(hi + i) = float4{h[0], h[1], h[2], h[3]};
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = __builtin_bit_cast(unsigned short, (__bf16)(float)v[k]);
    }
    *reinterpret_cast<uint2*>(lo + i) =
        uint2{w[0] | (unsigned)w[1] << 16, w[2] | (unsigned)w[3] << 16};
  }
}
}  // namespace

#define MSW_BY_NM(P, call, ...)                  \
  switch (nm) {                                  \
    case 4: call(P, 4)(__VA_ARGS__); break;      \
    case 8: call(P, 8)(__VA_ARGS__); break;      \
    default: call(P, 16)(__VA_ARGS__); break;    \
  }

void launch_mfma(int precision, const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  const int nm = ls.nm;
  if (precision == MSW_PREC_MIXED_FP32) {
    MSW_BY_NM(F32, MSW_LAUNCH_NAME, ls, sp, st)
  } else if (precision == MSW_PREC_MIXED_BF16) {
    MSW_BY_NM(BF16, MSW_LAUNCH_NAME, ls, sp, st)
  } else {
    MSW_BY_NM(F64, MSW_LAUNCH_NAME, ls, sp, st)
  }
}

void launch_mfma_multi(int precision, int nm, int nac, int nwa, int ny, unsigned grid,
                       const SweepParams* probs, const MultiLaunch& ml, hipStream_t st) {
  const size_t lds = mfma_lds_bytes(precision, nac, nwa, ny);
  if (precision == MSW_PREC_MIXED_FP32) {
    MSW_BY_NM(F32, MSW_LAUNCH_MULTI_NAME, nac, nwa, ny, grid, lds, probs, ml, st)
  } else if (precision == MSW_PREC_MIXED_BF16) {
    MSW_BY_NM(BF16, MSW_LAUNCH_MULTI_NAME, nac, nwa, ny, grid, lds, probs, ml, st)
  } else {
    MSW_BY_NM(F64, MSW_LAUNCH_MULTI_NAME, nac, nwa, ny, grid, lds, probs, ml, st)
  }
}
#undef MSW_BY_NM

size_t operand_copy_bytes(int precision, size_t n) {
  return precision == MSW_PREC_MIXED_FP32   ? n * (sizeof(float) + sizeof(unsigned short))
         : precision == MSW_PREC_MIXED_BF16 ? n * sizeof(unsigned short)
                                            : 0;
}

void launch_operand_copy(int precision, const MswDims& d, const double* vi, float* hi,
                         unsigned short* lo, int e_off, hipStream_t st) {
  if (precision != MSW_PREC_MIXED_FP32 && precision != MSW_PREC_MIXED_BF16) return;
  const int nm_log2 = d.nm == 16 ? 4 : d.nm == 8 ? 3 : 2;  // nm is 4, 8 or 16
  const long long nruns = (long long)d.ncell_x * d.ncell_y * d.nu * d.nm / 4;  // per plane
  const long long planes = (long long)d.ncell_z * d.ne;
  // about 4 runs per thread, and no more blocks than 2^31
  long long chunks = std::max(1LL, (nruns + 1023) / 1024);
  chunks = std::min(chunks, std::max(1LL, (long long)INT32_MAX / planes));
  float* h = precision == MSW_PREC_MIXED_FP32 ? hi : nullptr;
  if (nruns * 4 + (long long)chunks * 256 * 4 < (1LL << 32))
    hipLaunchKernelGGL(k_operand_copy<unsigned>, dim3((unsigned)(planes * chunks)), dim3(256), 0,
                       st, vi, h, lo, d.ncell_x, d.ncell_y, d.ne, nm_log2, d.nu, (int)chunks,
                       e_off);
  else
    hipLaunchKernelGGL(k_operand_copy<unsigned long long>, dim3((unsigned)(planes * chunks)),
                       dim3(256), 0, st, vi, h, lo, d.ncell_x, d.ncell_y, d.ne, nm_log2, d.nu,
                       (int)chunks, e_off);
}


}  // namespace msw
