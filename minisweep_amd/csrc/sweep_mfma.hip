// sweep_mfma.hip -- the sweep with m2a/a2m on the gfx950 matrix cores.
//
// Same tiles, schedule and face buffers as the per-lane kernel in
// sweep_plan.hip; what changes is where the contractions run and how the
// 64 cells x NAC angles of a wave step are spread over lanes:
//
//   * The 64 cells of a step (an 8x8 tile on its skewed hyperplane) form 4
//     cell blocks of 16 (b = 0..3: rows yl = 2b, 2b+1).  Lane l = j + 16g
//     owns cell 16b + j of every block.
//   * m2a: VS[angle][cell] = AT[angle][moment] . VI[moment][cell] as 16x16x4
//     MFMAs.  AT is the A operand (per-octant, held in registers for the whole
//     launch), VI the B operand, loaded straight from the state array with
//     each lane reading NM/4 consecutive moments of its cell.
//   * The MFMA result layout puts angle row(g, r) of block a, cell 16b + j in
//     lane (j, g), register r.  The solve runs in that layout: the z face is the
//     lane's own value from the previous step, the x face comes from lane j-1
//     (DPP row shift by 1), the y face from lane j-8 of the same block or lane
//     j+8 of block b-1 (DPP row shifts by 8); tile-edge inflow comes from LDS.
//   * a2m is done transposed, VO[cell][moment] = R^T . M, so the solved values
//     are consumed as the A operand directly from the registers they sit in
//     (the k index of each 16x16x4 step is chosen to be the lane's own angle),
//     and each 16-lane group stores 16 consecutive moments of one cell.
//
// Reference: the per-cell work is Sweeper_sweep_cell
// (src/3_sweeper/sweeper_kba_c_kernels.h:30-555), the solve
// Quantities_solve (src/3_sweeper/quantities_testing_kernels.h:380-446).
#include <hip/hip_runtime.h>

#include "msw_device.h"
#include "msw_internal.h"

namespace msw {

typedef double d4v __attribute__((ext_vector_type(4)));

// DPP control for "lane i reads lane i-1 within its 16-lane row"; lanes whose
// source is outside the row keep the update_dpp 'old' operand.
constexpr int kRowShr1 = 0x111;

// FP64 16x16x4 MFMA: A[i][k] at lane i+16k, B[k][n] at lane n+16k,
// D[i][n] at lane n + 16*(i%4), register i/4.
__device__ __forceinline__ constexpr int row_f64(int g, int r) { return g + 4 * r; }

// The solve rescaled by the cell's power-of-two spatial factor S: with
// czs = xfluxweight*S/Sz, cys = yfluxweight*S/Sy, cxs = zfluxweight*S/Sx,
//   r = vs + ((fz*czs + fy*cys) + fx*cxs) * oct_r
// rounds exactly as the reference's
//   ((vs/S + ((fz*cz + fy*cy) + fx*cx) * oct_r) * S
// because every rescaling is by a power of two.
__device__ __forceinline__ double solve_scaled_f64(double vs, double fz, double fy,
                                                   double fx, double czs,
                                                   double cys, double cxs,
                                                   double oct_r) {
#pragma clang fp contract(off)
  return vs + ((fz * czs + fy * cys) + fx * cxs) * oct_r;
}

// Tile of one wave: 16 cells along x (one 16-lane row) by 4 along y (the
// 4 cell blocks).  Lane (j, g) owns cells (x = j, y = b) of every block b.
constexpr int kTX = 16;
constexpr int kTY = 4;
constexpr int kSkewM = (kTX - 1) + (kTY - 1);

template <int NM, int NAC, int NWA>
__global__ __launch_bounds__(64 * NWA) void k_sweep_mfma_f64(SweepParams p,
                                                             int item_base) {
  constexpr int NAB = NAC / 16;  // 16-angle blocks per wave
  constexpr int KM = NM / 4;     // moments per lane group in m2a
  constexpr int LROW = NAC + 1;
  constexpr int NXIN = (kTY * NAC + 63) / 64;  // x-inflow values per lane per step
  constexpr int NYIN = (kTX * NAC + 63) / 64;  // y-inflow values per lane per step
  static_assert(NAC % 16 == 0 && NM % 4 == 0, "block shapes");

  const int lane = threadIdx.x & 63;
  const int wave = NWA == 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = lane & 15, g = lane >> 4;
  int bid = blockIdx.x;
  const int iu = bid % p.nu;
  bid /= p.nu;
  const int ie = bid % p.ne;
  bid /= p.ne;
  const TileItem it = p.items[item_base + bid];
  const int oct = it.octant;
  const int dx = (oct & 1) ? -1 : 1;
  const int dy = (oct & 2) ? -1 : 1;
  const int dz = (oct & 4) ? -1 : 1;
  const int nx = p.nx, ny = p.ny, nz = p.nz, nap = p.na_pad;
  const int gx = dx > 0 ? it.tx : p.ntx - 1 - it.tx;
  const int gy = dy > 0 ? it.ty : p.nty - 1 - it.ty;
  // A partial last tile swept downward starts at the first lane inside.
  const int xoff = (dx < 0 && it.tx == 0) ? p.ntx * kTX - nx : 0;
  const int yoff = (dy < 0 && it.ty == 0) ? p.nty * kTY - ny : 0;
  const int ia0 = wave * NAC;
  const bool first = it.first != 0;

  // ---- the lane's cells: (xl = j, yl = b) ----------------------------------
  const int x = dx > 0 ? gx * kTX + j : gx * kTX + kTX - 1 - j;
  int yb[kTY];
#pragma unroll
  for (int b = 0; b < kTY; ++b) yb[b] = dy > 0 ? gy * kTY + b : gy * kTY + kTY - 1 - b;
  const bool x_ok = x < nx;

  extern __shared__ double lds[];
  double* lx = lds + (size_t)wave * ((kTX + kTY) * LROW);  // x inflow [yl][ia]
  double* ly = lx + kTY * LROW;                           // y inflow [xl][ia]
  double* red = lds + (size_t)NWA * ((kTX + kTY) * LROW); // [NWA-1][16][64]

  const double oct_s = (double)(1 + oct);
  const double oct_r = 1.0 / oct_s;
  const double eu_oct = (double)sf_energy(ie) * (double)sf_unknown(iu) * oct_s;
  const int xb = dx > 0 ? -1 : nx;
  const int ybnd = dy > 0 ? -1 : ny;
  const int zb = dz > 0 ? -1 : nz;

  const size_t face_ouy = (((size_t)oct * p.ne + ie) * p.nu + iu);
  const double* __restrict__ fyz_in = p.fyz + face_ouy * ny * nz * nap;  // [y][z][ia]
  const double* __restrict__ fxz_in = p.fxz + face_ouy * nx * nz * nap;  // [x][z][ia]
  double* fyz = p.fyz + face_ouy * ny * nz * nap;
  double* fxz = p.fxz + face_ouy * nx * nz * nap;

  // ---- per-octant matrices as MFMA operands, held for the whole launch ---
  // m2a A operand, k-step s: AT[ia0 + 16a + j][KM*g + s]
  double aop[NAB][KM];
  // a2m B operand, k-step (a, r): M[ia0 + 16a + row(g, r)][moment j]
  double mop[NAB][4];
  {
    const double* at = p.at + ((size_t)oct * nap + ia0) * NM;
    const double* mf = p.mf + ((size_t)oct * nap + ia0) * NM;
#pragma unroll
    for (int a = 0; a < NAB; ++a) {
#pragma unroll
      for (int s = 0; s < KM; ++s) aop[a][s] = at[(16 * a + j) * NM + KM * g + s];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        mop[a][r] = j < NM ? mf[(16 * a + row_f64(g, r)) * NM + j] : 0.0;
    }
  }
  // zfluxweight of the lane's angles: 1/4 - 2^-(ia & 7); ia & 7 = row & 7
  double wz[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) wz[r] = 0.25 - ldexp(1.0, -(row_f64(g, r) & 7));

  double F[NAB][kTY][4];
#pragma unroll
  for (int a = 0; a < NAB; ++a)
#pragma unroll
    for (int b = 0; b < kTY; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) F[a][b][r] = 0.0;

  const size_t sstride_z = (size_t)p.ne * ny * nx * p.nu * NM;
  const size_t sstride_e = (size_t)ny * nx * p.nu * NM;
  const double* __restrict__ vi_e = p.vi + (size_t)ie * sstride_e + (size_t)iu * NM;
  double* vo_e = p.vo + (size_t)ie * sstride_e + (size_t)iu * NM;

  // ---- software pipeline: everything but the face recurrence is fetched
  // one step ahead (vi, the old vo of the output cells, the tile inflow) --
  double vi_n[kTY][KM];   // m2a B operand of the next step
  double vo_n[kTY][4];    // old vo of the next step's output cells
  double xin_n[NXIN];     // next step's x inflow (flattened [row][ia])
  double yin_n[NYIN];     // next step's y inflow (flattened [col][ia])

  auto fetch = [&](int t) {
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      const int zp = t - j - b;
      const bool ok = x_ok && yb[b] < ny && zp >= 0 && zp < nz;
      const int z = dz > 0 ? zp : nz - 1 - zp;
      const double* src = vi_e + (size_t)z * sstride_z + ((size_t)yb[b] * nx + x) * p.nu * NM +
                          KM * g;
#pragma unroll
      for (int s = 0; s < KM; ++s) vi_n[b][s] = ok ? src[s] : 0.0;
    }
    if (!first && j < NM) {
#pragma unroll
      for (int b = 0; b < kTY; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int xo_l = row_f64(g, r);
          const int zp = t - xo_l - b;
          const int xo = dx > 0 ? gx * kTX + xo_l : gx * kTX + kTX - 1 - xo_l;
          const bool ok = xo < nx && yb[b] < ny && zp >= 0 && zp < nz;
          const int zo = dz > 0 ? zp : nz - 1 - zp;
          vo_n[b][r] = ok ? vo_e[(size_t)zo * sstride_z + ((size_t)yb[b] * nx + xo) * p.nu * NM + j]
                          : 0.0;
        }
    }
#pragma unroll
    for (int k = 0; k < NXIN; ++k) {
      const int idx = lane + 64 * k;
      const int row = idx / NAC, i = idx % NAC;
      const int zq = t - xoff - row;
      const int yq = dy > 0 ? gy * kTY + row : gy * kTY + kTY - 1 - row;
      const bool ok = idx < kTY * NAC && zq >= 0 && zq < nz && yq < ny;
      const int zg = dz > 0 ? zq : nz - 1 - zq;
      xin_n[k] = !ok ? 0.0
                 : it.tx == 0 ? (double)((1 + ia0 + i) << (i & 7)) *
                                    ldexp(eu_oct, sf_space_exp(xb, yq, zg))
                              : fyz_in[((size_t)yq * nz + zg) * nap + ia0 + i];
    }
#pragma unroll
    for (int k = 0; k < NYIN; ++k) {
      const int idx = lane + 64 * k;
      const int col = idx / NAC, i = idx % NAC;
      const int zq = t - col - yoff;
      const int xq = dx > 0 ? gx * kTX + col : gx * kTX + kTX - 1 - col;
      const bool ok = idx < kTX * NAC && zq >= 0 && zq < nz && xq < nx;
      const int zg = dz > 0 ? zq : nz - 1 - zq;
      yin_n[k] = !ok ? 0.0
                 : it.ty == 0 ? (double)((1 + ia0 + i) << (i & 7)) *
                                    ldexp(eu_oct, sf_space_exp(xq, ybnd, zg))
                              : fxz_in[((size_t)xq * nz + zg) * nap + ia0 + i];
    }
  };

  fetch(0);
  const int nsteps = nz + kSkewM;
  for (int t = 0; t < nsteps; ++t) {
    // ---- take this step's prefetched data; start fetching the next -------
    double viop[kTY][KM];
    double voo[kTY][4];
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
#pragma unroll
      for (int s = 0; s < KM; ++s) viop[b][s] = vi_n[b][s];
#pragma unroll
      for (int r = 0; r < 4; ++r) voo[b][r] = vo_n[b][r];
    }
#pragma unroll
    for (int k = 0; k < NXIN; ++k) {
      const int idx = lane + 64 * k;
      if (idx < kTY * NAC) lx[(idx / NAC) * LROW + idx % NAC] = xin_n[k];
    }
#pragma unroll
    for (int k = 0; k < NYIN; ++k) {
      const int idx = lane + 64 * k;
      if (idx < kTX * NAC) ly[(idx / NAC) * LROW + idx % NAC] = yin_n[k];
    }
    if (t + 1 < nsteps) fetch(t + 1);

    // ---- per cell: z inflow on entry, rescaled solve coefficients ----------
    bool act[kTY];
    int zc[kTY];
    double czs[kTY], cys[kTY], sxs[kTY];
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      const int zp = t - j - b;
      act[b] = x_ok && yb[b] < ny && zp >= 0 && zp < nz;
      const int z = dz > 0 ? zp : nz - 1 - zp;
      zc[b] = z;
      if (x_ok && yb[b] < ny && zp == 0) {
        const double sc = ldexp(eu_oct, sf_space_exp(x, yb[b], zb));
#pragma unroll
        for (int a = 0; a < NAB; ++a)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int ia = ia0 + 16 * a + row_f64(g, r);
            F[a][b][r] = (double)((1 + ia) << (ia & 7)) * sc;
          }
      }
      const int e_s = sf_space_exp(x, yb[b], z);
      czs[b] = ldexp(0.5, e_s - sf_space_exp(x, yb[b], z - dz));
      cys[b] = ldexp(0.25, e_s - sf_space_exp(x, yb[b] - dy, z));
      sxs[b] = ldexp(1.0, e_s - sf_space_exp(x - dx, yb[b], z));
    }
    __builtin_amdgcn_wave_barrier();

    d4v vo_acc[kTY];
#pragma unroll
    for (int b = 0; b < kTY; ++b) vo_acc[b] = d4v{0.0, 0.0, 0.0, 0.0};

#pragma unroll
    for (int a = 0; a < NAB; ++a) {
      // m2a for the 4 cell blocks of angle block a
      d4v vs[kTY];
#pragma unroll
      for (int b = 0; b < kTY; ++b) {
        vs[b] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KM; ++s)
          vs[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[a][s], viop[b][s], vs[b], 0, 0, 0);
      }
      // solve.  The y face of block b is block b-1's face from the previous
      // step (same lane, same register): blocks go in descending order so
      // it is read before it is overwritten.  The x face is lane j-1's
      // (DPP row shift); lane j = 0 keeps the tile inflow read from LDS.
#pragma unroll
      for (int b = kTY - 1; b >= 0; --b) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ial = 16 * a + row_f64(g, r);
          const double xin = lx[b * LROW + ial];
          double fx = __hiloint2double(
              __builtin_amdgcn_update_dpp(__double2hiint(xin), __double2hiint(F[a][b][r]),
                                          kRowShr1, 0xF, 0xF, false),
              __builtin_amdgcn_update_dpp(__double2loint(xin), __double2loint(F[a][b][r]),
                                          kRowShr1, 0xF, 0xF, false));
          if (xoff) fx = (j == xoff) ? xin : fx;
          double fy;
          if (b == 0) {
            fy = ly[j * LROW + ial];
          } else {
            fy = F[a][b - 1][r];
            if (yoff) fy = (b == yoff) ? ly[j * LROW + ial] : fy;
          }
          const double res = solve_scaled_f64(vs[b][r], F[a][b][r], fy, fx, czs[b],
                                              cys[b], wz[r] * sxs[b], oct_r);
          F[a][b][r] = res * oct_s;
          vs[b][r] = res;
        }
      }
      // a2m (transposed): VO[cell][moment] += R^T[cell][angle] M[angle][moment]
#pragma unroll
      for (int b = 0; b < kTY; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          vo_acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(vs[b][r], mop[a][r], vo_acc[b],
                                                           0, 0, 0);
    }

    // ---- outflow faces for the downstream tiles ----------------------------
    if (j == kTX - 1 && it.tx + 1 < p.ntx) {
#pragma unroll
      for (int b = 0; b < kTY; ++b)
        if (act[b]) {
          double* dst = fyz + ((size_t)yb[b] * nz + zc[b]) * nap + ia0;
#pragma unroll
          for (int a = 0; a < NAB; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[16 * a + row_f64(g, r)] = F[a][b][r];
        }
    }
    if (it.ty + 1 < p.nty && act[kTY - 1]) {
      double* dst = fxz + ((size_t)x * nz + zc[kTY - 1]) * nap + ia0;
#pragma unroll
      for (int a = 0; a < NAB; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[16 * a + row_f64(g, r)] = F[a][kTY - 1][r];
    }

    // ---- vo: lane (moment j, group g) holds cells x = row(g, r) of block b --
    if constexpr (NWA > 1) {
      if (wave > 0) {
#pragma unroll
        for (int b = 0; b < kTY; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            red[((size_t)(wave - 1) * 16 + 4 * b + r) * 64 + lane] = vo_acc[b][r];
      }
      __syncthreads();
      if (wave == 0) {
        // partial sums of the angle chunks, added in chunk order
#pragma unroll
        for (int w = 1; w < NWA; ++w)
#pragma unroll
          for (int b = 0; b < kTY; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              vo_acc[b][r] += red[((size_t)(w - 1) * 16 + 4 * b + r) * 64 + lane];
      }
    }
    if (wave == 0 && j < NM) {
#pragma unroll
      for (int b = 0; b < kTY; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int xo_l = row_f64(g, r);
          const int zp = t - xo_l - b;
          const int xo = dx > 0 ? gx * kTX + xo_l : gx * kTX + kTX - 1 - xo_l;
          if (xo < nx && yb[b] < ny && zp >= 0 && zp < nz) {
            const int zo = dz > 0 ? zp : nz - 1 - zp;
            vo_e[(size_t)zo * sstride_z + ((size_t)yb[b] * nx + xo) * p.nu * NM + j] =
                first ? vo_acc[b][r] : voo[b][r] + vo_acc[b][r];
          }
        }
    }
    if constexpr (NWA > 1) __syncthreads();
  }
}

template <int NM, int NAC>
static void launch_nac(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  const dim3 block(64 * ls.nwa);
  switch (ls.nwa) {
    case 1:
      hipLaunchKernelGGL((k_sweep_mfma_f64<NM, NAC, 1>), dim3(ls.grid), block, ls.lds_bytes,
                         st, sp, ls.item_base);
      break;
    case 2:
      hipLaunchKernelGGL((k_sweep_mfma_f64<NM, NAC, 2>), dim3(ls.grid), block, ls.lds_bytes,
                         st, sp, ls.item_base);
      break;
    case 3:
      hipLaunchKernelGGL((k_sweep_mfma_f64<NM, NAC, 3>), dim3(ls.grid), block, ls.lds_bytes,
                         st, sp, ls.item_base);
      break;
    default:
      hipLaunchKernelGGL((k_sweep_mfma_f64<NM, NAC, 4>), dim3(ls.grid), block, ls.lds_bytes,
                         st, sp, ls.item_base);
      break;
  }
}

template <int NM>
static void launch_nm(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  if (ls.nac == 16)
    launch_nac<NM, 16>(ls, sp, st);
  else if (ls.nac == 32)
    launch_nac<NM, 32>(ls, sp, st);
  else
    launch_nac<NM, 64>(ls, sp, st);
}

void launch_mfma_f64(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  switch (ls.nm) {
    case 4: launch_nm<4>(ls, sp, st); break;
    case 8: launch_nm<8>(ls, sp, st); break;
    default: launch_nm<16>(ls, sp, st); break;
  }
}

size_t mfma_f64_lds_bytes(int nac, int nwa) {
  return sizeof(double) *
         ((size_t)nwa * (kTX + kTY) * (nac + 1) + (size_t)(nwa - 1) * 16 * 64);
}

int mfma_tile_x() { return kTX; }
int mfma_tile_y() { return kTY; }

}  // namespace msw
