// sweep_mfma.hip -- the sweep with m2a/a2m on the gfx950 matrix cores, in
// FP64 (plain) and in mixed precision (FP32 or BF16 contractions with FP32
// accumulation, FP32 faces and solve, FP64 input state; vo summed over the
// octants in FP32 and written in FP64 by each tile's last octant visit).
//
// One wave owns 16 (x) by 4 (y) cells for one (octant, run of energy groups,
// unknown, chunk of NAC angles) and walks the z extent of every group of the
// run as one stream.  Lane l = j + 16g owns cell (x = j, y = b) of every cell
// block b = 0..3; at step t that cell is at stream position t - j - b, so the
// 64 cells of a step lie on one diagonal hyperplane.  A workgroup is NWA
// angle-chunk waves times NY row groups (a 16 x 4NY tile); row group w runs
// 4w steps behind row group w-1 and takes its y inflow from it through LDS.
//
//   * m2a: VS[angle][cell] = AT[angle][moment] . VI[moment][cell] as 16x16
//     MFMAs.  AT is the A operand (per octant, in registers for the whole
//     launch), VI the B operand, loaded from the state array with each lane
//     reading consecutive moments of its own cell.
//   * The MFMA result puts angle row(g, r) of the 16-angle block in lane
//     (j, g), register r.  The solve runs in that layout: the z face is the
//     lane's own value from the previous step, the y face block b-1's value of
//     the same lane, the x face lane j-1's (DPP row shift).  Lane j = 0 gets
//     the tile's x inflow (loaded one value per lane of the row, gathered by
//     a DPP row rotation) and block 0 the y inflow, read from the face planes
//     one step ahead.
//   * The carried face value is the solved angular value itself; the
//     reference's faces hold it times (1 + octant), which cancels against its
//     1/(1 + octant), and the cell's power-of-two spatial scale is folded into
//     the three face coefficients, so the solve is three FMAs:
//        r = vs + Sz'*rz + Sy'*ry + Sx'*rx.
//   * a2m: VO[moment][cell] = M^T[moment][angle] . R[angle][cell], with the
//     solved values as the B operand straight from the registers they sit in
//     and M^T's rows ordered so that lane (j, g) ends with moments 4g..4g+3 of
//     cell j: vo is read and written as 32-byte runs.
//   * The angle chunks of a tile are the waves of one workgroup; each wave
//     stores vo for the cell blocks it owns after adding the other waves'
//     partial sums (LDS, double-buffered by step parity, fixed order).
//   * Face planes in HBM hold, per (octant, group, unknown), one yz plane for
//     x faces and one xz plane for y faces, angles in lane order.  The host
//     schedule never runs two tiles of one octant row (column) in one launch,
//     so a tile reads its inflow and writes its outflow in place.  A pre-pass
//     fills the planes with the domain-boundary inflow, so every tile takes
//     its inflow the same way; only the z boundary is computed in the kernel.
//
// MFMA layouts (gfx950):
//   f64 16x16x4 : A[i][k] lane i+16k, B[k][n] lane n+16k, D[i][n] lane n+16(i%4) reg i/4
//   f32 16x16x4 : A, B as f64;                           D[i][n] lane n+16(i/4) reg i%4
//   bf16 16x16x16: A[i][k] lane i+16(k/4) elem k%4, B[k][n] lane n+16(k/4) elem k%4, D as f32
//
// Reference: the per-cell work is Sweeper_sweep_cell
// (src/3_sweeper/sweeper_kba_c_kernels.h:30-555), the solve
// Quantities_solve (src/3_sweeper/quantities_testing_kernels.h:380-446), the
// boundary inflow Quantities_init_face (same file, :219-332).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "msw_device.h"
#include "msw_internal.h"
#include "../../include/msw_plan.h"

namespace msw {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));

// DPP control for "lane i reads lane i-1 within its 16-lane row"; lanes whose
// source is outside the row keep the update_dpp 'old' operand.
constexpr int kRowShr1 = 0x111;

constexpr int kTX = 16;  // tile cells along x: one 16-lane row
constexpr int kTY = 4;   // tile cells along y: the 4 cell blocks

__device__ __forceinline__ double shr1_keep(double old, double v) {
  return __hiloint2double(
      __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), kRowShr1, 0xF, 0xF,
                                  false),
      __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), kRowShr1, 0xF, 0xF,
                                  false));
}
__device__ __forceinline__ float shr1_keep(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v),
                                                    kRowShr1, 0xF, 0xF, false));
}

// Lane i of each 16-lane row reads lane (i - k) mod 16 (row_ror:k, k in 1..15).
template <int K>
__device__ __forceinline__ float row_ror(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + K, 0xF, 0xF, false));
}
template <int K>
__device__ __forceinline__ double row_ror(double v) {
  return __hiloint2double(__builtin_amdgcn_mov_dpp(__double2hiint(v), 0x120 + K, 0xF, 0xF, false),
                          __builtin_amdgcn_mov_dpp(__double2loint(v), 0x120 + K, 0xF, 0xF, false));
}
// Lane 0 of each row gets the row's lane s (s in 0..15, a constant after
// unrolling); other lanes get whatever the rotation brings (unused).
template <class T>
__device__ __forceinline__ T from_lane(T v, int s) {
  switch (s) {
    case 1: return row_ror<15>(v);
    case 2: return row_ror<14>(v);
    case 3: return row_ror<13>(v);
    case 4: return row_ror<12>(v);
    case 5: return row_ror<11>(v);
    case 6: return row_ror<10>(v);
    case 7: return row_ror<9>(v);
    case 8: return row_ror<8>(v);
    case 9: return row_ror<7>(v);
    case 10: return row_ror<6>(v);
    case 11: return row_ror<5>(v);
    case 12: return row_ror<4>(v);
    case 13: return row_ror<3>(v);
    case 14: return row_ror<2>(v);
    case 15: return row_ror<1>(v);
    default: return v;
  }
}

__device__ __forceinline__ double fmaT(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fmaT(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double ldexpT(double a, int e) { return __builtin_ldexp(a, e); }
__device__ __forceinline__ float ldexpT(float a, int e) { return __builtin_ldexpf(a, e); }

// Four consecutive values, 16-byte aligned.
__device__ __forceinline__ void load4(const double* s, double (&v)[4]) {
  const double2 a = reinterpret_cast<const double2*>(s)[0];
  const double2 b = reinterpret_cast<const double2*>(s)[1];
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void load4(const float* s, float (&v)[4]) {
  const float4 a = *reinterpret_cast<const float4*>(s);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void store4(double* d, const double (&v)[4]) {
  reinterpret_cast<double2*>(d)[0] = double2{v[0], v[1]};
  reinterpret_cast<double2*>(d)[1] = double2{v[2], v[3]};
}
__device__ __forceinline__ void store4(float* d, const float (&v)[4]) {
  *reinterpret_cast<float4*>(d) = float4{v[0], v[1], v[2], v[3]};
}

// ---- precision traits ------------------------------------------------------
// T: face / solve / operand type; A: MFMA accumulator element; row(g, r): the
// angle (within a 16-angle block) that MFMA result register r of lane group g
// holds; mom(i): the moment that a2m result row i stands for, chosen so that
// lane group g, register r holds moment 4g + r.  kPingPong: the z loop
// alternates two operand buffers (two steps per iteration) instead of copying
// the prefetched operands at the top of each step; measured faster for the
// mixed forms, slower for FP64 (register pressure).
struct PrecF64 {
  typedef double T;
  typedef double A;
  typedef double Vin;  // m2a B operand as held in registers
  typedef d4v acc4;
  static constexpr bool kPacked = false;
  static constexpr bool kPingPong = false;
  static constexpr int kWz = 2;  // (g + 4r) & 7 depends on r & 1 only
  static __device__ __forceinline__ constexpr int row(int g, int r) { return g + 4 * r; }
  static __device__ __forceinline__ constexpr int mom(int i) { return 4 * (i & 3) + (i >> 2); }
  static __device__ __forceinline__ acc4 mma(T a, T b, acc4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
struct PrecF32 {
  typedef float T;
  typedef float A;
  typedef float Vin;
  typedef f4v acc4;
  static constexpr bool kPacked = false;
  static constexpr bool kPingPong = true;
  static constexpr int kWz = 4;
  static __device__ __forceinline__ constexpr int row(int g, int r) { return 4 * g + r; }
  static __device__ __forceinline__ constexpr int mom(int i) { return i; }
  static __device__ __forceinline__ acc4 mma(T a, T b, acc4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
struct PrecBF16 {
  typedef float T;
  typedef float A;
  typedef short Vin;  // the BF16 bits of vi, read from the BF16 copy
  typedef f4v acc4;
  static constexpr bool kPacked = true;  // K = 16: 4 values per lane per MFMA
  static constexpr bool kPingPong = true;
  static constexpr int kWz = 4;
  static __device__ __forceinline__ constexpr int row(int g, int r) { return 4 * g + r; }
  static __device__ __forceinline__ constexpr int mom(int i) { return i; }
  static __device__ __forceinline__ acc4 mma16(const T (&a)[4], const T (&b)[4], acc4 c) {
    s4v pa, pb;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      pa[e] = __builtin_bit_cast(short, (__bf16)a[e]);
      pb[e] = __builtin_bit_cast(short, (__bf16)b[e]);
    }
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(pa, pb, c, 0, 0, 0);
  }
  // the same with B already in BF16
  static __device__ __forceinline__ acc4 mma16b(const T (&a)[4], const short (&b)[4], acc4 c) {
    s4v pa, pb;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      pa[e] = __builtin_bit_cast(short, (__bf16)a[e]);
      pb[e] = b[e];
    }
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(pa, pb, c, 0, 0, 0);
  }
};

// Exponent LUT of the face coefficients.  With e = (x + 3y + 7z + 2) & 3 the
// cell's spatial exponent, the upwind neighbours' exponents are (e + dz) & 3,
// (e + dy) & 3 and (e - dx) & 3, so each coefficient's exponent is a function
// of e alone: nibble e holds e - e_upwind + 8.
__host__ __device__ __forceinline__ unsigned coef_lut(int shift) {
  unsigned l = 0;
  for (int e = 0; e < 4; ++e) l |= (unsigned)(e - ((e + shift) & 3) + 8) << (4 * e);
  return l;
}

// Waves per SIMD the register allocation aims for: two where the working set
// fits in 256 registers without spilling to scratch, so that one wave's matrix
// work overlaps the other's solve and loads; else one (the compiler then
// keeps the overflow in accumulation registers).  The shapes that fit were
// read off the compiler's resource report for every instantiation: 16 angles
// per wave in FP32; 32 in FP32 with one row group (or two with two angle
// waves) below 16 moments, and in BF16 likewise at any moment count; 16 in
// FP64 with several angle waves, one row group and below 16 moments.
template <class PR, int NM, int NAC, int NWA, int NY>
struct Occupancy {
  static constexpr int bytes = (int)sizeof(typename PR::T) * NAC;
  static constexpr bool rows_ok = NY == 1 || (NWA == 2 && NY == 2 && NM < 16);
  static constexpr int value =
      bytes <= 64                                                          ? 2
      : bytes <= 128 && sizeof(typename PR::T) == 4 && rows_ok &&
              (NM < 16 || PR::kPacked)                                     ? 2
      : bytes <= 128 && NWA >= 2 && NY == 1 && NM < 16                     ? 2
                                                                           : 1;
  // the many-problem kernel needs a few registers more (problem lookup): two
  // waves only at 16 angles per wave in FP32 / BF16
  static constexpr int multi = bytes <= 64 ? 2 : 1;
};

// One workgroup's work: tile item `it` of a sweep, energy groups
// [ie0, ie0 + nge) streamed back to back through one z pipeline, unknown iu,
// angle chunk = wave index.  Lane (j, b) is at stream position
// t - j - b = (group - ie0) * nz + z, so the tile's fill and drain are paid
// once per nge groups.  ZB: built for z blocks (z faces between blocks);
// the whole-grid kernels carry none of that code.
template <class PR, int NM, int NAC, int NWA, int NY, bool ZB>
__device__ __forceinline__ void sweep_tile(const SweepParams& p, const TileItem it,
                                           const int ie0, const int nge, const int iu) {
  typedef typename PR::T T;
  typedef typename PR::A A;
  typedef typename PR::acc4 acc4;
  constexpr int NAB = NAC / 16;                // 16-angle blocks per wave
  constexpr int KM = NM / 4;                   // m2a k-steps (unpacked)
  constexpr int KA = PR::kPacked ? 4 : KM;     // m2a operand values per lane and block
  constexpr int MSTEP = PR::kPacked ? 4 : KM;  // first m2a moment of lane group g: MSTEP*g
  constexpr int NOWN = (kTY + NWA - 1) / NWA;  // cell blocks whose vo one wave stores
  constexpr bool kF64 = std::is_same<typename PR::T, double>::value;
  static_assert(NAC % 16 == 0 && NM % 4 == 0 && NM <= 16, "block shapes");

  constexpr int kTYT = kTY * NY;                // tile rows: NY waves of 4 rows
  const int lane = threadIdx.x & 63;
  const int wave = NWA * NY == 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wa = NWA == 1 ? 0 : wave % NWA;     // angle chunk of this wave
  const int wyi = NY == 1 ? 0 : wave / NWA;     // row group of this wave: rows 4 wyi..4 wyi+3
  const int j = lane & 15, g = lane >> 4;
  const int oct = it.octant;
  const int dx = (oct & 1) ? -1 : 1;
  const int dy = (oct & 2) ? -1 : 1;
  const int dz = (oct & 4) ? -1 : 1;
  const int nx = p.nx, ny = p.ny, nz = p.nz, nu = p.nu, nap = p.na_pad;
  const int gx = dx > 0 ? it.tx : p.ntx - 1 - it.tx;
  const int gy = dy > 0 ? it.ty : p.nty - 1 - it.ty;
  const int ia0 = wa * NAC;
  const bool first = (it.flags & kVisitFirst) != 0;
  const bool last = (it.flags & kVisitLast) != 0;
  const bool x_out = it.tx + 1 < p.ntx, y_out = it.ty + 1 < p.nty;

  // ---- the lane's cells ------------------------------------------------------
  // The tile covers grid cells [16 gx, 16 gx + wx) x [kTYT gy, kTYT gy + wy)
  // for every octant; lane j = 0 and row 0 are its upwind-most column and row
  // in the octant's direction, so a partial tile's missing cells are always
  // downstream (lanes j >= wx, rows >= wy).  Wave row group wyi owns rows
  // 4 wyi + b and runs 4 wyi steps behind row group 0, so that its row 0
  // meets the row above it one step after that row was solved.
  const int wx = min(kTX, nx - gx * kTX), wy = min(kTYT, ny - gy * kTYT);
  const bool x_ok = j < wx;
  const int jc = x_ok ? j : wx - 1;  // clamped: loads never leave the arrays
  const int x = dx > 0 ? gx * kTX + jc : gx * kTX + wx - 1 - jc;
  int yb[kTY], cst[kTY];
  const size_t Se = (size_t)ny * nx * nu * NM;  // state stride of a group
  const size_t Sz = (size_t)p.ne * Se;          // state stride of a z plane
  // State offset of each block's cell, kept incrementally (no multiplies in
  // the z loop): it starts at z = 0 of group ie0 in the octant's frame and
  // moves by dS per step, by dW where the stream enters the next group, and
  // stays where it is before and after the stream (clamped, never used).
  const size_t cxo = (size_t)ie0 * Se + ((size_t)x * nu + iu) * NM +
                     (dz > 0 ? (size_t)0 : (size_t)(nz - 1) * Sz);
  const long long dS = dz > 0 ? (long long)Sz : -(long long)Sz;
  const long long dW = (long long)Se - (long long)(nz - 1) * dS;
  size_t sc[kTY];  // this step's offsets
#pragma unroll
  for (int b = 0; b < kTY; ++b) {
    const int r = kTY * wyi + b;
    const int bc = r < wy ? r : wy - 1;
    yb[b] = dy > 0 ? gy * kTYT + bc : gy * kTYT + wy - 1 - bc;
    sc[b] = cxo + (size_t)yb[b] * nx * nu * NM;
    cst[b] = x + 3 * yb[b] + 2;
  }
  // z blocks: this block's planes are grid planes z_off .. z_off + nz - 1;
  // the domain's z boundary is outside the grid, not the block
  const int zb = dz > 0 ? -1 : (p.nz_grid ? p.nz_grid : nz);
  const int zoff = p.z_off;

  // ---- LDS: vo partial sums [2][NY][kTY][NWA][64][4], y faces between row
  // groups [2][NY-1][NWA][NAB][64][4] --
  extern __shared__ double lds_raw[];
  A* red = reinterpret_cast<A*>(lds_raw);
  T* yring = reinterpret_cast<T*>(red + (NWA > 1 ? 2 * NY * kTY * NWA * 64 * 4 : 0));

  // ---- face planes of this (octant, unknown): [y or x][group][z][angle],
  // z in the octant's frame so that a stream position maps to one offset.
  // Offsets within one (octant, unknown) plane fit 32 bits (the plan checks
  // max(nx, ny) * ne * nz * na_pad < 2^31). --
  const size_t nze = (size_t)p.ne * nz;  // stream length of all groups
  const int fplane = (int)nze * nap;     // one y (x face) or x (y face) row
  const size_t face_ou = (size_t)oct * nu + iu;
  T* fyz = reinterpret_cast<T*>(p.fyz) + (face_ou * ny * nze + (size_t)ie0 * nz) * nap + ia0 +
           4 * g;
  T* fxz = reinterpret_cast<T*>(p.fxz) + (face_ou * nx * nze + (size_t)ie0 * nz) * nap + ia0 +
           4 * g + x * fplane;  // this lane's column
  // z faces between z blocks: [octant & 3][iu][group][y][x][angle], the lane's
  // column; fz_cell(b) addresses block b's cell in its current group
  T* fzl = p.fz ? reinterpret_cast<T*>(p.fz) +
                      ((size_t)((oct & 3) * nu + iu) * p.ne + ie0) * ny * nx * nap +
                      (size_t)x * nap + ia0 + 4 * g
                : nullptr;

  // ---- per-octant operands, in registers for the whole launch --------------
  const int moff = MSTEP * g;
  T aop[NAB][KA];      // m2a A: AT[ia0 + 16a + j][moff + s]
  T mop[NAB][4];       // a2m A: M[ia0 + 16a + row(g, s)][mom(j)]
  T wz[PR::kWz];       // zfluxweight of the lane's angles: 1/4 - 2^-(ia & 7), by r % kWz
  float iaf[NAB][4];   // (1 + ia) 2^(ia & 7) (exact in FP32): boundary inflow angle factor
  {
    const double* at = p.at + ((size_t)oct * nap + ia0) * NM;
    const double* mf = p.mf + ((size_t)oct * nap + ia0) * NM;
    const int mm = PR::mom(j);
#pragma unroll
    for (int a = 0; a < NAB; ++a) {
#pragma unroll
      for (int s = 0; s < KA; ++s) {
        const int im = moff + s;
        aop[a][s] = im < NM ? (T)at[(16 * a + j) * NM + (im < NM ? im : 0)] : (T)0;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        mop[a][s] = mm < NM ? (T)mf[(16 * a + PR::row(g, s)) * NM + (mm < NM ? mm : 0)] : (T)0;
        const int ia = ia0 + 16 * a + PR::row(g, s);
        iaf[a][s] = ia < p.na ? (float)((1 + ia) << (ia & 7)) : 0.0f;
      }
    }
#pragma unroll
    for (int s = 0; s < PR::kWz; ++s) wz[s] = (T)0.25 - ldexpT((T)1, -(PR::row(g, s) & 7));
  }
  const int su = sf_unknown(iu);
  const unsigned lut_z = coef_lut(dz), lut_y = coef_lut(dy), lut_x = coef_lut(-dx);

  T F[NAB][kTY][4];  // outgoing face value (= solved value) of the lane's cells
#pragma unroll
  for (int a = 0; a < NAB; ++a)
#pragma unroll
    for (int b = 0; b < kTY; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) F[a][b][r] = (T)0;

  // ---- data fetched one step ahead; each item is refetched for the next step
  // right after its last use in this one.  Every prefetch is issued on every
  // path, the last step's too (z is clamped), so that the memory counters
  // agree across paths and each wait covers only the data it needs. ---------
  // m2a B operand, two buffers: one step's while the next step's loads are in
  // flight (a whole step ahead)
  typedef typename PR::Vin Vin;
  Vin vbA[kTY][KA], vbB[kTY][KA];
  // running vo sum of the owned blocks' cells from earlier octant visits:
  // FP64 vo in the FP64 form, the FP32 sum in the mixed forms
  typedef typename std::conditional<kF64, double, float>::type V;
  V von[NOWN][4];
  // x inflow, spread: lane (j, g) holds block j/4, register j%4, and lane
  // (0, g) gathers it by DPP; two buffers as for vi
  T xbA[NAB], xbB[NAB];
  T yinn[NAB][4];        // y inflow of block 0
  const int ns = nge * nz;  // stream length of this workgroup
  const int nsteps = ns + kTX + kTYT - 2;
  const bool mpad = moff + KA > NM;  // packed operand past the last moment

  // Stream position of block b's cell: (group offset es, z in the octant's
  // frame zs), advanced by one each step.  Before the stream zs < 0, after it
  // es >= nge; addresses use the clamped position.
  int zs[kTY], es[kTY];
#pragma unroll
  for (int b = 0; b < kTY; ++b) {
    zs[b] = -j - b - kTY * wyi;
    es[b] = 0;
  }
  auto fz_cell = [&](int b) {
    const int ec = es[b] < nge ? es[b] : nge - 1;
    return fzl + ((size_t)ec * ny + yb[b]) * nx * nap;
  };
  int zn[kTY], en[kTY];  // next step's stream position per block
  size_t sn[kTY];        // next step's state offsets
  auto advance = [&]() {
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      const bool wrap = zs[b] + 1 == nz;
      zn[b] = wrap ? 0 : zs[b] + 1;
      en[b] = wrap ? es[b] + 1 : es[b];
      const bool move = zs[b] >= 0 && en[b] < nge;
      sn[b] = sc[b] + (size_t)(move ? (wrap ? dW : dS) : 0ll);
    }
  };
  // Face-plane offsets.  Stores of block b's outflow go to stream position
  // t - j - b - 4 wyi, valid whenever they happen: fo0 - b * nap with fo0 the
  // position t - j - 4 wyi times nap.  The inflow loads of the next step use
  // clamped positions, kept as (position, offset) pairs.
  int fo0 = (-j - kTY * wyi) * nap;
  const int bb = j >> 2;  // x inflow: lane (j, g) fetches block j/4's value
  int pxs = -bb - kTY * wyi, oxs = 0;  // x inflow position of the step fetched
  int pyi = -j, oyi = 0;               // y inflow position (row group 0)
  auto bump = [&](int& pos, int& off) {  // position + 1, offset clamped to [0, ns)
    ++pos;
    off += pos >= 1 && pos < ns ? nap : 0;
  };
  const T* xs_src;
  {
    const int r = kTY * wyi + bb;
    const int bc = r < wy ? r : wy - 1;
    const int y = dy > 0 ? gy * kTYT + bc : gy * kTYT + wy - 1 - bc;
    xs_src = fyz + y * fplane + (j & 3);
  }
  // the input state: FP64, or its BF16 copy in the BF16 form (4 values of a
  // lane in 8 bytes)
  auto fetch_vi = [&](const size_t (&so)[kTY], Vin (&vnx)[kTY][KA]) {
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      if constexpr (PR::kPacked) {
        uint2 w = {0u, 0u};
        if (!mpad) w = *reinterpret_cast<const uint2*>(p.vib + so[b] + moff);
        vnx[b][0] = (short)(w.x & 0xffffu);
        vnx[b][1] = (short)(w.x >> 16);
        vnx[b][2] = (short)(w.y & 0xffffu);
        vnx[b][3] = (short)(w.y >> 16);
      } else if constexpr (KA == 4) {
        double v[4] = {0, 0, 0, 0};
        if (!mpad) load4(p.vi + so[b] + moff, v);
#pragma unroll
        for (int s = 0; s < 4; ++s) vnx[b][s] = (T)v[s];
      } else {
        const double* src = p.vi + so[b] + moff;
#pragma unroll
        for (int s = 0; s < KA; ++s) vnx[b][s] = (T)src[s];
      }
    }
  };
  // running sum: vo (FP64 form) or vo32 (mixed).  A tile's first visit has
  // no sum to read; it loads 16 bytes of the octant matrices instead (cached,
  // never used), so the load is still issued on every path.
  const V* vrd = first ? (const V*)p.at : kF64 ? (const V*)p.vo : (const V*)p.vo32;
  auto fetch_vo = [&](const size_t (&so)[kTY]) {
#pragma unroll
    for (int b = 0; b < kTY; ++b)
      if (b % NWA == wa && 4 * g < NM)
        load4(vrd + (first ? 0 : so[b]) + 4 * g, von[b / NWA]);
  };
  auto fetch_xs = [&](T (&xsn)[NAB]) {
#pragma unroll
    for (int a = 0; a < NAB; ++a) xsn[a] = xs_src[oxs + 16 * a];
  };
  auto fetch_yin = [&]() {  // row group 0 only
#pragma unroll
    for (int a = 0; a < NAB; ++a) load4(fxz + oyi + 16 * a, yinn[a]);
  };

  // m2a of angle block a: VS[angle][cell] = AT . VI
  auto m2a = [&](int a, const Vin (&vin)[kTY][KA], acc4 (&vs)[kTY]) {
#pragma unroll
    for (int b = kTY - 1; b >= 0; --b) {
      vs[b] = acc4{0, 0, 0, 0};
      if constexpr (PR::kPacked) {
        vs[b] = PR::mma16b(aop[a], vin[b], vs[b]);
      } else {
#pragma unroll
        for (int s = 0; s < KM; ++s) vs[b] = PR::mma(aop[a][s], vin[b][s], vs[b]);
      }
    }
  };

  fetch_xs(PR::kPingPong ? xbA : xbB);
  fetch_vi(sc, PR::kPingPong ? vbA : vbB);
  if (wyi == 0) fetch_yin();
  fetch_vo(sc);
  // BF16: m2a of angle block 0 for the coming step is issued at the end of
  // the step before, so the matrix cores run it during that step's vo
  // reduction (41.5 vs 43.0 ms per sweep at 64^3; FP64 and FP32 ran slower
  // that way, 92.8 vs 88.2 and 53.1 vs 53.0 ms)
  constexpr bool kEarlyM2a = PR::kPacked;
  acc4 vs0[kTY];
  if constexpr (kEarlyM2a) m2a(0, PR::kPingPong ? vbA : vbB, vs0);

  // One step: this step's operands are vin and xsc, the next step's are
  // fetched into vnx and xsn first thing.
  auto step = [&](const int t, const Vin (&vin)[kTY][KA], Vin (&vnx)[kTY][KA],
                  const T (&xsc)[NAB], T (&xsn)[NAB]) {
    advance();
    fetch_vi(sn, vnx);
    bump(pxs, oxs);
    fetch_xs(xsn);

    // ---- per cell block: validity, z, spatial phase; z inflow on entry -------
    // A lane whose cell starts a group (z = 0) takes the domain-boundary z
    // inflow instead of its own face register; the register itself still
    // holds the previous group's last cell, which its x and y neighbours read
    // as their inflow in this step.
    bool val[kTY], ent[kTY];
    int zgs[kTY], sh[kTY];
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      val[b] = x_ok && kTY * wyi + b < wy && zs[b] >= 0 && es[b] < nge;
      ent[b] = zs[b] == 0 && es[b] < nge;
      const int zc = zs[b] < 0 ? 0 : zs[b];
      zgs[b] = dz > 0 ? zc : nz - 1 - zc;
      sh[b] = ((cst[b] + 7 * (zoff + zgs[b])) & 3) * 4;
    }

    acc4 vacc[kTY];
#pragma unroll
    for (int b = 0; b < kTY; ++b) vacc[b] = acc4{0, 0, 0, 0};

    bool any_ent[kTY];
#pragma unroll
    for (int b = 0; b < kTY; ++b) any_ent[b] = __ballot(ent[b]) != 0;

    // The m2a of angle block a+1 is issued before the solve of block a, so the
    // matrix cores work while the vector units solve.
    acc4 vsc[kTY], vsn[kTY];
    if constexpr (kEarlyM2a) {
#pragma unroll
      for (int b = 0; b < kTY; ++b) vsc[b] = vs0[b];
    } else {
      m2a(0, vin, vsc);
    }
#pragma unroll
    for (int a = 0; a < NAB; ++a) {
      if (a + 1 < NAB) m2a(a + 1 < NAB ? a + 1 : 0, vin, vsn);
      // solve, blocks in descending order: block b's y face is block b-1's
      // value from the previous step, read before it is overwritten
      // y inflow of row 0 of a lower row group: the row group above wrote it
      // into the ring in the previous step
      T yr[4];
      if (NY > 1 && wyi > 0)
        load4(yring + ((((size_t)((t + 1) & 1) * (NY - 1) + (wyi - 1)) * NWA + wa) * NAB + a) *
                          256 + lane * 4,
              yr);
#pragma unroll
      for (int b = kTY - 1; b >= 0; --b) {
        // x inflow of lane (0, g): row (g, r) of block b sits in lane 4b + r
        T xin[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) xin[r] = from_lane(xsc[a], 4 * b + r);
        const T cz = ldexpT((T)0x1p-9, (int)((lut_z >> sh[b]) & 15));
        const T cy = ldexpT((T)0x1p-10, (int)((lut_y >> sh[b]) & 15));
        const int xe = (int)((lut_x >> sh[b]) & 15) - 8;
        T rz[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) rz[r] = F[a][b][r];
        if (any_ent[b]) {  // some lane of this block starts a group
          if (ZB && p.z_in) {  // inflow: the previous z block's last plane
            T zi[4];
            load4(fz_cell(b) + 16 * a, zi);
#pragma unroll
            for (int r = 0; r < 4; ++r) rz[r] = ent[b] ? zi[r] : rz[r];
          } else {
            const T sc = ent[b] ? (T)ldexp((double)(sf_energy(ie0 + es[b]) * su),
                                           sf_space_exp(x, yb[b], zb))
                                : (T)0;
#pragma unroll
            for (int r = 0; r < 4; ++r) rz[r] = ent[b] ? (T)iaf[a][r] * sc : rz[r];
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const T rx = shr1_keep(xin[r], F[a][b][r]);
          const T ry = b > 0 ? F[a][b > 0 ? b - 1 : 0][r]
                             : (NY > 1 && wyi > 0) ? yr[r] : yinn[a][r];
          const T res = fmaT(rx, ldexpT(wz[r % PR::kWz], xe),
                             fmaT(ry, cy, fmaT(rz[r], cz, (T)vsc[b][r])));
          F[a][b][r] = res;
          vsc[b][r] = res;
        }
      }
      // a2m: VO[moment][cell] += M^T[moment][angle] R[angle][cell]
#pragma unroll
      for (int b = 0; b < kTY; ++b) {
        if constexpr (PR::kPacked) {
          T rv[4] = {(T)vsc[b][0], (T)vsc[b][1], (T)vsc[b][2], (T)vsc[b][3]};
          vacc[b] = PR::mma16(mop[a], rv, vacc[b]);
        } else {
#pragma unroll
          for (int s = 0; s < 4; ++s) vacc[b] = PR::mma(mop[a][s], (T)vsc[b][s], vacc[b]);
        }
      }
      if (a + 1 < NAB) {
#pragma unroll
        for (int b = 0; b < kTY; ++b) vsc[b] = vsn[b];
      }
    }
    // the next step's block 0 (its operand was fetched at this step's start;
    // past the last step it runs on clamped data and is never used)
    if constexpr (kEarlyM2a) m2a(0, vnx, vs0);
    if (wyi == 0) {
      bump(pyi, oyi);
      fetch_yin();
    }
    // the last row of this row group is the next row group's y inflow
    if (NY > 1 && wyi < NY - 1) {
#pragma unroll
      for (int a = 0; a < NAB; ++a)
        store4(yring + ((((size_t)(t & 1) * (NY - 1) + wyi) * NWA + wa) * NAB + a) * 256 +
                   lane * 4,
               F[a][kTY - 1]);
    }

    // ---- outflow faces for the downstream tiles ------------------------------
    if (x_out && j == wx - 1) {
#pragma unroll
      for (int b = 0; b < kTY; ++b)
        if (val[b]) {
          T* dst = fyz + (yb[b] * fplane + fo0 - b * nap);
#pragma unroll
          for (int a = 0; a < NAB; ++a) store4(dst + 16 * a, F[a][b]);
        }
    }
    if (y_out) {
#pragma unroll
      for (int b = 0; b < kTY; ++b)
        if (kTY * wyi + b == wy - 1 && val[b]) {
          T* dst = fxz + (fo0 - b * nap);
#pragma unroll
          for (int a = 0; a < NAB; ++a) store4(dst + 16 * a, F[a][b]);
        }
    }
    if (ZB && p.z_out) {  // the block's last plane is the next block's z inflow
#pragma unroll
      for (int b = 0; b < kTY; ++b)
        if (val[b] && zs[b] == nz - 1) {
          T* dst = fz_cell(b);
#pragma unroll
          for (int a = 0; a < NAB; ++a) store4(dst + 16 * a, F[a][b]);
        }
    }

    // ---- vo: lane (j, g) holds moments 4g..4g+3 of cell (j, b) ---------------
    A* rbuf = red + ((size_t)(t & 1) * NY + wyi) * kTY * NWA * 64 * 4;
    if constexpr (NWA > 1) {
#pragma unroll
      for (int b = 0; b < kTY; ++b)
        if (b % NWA != wa) {
          A v[4] = {vacc[b][0], vacc[b][1], vacc[b][2], vacc[b][3]};
          store4(rbuf + ((size_t)(b * NWA + wa) * 64 + lane) * 4, v);
        }
    }
    if constexpr (NWA > 1 || NY > 1) __syncthreads();
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      if (b % NWA != wa) continue;
      // Mixed forms with one angle chunk, before the last visit: the FP32
      // running sum takes one FP32 add, which rounds as the sum in FP64
      // rounded to FP32 does (FP64 holds more than twice FP32's bits).
      if constexpr (!kF64 && NWA == 1) {
        if (!last) {
          if (val[b] && 4 * g < NM) {
            float f[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
              f[r] = first ? (float)vacc[b][r] : (float)vacc[b][r] + (float)von[b][r];
            store4(p.vo32 + sc[b] + 4 * g, f);
          }
          continue;
        }
      }
      double sum[4];
#pragma unroll
      for (int w = 0; w < NWA; ++w) {
        A v[4];
        if (w == wa) {
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = vacc[b][r];
        } else {
          load4(rbuf + ((size_t)(b * NWA + w) * 64 + lane) * 4, v);
        }
        // angle chunks added in chunk order
#pragma unroll
        for (int r = 0; r < 4; ++r) sum[r] = w == 0 ? (double)v[r] : sum[r] + (double)v[r];
      }
      if (!first) {
#pragma unroll
        for (int r = 0; r < 4; ++r) sum[r] += (double)von[b / NWA][r];
      }
      if (val[b] && 4 * g < NM) {
        if (kF64 || last) {
          store4(p.vo + sc[b] + 4 * g, sum);
        } else {
          float f[4] = {(float)sum[0], (float)sum[1], (float)sum[2], (float)sum[3]};
          store4(p.vo32 + sc[b] + 4 * g, f);
        }
      }
    }
    fetch_vo(sn);
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      zs[b] = zn[b];
      es[b] = en[b];
      sc[b] = sn[b];
    }
    fo0 += nap;
  };
  if constexpr (PR::kPingPong) {
    for (int t = 0; t < nsteps; t += 2) {
      step(t, vbA, vbB, xbA, xbB);
      if (t + 1 < nsteps) step(t + 1, vbB, vbA, xbB, xbA);
    }
  } else {
    for (int t = 0; t < nsteps; ++t) {
#pragma unroll
      for (int a = 0; a < NAB; ++a) xbA[a] = xbB[a];
#pragma unroll
      for (int b = 0; b < kTY; ++b)
#pragma unroll
        for (int s = 0; s < KA; ++s) vbA[b][s] = vbB[b][s];
      step(t, vbA, vbB, xbA, xbB);
    }
  }
}

// Workgroup bid of a launch -> (item, chunk of epw energy groups, unknown).
template <class PR, int NM, int NAC, int NWA, int NY, bool ZB = false>
__device__ __forceinline__ void sweep_block(const SweepParams& p, int item_base, int bid,
                                            int epw) {
  const int iu = bid % p.nu;
  bid /= p.nu;
  const int nec = (p.ne + epw - 1) / epw;
  const int ec = bid % nec;
  bid /= nec;
  const int ie0 = ec * epw;
  sweep_tile<PR, NM, NAC, NWA, NY, ZB>(p, p.items[item_base + bid], ie0, min(epw, p.ne - ie0),
                                       iu);
}

// Single sweep (ZB: one z block of it).
template <class PR, int NM, int NAC, int NWA, int NY, bool ZB>
__global__ __launch_bounds__(64 * NWA * NY) __attribute__((amdgpu_waves_per_eu(
    Occupancy<PR, NM, NAC, NWA, NY>::value))) void k_sweep_mfma(SweepParams p, int item_base,
                                                                int epw) {
  sweep_block<PR, NM, NAC, NWA, NY, ZB>(p, item_base, (int)blockIdx.x, epw);
}

// Many sweeps in one launch: segment k of the launch is one launch of the
// schedule of problem prob[k] (its items from item_base[k]); workgroups
// [start[k], start[k+1]) belong to it.
template <class PR, int NM, int NAC, int NWA, int NY>
__global__ __launch_bounds__(64 * NWA * NY) __attribute__((amdgpu_waves_per_eu(
    Occupancy<PR, NM, NAC, NWA, NY>::multi))) void k_sweep_mfma_multi(
    const SweepParams* __restrict__ probs, MultiLaunch ml) {
  const int blk = (int)blockIdx.x;
  int k = 0;
  for (int step = kMaxSegments / 2; step > 0; step >>= 1)
    if (k + step < ml.nseg && ml.start[k + step] <= blk) k += step;
  const SweepParams p = probs[ml.prob[k]];
  sweep_block<PR, NM, NAC, NWA, NY>(p, ml.item_base[k], blk - ml.start[k], ml.epw[k]);
}

// Boundary inflow of every tile row and column: before the sweep, the x face
// plane of each (octant, group, unknown) holds the x-boundary value of every
// (y, z, angle) and the y face plane the y-boundary value of every (x, z,
// angle); the first tile of a row (column) reads them like any other inflow.
// Values are Quantities_init_face's divided by (1 + octant), angles in the
// lane order of the sweep kernel.
// Face row `row` of a sweep: rows [0, 8 nu ny) are (octant, unknown, y) of
// the x planes, the rest (octant, unknown, x) of the y planes; each holds
// ne * nz * na_pad values ordered (group, octant-frame z, lane-order angle).
template <class PR>
__device__ __forceinline__ void fill_row(const SweepParams& p, int row) {
  typedef typename PR::T T;
  const int nap = p.na_pad, nz = p.nz;
  const int nrow_x = kNoct * p.nu * p.ny;
  const bool xface = row < nrow_x;
  int k = xface ? row : row - nrow_x;
  const int ext = xface ? p.ny : p.nx;
  const int c = k % ext;  // y for x faces, x for y faces
  k /= ext;
  const int iu = k % p.nu;
  const int oct = k / p.nu;
  const int dz = (oct & 4) ? -1 : 1;
  const size_t len = (size_t)p.ne * nz * nap;
  T* dst = xface ? reinterpret_cast<T*>(p.fyz) + (size_t)row * len
                 : reinterpret_cast<T*>(p.fxz) + (size_t)(row - nrow_x) * len;
  const int su = sf_unknown(iu);
  // boundary cell coordinates outside the grid: x (x faces) or y (y faces)
  const int bx = xface ? ((oct & 1) ? p.nx : -1) : c;
  const int by = xface ? c : ((oct & 2) ? p.ny : -1);
  // Lanes cover runs of 4 angles (nap is a multiple of 16): with nap/4 <= 64
  // runs per (group, z) row a wave takes 64 / (nap/4) rows per pass, lane
  // (row qo, run r0); wider rows take one row per pass.  The row's scale
  // sf_energy * sf_unknown * 2^space is a power of two, so (1 + ia) 2^(ia & 7)
  // times it is exact, the value Quantities_init_face gives.
  const int nap4 = nap >> 2, nq = p.ne * nz;
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  const int nwave = (int)blockDim.x >> 6;
  const int per = nap4 <= 64 ? 64 / nap4 : 1;
  const int qo = nap4 <= 64 ? lane / nap4 : 0;
  const int r0 = nap4 <= 64 ? lane % nap4 : lane, rstep = nap4 <= 64 ? nap4 : 64;
  if (qo >= per) return;  // lanes past the last whole row of a pass
  for (int q = wave * per + qo; q < nq; q += nwave * per) {
    const int zq = q % nz, ie = q / nz;
    const int z = p.z_off + (dz > 0 ? zq : nz - 1 - zq);  // grid plane
    const double sc = ldexp((double)(sf_energy(ie) * su), sf_space_exp(bx, by, z));
    T* row_dst = dst + (size_t)q * nap;
    for (int r4 = r0; r4 < nap4; r4 += rstep) {
      T v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ia = 16 * (r4 >> 2) + PR::row(r4 & 3, r);
        v[r] = ia < p.na ? (T)((double)((1 + ia) << (ia & 7)) * sc) : (T)0;
      }
      store4(row_dst + 4 * r4, v);
    }
  }
}

template <class PR>
__global__ __launch_bounds__(256) void k_fill_faces(SweepParams p) {
  fill_row<PR>(p, (int)blockIdx.x);
}

// Fill for several sweeps: segment k's workgroups are problem prob[k]'s rows.
template <class PR>
__global__ __launch_bounds__(256) void k_fill_faces_multi(const SweepParams* __restrict__ probs,
                                                          MultiLaunch ml) {
  const int blk = (int)blockIdx.x;
  int k = 0;
  for (int step = kMaxSegments / 2; step > 0; step >>= 1)
    if (k + step < ml.nseg && ml.start[k + step] <= blk) k += step;
  const SweepParams p = probs[ml.prob[k]];
  fill_row<PR>(p, blk - ml.start[k]);
}

// Workgroup shapes built: (waves of angle chunks, row groups of 4 rows), at
// most 4 waves per workgroup.
#define MSW_SHAPES(X) X(1, 1) X(2, 1) X(3, 1) X(4, 1) X(1, 2) X(2, 2) X(1, 4)

bool mfma_shape_supported(int nwa, int ny) {
#define MSW_CASE(W, Y) \
  if (nwa == W && ny == Y) return true;
  MSW_SHAPES(MSW_CASE)
#undef MSW_CASE
  return false;
}

template <class PR, int NM, int NAC>
static void launch_nac(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  const dim3 block(64 * ls.nwa * ls.ny);
#define MSW_CASE(W, Y)                                                                     \
  if (ls.nwa == W && ls.ny == Y) {                                                         \
    if constexpr (std::is_same<PR, PrecF64>::value) {                                     \
      if (sp.fz) {                                                                         \
        hipLaunchKernelGGL((k_sweep_mfma<PR, NM, NAC, W, Y, true>), dim3(ls.grid), block,   \
                           ls.lds_bytes, st, sp, ls.item_base, ls.epw);                   \
        return;                                                                            \
      }                                                                                    \
    }                                                                                      \
    hipLaunchKernelGGL((k_sweep_mfma<PR, NM, NAC, W, Y, false>), dim3(ls.grid), block,      \
                       ls.lds_bytes, st, sp, ls.item_base, ls.epw);                       \
    return;                                                                                \
  }
  MSW_SHAPES(MSW_CASE)
#undef MSW_CASE
}

template <class PR, int NM>
static void launch_nm(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  if (ls.nac == 16)
    launch_nac<PR, NM, 16>(ls, sp, st);
  else if (ls.nac == 32)
    launch_nac<PR, NM, 32>(ls, sp, st);
  else
    launch_nac<PR, NM, 64>(ls, sp, st);
}

template <class PR>
static void launch_prec(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  switch (ls.nm) {
    case 4: launch_nm<PR, 4>(ls, sp, st); break;
    case 8: launch_nm<PR, 8>(ls, sp, st); break;
    default: launch_nm<PR, 16>(ls, sp, st); break;
  }
}

void launch_mfma(int precision, const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  if (precision == MSW_PREC_MIXED_FP32)
    launch_prec<PrecF32>(ls, sp, st);
  else if (precision == MSW_PREC_MIXED_BF16)
    launch_prec<PrecBF16>(ls, sp, st);
  else
    launch_prec<PrecF64>(ls, sp, st);
}

void launch_mfma_fill(int precision, const SweepParams& sp, hipStream_t st) {
  const unsigned blocks = (unsigned)mfma_fill_rows(sp);
  if (precision == MSW_PREC_FP64)
    hipLaunchKernelGGL(k_fill_faces<PrecF64>, dim3(blocks), dim3(256), 0, st, sp);
  else
    hipLaunchKernelGGL(k_fill_faces<PrecF32>, dim3(blocks), dim3(256), 0, st, sp);
}

template <class PR, int NM, int NAC>
static void launch_multi_nac(int nwa, int ny, unsigned grid, size_t lds, const SweepParams* probs,
                             const MultiLaunch& ml, hipStream_t st) {
  const dim3 block(64 * nwa * ny);
#define MSW_CASE(W, Y)                                                                     \
  if (nwa == W && ny == Y) {                                                               \
    hipLaunchKernelGGL((k_sweep_mfma_multi<PR, NM, NAC, W, Y>), dim3(grid), block, lds, st, \
                       probs, ml);                                                         \
    return;                                                                                \
  }
  MSW_SHAPES(MSW_CASE)
#undef MSW_CASE
}

template <class PR, int NM>
static void launch_multi_nm(int nac, int nwa, int ny, unsigned grid, size_t lds,
                            const SweepParams* probs, const MultiLaunch& ml, hipStream_t st) {
  if (nac == 16)
    launch_multi_nac<PR, NM, 16>(nwa, ny, grid, lds, probs, ml, st);
  else if (nac == 32)
    launch_multi_nac<PR, NM, 32>(nwa, ny, grid, lds, probs, ml, st);
  else
    launch_multi_nac<PR, NM, 64>(nwa, ny, grid, lds, probs, ml, st);
}

template <class PR>
static void launch_multi_prec(int nm, int nac, int nwa, int ny, unsigned grid, size_t lds,
                              const SweepParams* probs, const MultiLaunch& ml, hipStream_t st) {
  switch (nm) {
    case 4: launch_multi_nm<PR, 4>(nac, nwa, ny, grid, lds, probs, ml, st); break;
    case 8: launch_multi_nm<PR, 8>(nac, nwa, ny, grid, lds, probs, ml, st); break;
    default: launch_multi_nm<PR, 16>(nac, nwa, ny, grid, lds, probs, ml, st); break;
  }
}

void launch_mfma_multi(int precision, int nm, int nac, int nwa, int ny, unsigned grid,
                       const SweepParams* probs, const MultiLaunch& ml, hipStream_t st) {
  const size_t lds = mfma_lds_bytes(precision, nac, nwa, ny);
  if (precision == MSW_PREC_MIXED_FP32)
    launch_multi_prec<PrecF32>(nm, nac, nwa, ny, grid, lds, probs, ml, st);
  else if (precision == MSW_PREC_MIXED_BF16)
    launch_multi_prec<PrecBF16>(nm, nac, nwa, ny, grid, lds, probs, ml, st);
  else
    launch_multi_prec<PrecF64>(nm, nac, nwa, ny, grid, lds, probs, ml, st);
}

void launch_mfma_fill_multi(int precision, unsigned grid, const SweepParams* probs,
                            const MultiLaunch& ml, hipStream_t st) {
  if (precision == MSW_PREC_FP64)
    hipLaunchKernelGGL(k_fill_faces_multi<PrecF64>, dim3(grid), dim3(256), 0, st, probs, ml);
  else
    hipLaunchKernelGGL(k_fill_faces_multi<PrecF32>, dim3(grid), dim3(256), 0, st, probs, ml);
}

int mfma_fill_rows(const SweepParams& sp) { return kNoct * sp.nu * (sp.nx + sp.ny); }

size_t mfma_face_elem_bytes(int precision) {
  return precision == MSW_PREC_FP64 ? sizeof(double) : sizeof(float);
}

size_t mfma_lds_bytes(int precision, int nac, int nwa, int ny) {
  const size_t t = mfma_face_elem_bytes(precision);  // accumulator = face type
  const size_t nab = nac / 16;
  const size_t red = nwa > 1 ? 2 * (size_t)ny * kTY * nwa * 64 * 4 * t : 0;
  const size_t ring = ny > 1 ? 2 * (size_t)(ny - 1) * nwa * nab * 64 * 4 * t : 0;
  return red + ring;
}

bool mfma_face_offsets_fit(const MswDims& d, int na_pad) {
  const long long ext = d.ncell_x > d.ncell_y ? d.ncell_x : d.ncell_y;
  return ext * d.ne * d.ncell_z * na_pad < (1ll << 31);
}

int mfma_tile_x() { return kTX; }
int mfma_tile_rows() { return kTY; }

}  // namespace msw
