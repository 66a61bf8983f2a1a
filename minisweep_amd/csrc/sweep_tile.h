// sweep_tile.h -- the tile sweep on the gfx950 matrix cores, in FP64 (plain)
// and in mixed precision (FP32 or BF16 contractions with FP32 accumulation,
// FP32 faces and solve, FP64 input state; vo summed over the octants in FP32
// and written in FP64 by each tile's last octant visit).  Device templates
// only: the per-precision files instantiate them.
//
// One wave owns 16 (x) by 4 (y) cells for one (octant, run of energy groups,
// unknown, chunk of NAC angles) and walks the z extent of every group of the
// run as one stream.  Lane l = j + 16g owns cell (x = j, y = b) of every cell
// block b = 0..3; at step t that cell is at stream position t - j - b, so the
// 64 cells of a step lie on one diagonal hyperplane.  A workgroup is NWA
// angle-chunk waves times NY row groups (a 16 x 4NY tile); row group w runs
// 4w steps behind row group w-1 and takes its y inflow from it through LDS.
//
//   * Carried values.  The reference's solve (quantities_testing_kernels.h:
//     380-446) is r = (vs/S + (Fz/2/Sz + Fy/4/Sy + w(a) Fx/Sx)/(1+o)) S,
//     with the faces holding r (1+o) of the upwind cell and S = 2^((x + 3y +
//     7z + 2) & 3) the cell's spatial scale.  The (1+o) factors cancel.  With
//     K = S E(e) U(u) the cell's scale (E, U the group's and the unknown's
//     factors, equal for a cell and its upwind neighbours) and q = r/K carried
//     instead of r, the solve is
//        q = vs/K + q_z/2 + q_y/4 + w(a) q_x
//     with coefficients that do not depend on the cell.  K is a power of two,
//     so scaling by it is exact and commutes with every rounding: the m2a
//     operand is divided by K (the mixed forms' operand copies come divided),
//     the a2m result multiplied by K, and the bits are those of the solve on
//     r.  A boundary face value, the reference's (1+a) 2^(a&7) S_out E U
//     (1+o), becomes (1+a) 2^(a&7): a constant of the lane's angles, so the
//     domain's edge tiles and each group's first z plane take it from
//     registers.
//   * m2a: VS[angle][cell] = AT[angle][moment] . VI[moment][cell] as 16x16
//     MFMAs.  AT is the A operand (per octant, in registers for the whole
//     launch), VI the B operand, loaded from the state with each lane
//     reading consecutive moments of its own cell.  The FP32 form adds the
//     BF16 rounding of what FP32 drops of vi (one BF16 MFMA on the operand
//     copy's second part), so vi enters to 2^-33 instead of 2^-24.
//   * The MFMA result puts angle row(g, r) of the 16-angle block in lane
//     (j, g), register r.  The solve runs in that layout: the z face is the
//     lane's own value from the previous step, the y face block b-1's value of
//     the same lane, the x face lane j-1's (DPP row shift).  Lane j = 0 gets
//     the tile's x inflow (loaded one value per lane of the row, gathered by
//     a DPP row rotation) and block 0 the y inflow, read from the face planes
//     one step ahead, or formed from the boundary value at the domain's edge.
//   * a2m: VO[moment][cell] = M^T[moment][angle] . Q[angle][cell], with the
//     solved values as the B operand straight from the registers they sit in
//     and M^T's rows ordered so that lane (j, g) ends with moments 4g..4g+3 of
//     cell j: vo is read and written as 32-byte runs.
//   * The angle chunks of a tile are the waves of one workgroup; each wave
//     stores vo for the cell blocks it owns after adding the other waves'
//     partial sums (LDS, double-buffered by step parity, fixed order).
//   * Face planes in HBM hold, per (octant, group, unknown), one yz plane for
//     x faces and one xz plane for y faces, angles in lane order.  The host
//     schedule never runs two tiles of one octant row (column) in one launch,
//     so a tile reads its inflow and writes its outflow in place.
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
#pragma once

#include <hip/hip_runtime.h>

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
__device__ __forceinline__ short bf16_bits(float v) { return __builtin_bit_cast(short, (__bf16)v); }

// Four BF16 values as the MFMA operand, kept two to a 32-bit register.
typedef unsigned bf16x4[2];

// Up to 4 BF16 values (n = 1, 2 or 4; 2n-byte aligned) into elements 0..n-1,
// zeros after.
template <int N>
__device__ __forceinline__ void load_bf16(const unsigned short* s, bf16x4& v) {
  if constexpr (N == 4) {
    const uint2 w = *reinterpret_cast<const uint2*>(s);
    v[0] = w.x;
    v[1] = w.y;
  } else if constexpr (N == 2) {
    v[0] = *reinterpret_cast<const unsigned*>(s);
    v[1] = 0;
  } else {
    v[0] = s[0];
    v[1] = 0;
  }
}

// ---- precision traits ------------------------------------------------------
// T: face / solve type; A: MFMA accumulator element; row(g, r): the angle
// (within a 16-angle block) that MFMA result register r of lane group g
// holds; mom(i): the moment that a2m result row i stands for, chosen so that
// lane group g, register r holds moment 4g + r.  kPingPong: the z loop
// alternates two operand buffers (two steps per iteration) instead of copying
// the prefetched operands at the top of each step; measured faster for the
// mixed forms, slower for FP64 (register pressure).  kScaleVi: the m2a
// operand is divided by the cell's scale in the kernel (FP64) rather than in
// an operand copy made before the sweep (mixed forms).
struct PrecF64 {
  typedef double T;
  typedef double A;
  typedef d4v acc4;
  static constexpr bool kPacked = false;
  static constexpr bool kSplit = false;
  static constexpr bool kPingPong = false;
  static constexpr bool kScaleVi = true;
  static constexpr int kValuPerMfma = 6;  // vector instructions issued per matrix one
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
  typedef f4v acc4;
  static constexpr bool kPacked = false;
  static constexpr bool kSplit = true;  // + BF16 MFMA on the rest of vi
  static constexpr bool kPingPong = true;
  static constexpr bool kScaleVi = false;
  static constexpr int kValuPerMfma = 6;
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
  typedef f4v acc4;
  static constexpr bool kPacked = true;  // K = 16: 4 values per lane per MFMA
  static constexpr bool kSplit = false;
  static constexpr bool kPingPong = true;
  static constexpr bool kScaleVi = false;
  static constexpr int kValuPerMfma = 16;
  static constexpr int kWz = 4;
  static __device__ __forceinline__ constexpr int row(int g, int r) { return 4 * g + r; }
  static __device__ __forceinline__ constexpr int mom(int i) { return i; }
};

// BF16 16x16x16 MFMA, A given as FP32 values (converted; exact for the
// octant matrices, whose entries are integers below 256 times powers of two),
// B as BF16 bits.
__device__ __forceinline__ s4v bf16_pack(const float (&a)[4]) {
  s4v p;
#pragma unroll
  for (int e = 0; e < 4; ++e) p[e] = bf16_bits(a[e]);
  return p;
}
__device__ __forceinline__ f4v mma_bf16(const float (&a)[4], const bf16x4& b, f4v c) {
  const uint2 w = {b[0], b[1]};
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16_pack(a), __builtin_bit_cast(s4v, w), c,
                                                    0, 0, 0);
}
__device__ __forceinline__ f4v mma_bf16(const float (&a)[4], const float (&b)[4], f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf16_pack(a), bf16_pack(b), c, 0, 0, 0);
}

// The m2a B operand of one step, for the 4 cell blocks of a wave: FP64 vi
// (KA = nm/4 values per lane and block), FP32 vi and the BF16 rest (FP32
// form), or BF16 vi (4 values, BF16 form).
template <class PR, int KA>
struct VinBuf {
  typename PR::T v[kTY][KA];
};
template <int KA>
struct VinBuf<PrecF32, KA> {
  float v[kTY][KA];
  bf16x4 lo[kTY];
};
template <int KA>
struct VinBuf<PrecBF16, KA> {
  bf16x4 v[kTY];
};

// Waves per SIMD the register allocation aims for: two where the working set
// fits in 256 registers without spilling to scratch, so that one wave's matrix
// work overlaps the other's solve and loads; else one (the compiler then
// keeps the overflow in accumulation registers).  The shapes that fit were
// read off the compiler's resource report for every instantiation.
template <class PR, int NM, int NAC, int NWA, int NY>
struct Occupancy {
  static constexpr int bytes = (int)sizeof(typename PR::T) * NAC;
  static constexpr bool rows_ok = NY == 1 || (NWA == 2 && NY == 2 && NM < 16);
  // FP32 form, 8 moments, one angle wave and one row group: 2 registers past 256
  static constexpr bool split_ok = !PR::kSplit || NM < 8 || NWA > 1 || NY > 1;
  static constexpr int value =
      bytes <= 64                                                          ? 2
      : bytes <= 128 && sizeof(typename PR::T) == 4 && rows_ok && split_ok &&
              (NM < 16 || PR::kPacked)                                     ? 2
      : bytes <= 128 && NWA >= 2 && NY == 1 && (NM < 16 || sizeof(typename PR::T) == 8)
                                                                           ? 2
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
  // the tile's upwind x edge (y edge) is the domain's: its inflow is the
  // boundary value, formed here rather than read
  const bool x_bnd = it.tx == 0;
  const bool y_bnd = it.ty == 0 && wyi == 0;

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
  // the FP32 form's second m2a part: the A operand as BF16 (lane group g's
  // moments in elements 0..KM-1, zeros after)
  float aopb[NAB][4];
  if constexpr (PR::kSplit) {
#pragma unroll
    for (int a = 0; a < NAB; ++a)
#pragma unroll
      for (int e = 0; e < 4; ++e) aopb[a][e] = e < KA ? (float)aop[a][e < KA ? e : 0] : 0.0f;
  }
  // exponents of the group factors E of the (at most kMaxStreamGroups)
  // groups streamed, 2 bits each; + the unknown's: the scale K of a cell is
  // 2^(space + group + unknown exponent)
  unsigned long long emask = 0;
  for (int k = 0; k < nge; ++k)
    emask |= (unsigned long long)sf_energy_exp(p.e_off + ie0 + k) << (2 * k);
  const int su = sf_unknown_exp(iu);

  T F[NAB][kTY][4];  // outgoing face value (= solved value q) of the lane's cells
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
  VinBuf<PR, KA> vbA, vbB;  // m2a B operand, one step's while the next's loads are in flight
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
  // exponent of the scale K of block b's cell at stream position (z, e),
  // clamped like the addresses
  auto scale_exp = [&](int b, int z, int e) {
    const int zc = z < 0 ? 0 : z;
    const int zg = dz > 0 ? zc : nz - 1 - zc;
    const int ec = e < nge ? e : nge - 1;
    return ((cst[b] + 7 * (zoff + zg)) & 3) + su + (int)((emask >> (2 * ec)) & 3);
  };
  int exs[kTY];          // this step's scale exponents
  int zn[kTY], en[kTY];  // next step's stream position per block
  int exn[kTY];          // next step's scale exponents
  size_t sn[kTY];        // next step's state offsets
  auto advance = [&]() {
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      const bool wrap = zs[b] + 1 == nz;
      zn[b] = wrap ? 0 : zs[b] + 1;
      en[b] = wrap ? es[b] + 1 : es[b];
      const bool move = zs[b] >= 0 && en[b] < nge;
      sn[b] = sc[b] + (size_t)(move ? (wrap ? dW : dS) : 0ll);
      exn[b] = scale_exp(b, zn[b], en[b]);
    }
  };
  // FP64: the m2a operand divided by the cell's scale in place, once it has
  // arrived (the copies of the mixed forms come divided)
  auto scale_vi = [&](VinBuf<PR, KA>& v, const int (&ex)[kTY]) {
    if constexpr (PR::kScaleVi) {
#pragma unroll
      for (int b = 0; b < kTY; ++b)
#pragma unroll
        for (int s = 0; s < KA; ++s) v.v[b][s] = ldexpT(v.v[b][s], -ex[b]);
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
  // the m2a operand for the cells at offsets so
  auto fetch_vi = [&](const size_t (&so)[kTY], VinBuf<PR, KA>& vnx) {
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      if constexpr (PR::kPacked) {  // BF16 form: 4 values of a lane in 8 bytes
        if (!mpad) {
          load_bf16<4>(p.vib + so[b] + moff, vnx.v[b]);
        } else {
          vnx.v[b][0] = 0;
          vnx.v[b][1] = 0;
        }
      } else if constexpr (PR::kSplit) {  // FP32 form: FP32 part and BF16 rest
        const float* src = p.vhi + so[b] + moff;
        if constexpr (KA == 4) {
          load4(src, vnx.v[b]);
        } else {
#pragma unroll
          for (int s = 0; s < KA; ++s) vnx.v[b][s] = src[s];
        }
        load_bf16<KA>(p.vlo + so[b] + moff, vnx.lo[b]);
      } else if constexpr (KA == 4) {
        load4(p.vi + so[b] + moff, vnx.v[b]);
      } else {
        const double* src = p.vi + so[b] + moff;
#pragma unroll
        for (int s = 0; s < KA; ++s) vnx.v[b][s] = src[s];
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
  // The inflow loads are issued on every path, at the domain's edge too (they
  // read the planes at clamped offsets and are not used there), so that each
  // wait covers only the data it needs; the edge takes the boundary constant
  // of the lane's angles instead, by a register select at the use: spread for
  // the x gather (lane (j, g) holds the value of row (g, j % 4)).
  auto fetch_xs = [&](T (&xsn)[NAB]) {
#pragma unroll
    for (int a = 0; a < NAB; ++a) xsn[a] = xs_src[oxs + 16 * a];
  };
  auto fetch_yin = [&]() {  // row group 0
#pragma unroll
    for (int a = 0; a < NAB; ++a) load4(fxz + oyi + 16 * a, yinn[a]);
  };
  T xbnd[NAB];
#pragma unroll
  for (int a = 0; a < NAB; ++a) {
    const int ia = ia0 + 16 * a + PR::row(g, j & 3);
    xbnd[a] = ia < p.na ? (T)((1 + ia) << (ia & 7)) : (T)0;
  }

  // m2a of angle block a: VS[angle][cell] = AT . VI (VI divided by the cell's
  // scale K: in the operand copies, or in place at the step's start in the
  // FP64 form)
  auto m2a = [&](int a, const VinBuf<PR, KA>& vin, acc4 (&vs)[kTY]) {
#pragma unroll
    for (int b = kTY - 1; b >= 0; --b) {
      vs[b] = acc4{0, 0, 0, 0};
      if constexpr (PR::kPacked) {
        vs[b] = mma_bf16(aop[a], vin.v[b], vs[b]);
      } else if constexpr (PR::kSplit) {
        vs[b] = mma_bf16(aopb[a], vin.lo[b], vs[b]);
#pragma unroll
        for (int s = 0; s < KM; ++s) vs[b] = PR::mma(aop[a][s], vin.v[b][s], vs[b]);
      } else {
#pragma unroll
        for (int s = 0; s < KM; ++s) vs[b] = PR::mma(aop[a][s], vin.v[b][s], vs[b]);
      }
    }
  };

  fetch_xs(PR::kPingPong ? xbA : xbB);
  fetch_vi(sc, PR::kPingPong ? vbA : vbB);
  if (wyi == 0) fetch_yin();
  fetch_vo(sc);
#pragma unroll
  for (int b = 0; b < kTY; ++b) exs[b] = scale_exp(b, zs[b], es[b]);
  scale_vi(PR::kPingPong ? vbA : vbB, exs);
  // Mixed forms: m2a of angle block 0 for the coming step is issued at the
  // end of the step before, so the matrix cores run it during that step's vo
  // reduction (measured at 64^3: FP32 kernel 48.1 -> 46.4 ms, BF16 41.6 ->
  // 40.0 ms when it was introduced; FP64 runs slower that way, 120 vs 90 ms)
  constexpr bool kEarlyM2a = !kF64;
  acc4 vs0[kTY];
  if constexpr (kEarlyM2a) m2a(0, PR::kPingPong ? vbA : vbB, vs0);

  // One step: this step's operands are vin and xsc, the next step's are
  // fetched into vnx and xsn first thing.
  auto step = [&](const int t, VinBuf<PR, KA>& vin, VinBuf<PR, KA>& vnx, const T (&xsc)[NAB],
                  T (&xsn)[NAB]) {
    advance();
    fetch_vi(sn, vnx);
    bump(pxs, oxs);
    fetch_xs(xsn);

    // ---- per cell block: validity, z, scale K; z inflow on entry -----------
    // A lane whose cell starts a group (z = 0) takes the domain-boundary z
    // inflow instead of its own face register; the register itself still
    // holds the previous group's last cell, which its x and y neighbours read
    // as their inflow in this step.
    bool val[kTY], ent[kTY];
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      val[b] = x_ok && kTY * wyi + b < wy && zs[b] >= 0 && es[b] < nge;
      ent[b] = zs[b] == 0 && es[b] < nge;
    }

    acc4 vacc[kTY];
#pragma unroll
    for (int b = 0; b < kTY; ++b) vacc[b] = acc4{0, 0, 0, 0};

    bool any_ent[kTY];
    if constexpr (ZB) {
#pragma unroll
      for (int b = 0; b < kTY; ++b) any_ent[b] = __ballot(ent[b]) != 0;
    }

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
      // the edge's boundary constants in place of the loaded inflow
      const T xsel = x_bnd ? xbnd[a] : xsc[a];
      T ysel[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) ysel[r] = y_bnd ? (T)iaf[a][r] : yinn[a][r];
#pragma unroll
      for (int b = kTY - 1; b >= 0; --b) {
        // x inflow of lane (0, g): row (g, r) of block b sits in lane 4b + r
        T xin[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) xin[r] = from_lane(xsel, 4 * b + r);
        // z inflow: the lane's own value, or the boundary constant where its
        // cell starts a group (a select, so that the solve stays one
        // straight-line region the scheduler can interleave with the matrix
        // work); z-blocked kernels read the previous block's last plane
        T rz[4];
        if constexpr (ZB) {
#pragma unroll
          for (int r = 0; r < 4; ++r) rz[r] = F[a][b][r];
          if (any_ent[b]) {
            if (p.z_in) {
              T zi[4];
              load4(fz_cell(b) + 16 * a, zi);
#pragma unroll
              for (int r = 0; r < 4; ++r) rz[r] = ent[b] ? zi[r] : rz[r];
            } else {
#pragma unroll
              for (int r = 0; r < 4; ++r) rz[r] = ent[b] ? (T)iaf[a][r] : rz[r];
            }
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) rz[r] = ent[b] ? (T)iaf[a][r] : F[a][b][r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const T rx = shr1_keep(xin[r], F[a][b][r]);
          const T ry = b > 0 ? F[a][b > 0 ? b - 1 : 0][r] : (NY > 1 && wyi > 0) ? yr[r] : ysel[r];
          const T res = fmaT(rx, wz[r % PR::kWz], fmaT(ry, (T)0.25, fmaT(rz[r], (T)0.5,
                                                                          (T)vsc[b][r])));
          F[a][b][r] = res;
          vsc[b][r] = res;
        }
      }
      // a2m: VO[moment][cell] += M^T[moment][angle] Q[angle][cell]
#pragma unroll
      for (int b = 0; b < kTY; ++b) {
        if constexpr (PR::kPacked) {
          T rv[4] = {(T)vsc[b][0], (T)vsc[b][1], (T)vsc[b][2], (T)vsc[b][3]};
          vacc[b] = mma_bf16(mop[a], rv, vacc[b]);
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
    if constexpr (kEarlyM2a) {
      scale_vi(vnx, exn);
      m2a(0, vnx, vs0);
    }
    // Instruction order of the step so far: each matrix instruction followed
    // by a few vector ones, so that the matrix cores run while the wave
    // issues the solve (a wave issues in order; a run of matrix instructions
    // would hold the vector work back for as long as they take).
    constexpr int kMfmaPerStep =
        PR::kPacked ? 2 * kTY * NAB : kTY * NAB * (KM + 4 + (PR::kSplit ? 1 : 0));
#pragma unroll
    for (int i = 0; i < kMfmaPerStep; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, PR::kValuPerMfma, 0);
    }
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
    // the domain's downstream faces in the reference's layouts, on request:
    // the solved value r = q K (K = 2^exs) times (1+octant), per angle
    if constexpr (!ZB) {
      if (p.oyz) {
        auto put = [&](double* base, const int b) {
#pragma unroll
          for (int a = 0; a < NAB; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int ia = ia0 + 16 * a + PR::row(g, r);
              if (ia < p.na)
                base[ia] = ldexpT((double)F[a][b][r], exs[b]) * (double)(1 + oct);
            }
        };
#pragma unroll
        for (int b = 0; b < kTY; ++b) {
          if (!val[b]) continue;
          const int ie = ie0 + es[b];
          const int zg = dz > 0 ? zs[b] : nz - 1 - zs[b];
          const size_t oe = (size_t)oct * p.ne + ie;
          if (!x_out && j == wx - 1)
            put(p.oyz + (((oe * nz + zg) * ny + yb[b]) * nu + iu) * p.na, b);
          if (!y_out && kTY * wyi + b == wy - 1)
            put(p.oxz + (((oe * nz + zg) * nx + x) * nu + iu) * p.na, b);
          if (zs[b] == nz - 1)
            put(p.oxy + (((oe * ny + yb[b]) * nx + x) * nu + iu) * p.na, b);
        }
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

    // ---- vo: lane (j, g) holds moments 4g..4g+3 of cell (j, b), in units of
    // the cell's scale K (times 2^exs[b] below) -------------------------------
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
            for (int r = 0; r < 4; ++r) {
              const float v = ldexpT((float)vacc[b][r], exs[b]);
              f[r] = first ? v : v + (float)von[b][r];
            }
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
#pragma unroll
      for (int r = 0; r < 4; ++r) sum[r] = ldexpT(sum[r], exs[b]);
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
    if constexpr (!kEarlyM2a) scale_vi(vnx, exn);
#pragma unroll
    for (int b = 0; b < kTY; ++b) {
      zs[b] = zn[b];
      es[b] = en[b];
      exs[b] = exn[b];
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
      vbA = vbB;
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

// Workgroup ids and the 8 XCDs.  The dispatcher deals workgroups to the XCDs
// round-robin by id, and each XCD has its own L2.  The nu unknowns of one
// (item, group chunk) G read and write the same cache lines (a cell's
// unknowns are adjacent in the state and in the mixed forms' FP32 and BF16
// copies), so they get ids on one XCD: id = ((G / 8) nu + iu) 8 + G % 8.  A
// launch of ng groups is padded to xcd_round_ids(ng, nu) ids; the ids past the
// last group exit at once.  Returns the (G nu + iu) index sweep_block takes,
// or -1.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_unmap(int id, int nu, int ngroups) {
  const int k = id / kXcds;
  const int g = (k / nu) * kXcds + id % kXcds;
  return g < ngroups ? g * nu + k % nu : -1;
}

// Single sweep (ZB: one z block of it).
template <class PR, int NM, int NAC, int NWA, int NY, bool ZB>
__global__ __launch_bounds__(64 * NWA * NY) __attribute__((amdgpu_waves_per_eu(
    Occupancy<PR, NM, NAC, NWA, NY>::value))) void k_sweep_mfma(SweepParams p, int item_base,
                                                                int epw, int ngroups) {
  const int w = xcd_unmap((int)blockIdx.x, p.nu, ngroups);
  if (w < 0) return;
  sweep_block<PR, NM, NAC, NWA, NY, ZB>(p, item_base, w, epw);
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
  const int epw = ml.epw_ng[k] & kEpwMask;
  const int local = blk - ml.start[k];
  const int w = ml.epw_ng[k] & kXcdFlag
                    ? xcd_unmap(local, p.nu, ml.epw_ng[k] >> (kEpwBits + 1))
                    : local;
  if (w < 0) return;
  sweep_block<PR, NM, NAC, NWA, NY>(p, ml.item_base[k], w, epw);
}

// Workgroup shapes built: (waves of angle chunks, row groups of 4 rows), at
// most 4 waves per workgroup.
#define MSW_SHAPES(X) X(1, 1) X(2, 1) X(3, 1) X(4, 1) X(1, 2) X(2, 2) X(1, 4)

template <class PR, int NM, int NAC>
static void launch_nac(const LaunchSpec& ls, const SweepParams& sp, hipStream_t st) {
  const dim3 block(64 * ls.nwa * ls.ny);
#define MSW_CASE(W, Y)                                                                     \
  if (ls.nwa == W && ls.ny == Y) {                                                         \
    if constexpr (std::is_same<PR, PrecF64>::value) {                                     \
      if (sp.fz) {                                                                         \
        hipLaunchKernelGGL((k_sweep_mfma<PR, NM, NAC, W, Y, true>), dim3(ls.grid), block,   \
                           ls.lds_bytes, st, sp, ls.item_base, ls.epw, ls.ngroups);       \
        return;                                                                            \
      }                                                                                    \
    }                                                                                      \
    hipLaunchKernelGGL((k_sweep_mfma<PR, NM, NAC, W, Y, false>), dim3(ls.grid), block,      \
                       ls.lds_bytes, st, sp, ls.item_base, ls.epw, ls.ngroups);           \
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

}  // namespace msw
