// plan_host.cpp -- the host-only parts of the sweep library: error text, the
// default kernel shapes, MFMA tile geometry, the tile schedule and the
// manufactured problem's matrices.  Plain C++ (no device code), so that the
// plan and batch host logic also builds and runs without a GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "msw_device.h"
#include "msw_internal.h"
#include "../../include/msw_plan.h"

namespace msw {

namespace {
constexpr int kTileX = 16;    // tile cells along x: one 16-lane row
constexpr int kTileRows = 4;  // cell rows per row group
}  // namespace

static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int set_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Energy groups streamed through one MFMA tile pipeline: about 128 stream
// steps per workgroup, so the 18-step fill and drain of a tile cost ~12%
// rather than 18/(nz+18).
int default_mfma_epw(int nz, int ne) {
  const int e = std::min((128 + nz / 2) / nz, kMaxStreamGroups);
  return e < 1 ? 1 : e > ne ? ne : e;
}

// Row groups of 4 rows per MFMA tile.  Measured at 64^3, na=64: FP64 (32
// angles per wave, two waves per tile) is fastest with one row group (the
// per-step barrier of four waves costs more than the y-face traffic it
// saves); the mixed forms (64 angles per wave) with two.  Never more rows
// than the grid has.
int default_mfma_ny(int precision, int nwa, int ncell_y) {
  const int want = precision == MSW_PREC_FP64 ? 1 : 2;
  for (int ny = want; ny > 1; ny /= 2)
    if (mfma_shape_supported(nwa, ny) && 4 * ny <= ((ncell_y + 3) / 4) * 4) return ny;
  return 1;
}

// Energy groups per workgroup for one launch of waves1 waves per group: the
// preferred count, lowered while the launch would have fewer waves than the
// target (16 per SIMD for FP64, 8 for the mixed forms, whose longer streams
// pay off more: measured at 64^3).
int launch_mfma_epw(int preferred, int ne, long long waves1, int precision) {
  const long long target = precision == MSW_PREC_FP64 ? 4096 : 2048;
  int e = std::min(preferred < 1 ? 1 : preferred > ne ? ne : preferred, kMaxStreamGroups);
  while (e > 1 && waves1 * ((ne + e - 1) / e) < target) --e;
  return e;
}

// Default angles per wave of the MFMA kernels: measured at 64^3, na=64, 32
// per wave is fastest in FP64 (16 loses more to the cross-wave vo reduction
// than it gains in occupancy) and 64 in the mixed forms; never more than na
// rounded up to a 16-angle block, and at most 4 waves per tile.
int default_mfma_nac(int precision, const MswDims& d, bool alone) {
  const int na = d.na;
  const int cap = na <= 16 ? 16 : na <= 32 ? 32 : 64;
  int nac = precision == MSW_PREC_FP64 ? 32 : 64;
  // FP64 with at most 8192 (tile, group, unknown) streams, its launches run
  // alone: 16 angles per wave, two waves per SIMD.  The extra operand loads
  // and the reduction over more angle waves cost less than the second wave
  // gains until the launches hold enough streams to hide latency at one wave
  // (one MI355X, ms per sweep at 32 / 16 angles: 64^3 ne64 16384 streams
  // 91.9 / 94.2; ne32 51.0 / 49.2; ne16 27.8 / 26.0; 48^3 ne32 23.4 / 22.0;
  // 32^3 ne16 5.0 / 4.4; 16^3 ne8 1.44 / 1.03; the 256 uneven problems of
  // the batch workload one after another 410 / 471 problems/s).  In launches
  // combined with other problems' (batch ticks) the waves are there already:
  // the 256 problems batched in FP64 run 743 / 598 problems/s.
  if (precision == MSW_PREC_FP64 && alone) {
    const long long tiles = (long long)((d.ncell_x + mfma_tile_x() - 1) / mfma_tile_x()) *
                            ((d.ncell_y + mfma_tile_rows() - 1) / mfma_tile_rows());
    if (tiles * d.ne * d.nu <= 8192) nac = 16;
  }
  nac = nac < cap ? nac : cap;
  while ((na + nac - 1) / nac > 4 && nac < 64) nac *= 2;
  return nac;
}

bool mfma_shape_supported(int nwa, int ny) {
  return (ny == 1 && nwa >= 1 && nwa <= 4) || (ny == 2 && (nwa == 1 || nwa == 2)) ||
         (ny == 4 && nwa == 1);
}

size_t mfma_face_elem_bytes(int precision) {
  return precision == MSW_PREC_FP64 ? sizeof(double) : sizeof(float);
}

size_t mfma_lds_bytes(int precision, int nac, int nwa, int ny) {
  const size_t t = mfma_face_elem_bytes(precision);  // accumulator = face type
  const size_t nab = nac / 16;
  const size_t red = nwa > 1 ? 2 * (size_t)ny * kTileRows * nwa * 64 * 4 * t : 0;
  const size_t ring = ny > 1 ? 2 * (size_t)(ny - 1) * nwa * nab * 64 * 4 * t : 0;
  return red + ring;
}

bool mfma_face_offsets_fit(const MswDims& d, int na_pad) {
  const long long ext = d.ncell_x > d.ncell_y ? d.ncell_x : d.ncell_y;
  return ext * d.ne * d.ncell_z * na_pad < (1ll << 31);
}

int mfma_tile_x() { return kTileX; }
int mfma_tile_rows() { return kTileRows; }

// ---------------------------------------------------------------------------
// Host-side schedule: list scheduling of (octant, tile) visits.
//
// Octant o's tile (i, j) in its own frame needs (i-1, j) and (i, j-1) done.
// A launch takes every ready visit whose physical tile is not yet taken in
// that launch.  Octant priority: kind 0, most remaining visits first; kind 1,
// the reference's fixed octant order {0,4,2,6,3,7,1,5}
// (src/3_sweeper/stepscheduler_kba.c:155).  Returns the launches in order;
// marks the first and the last visit of every physical tile.
std::vector<std::vector<TileItem>> build_schedule(int ntx, int nty, int kind, unsigned oct_mask,
                                                  int touched0) {
  static const int kRefOrder[kNoct] = {0, 4, 2, 6, 3, 7, 1, 5};
  std::vector<std::vector<TileItem>> launches;
  std::vector<std::vector<int>> done(kNoct, std::vector<int>(ntx * nty, 0));
  std::vector<int> remaining(kNoct, 0);
  int left = 0;
  for (int o = 0; o < kNoct; ++o)
    if (oct_mask >> o & 1) {
      remaining[o] = ntx * nty;
      left += ntx * nty;
    }
  // visits so far per physical tile (touched0 from earlier passes)
  std::vector<int> touched(ntx * nty, touched0);
  while (left > 0) {
    std::vector<int> order;
    for (int o = 0; o < kNoct; ++o) {
      const int oc = kind == 1 ? kRefOrder[o] : o;
      if (oct_mask >> oc & 1) order.push_back(oc);
    }
    if (kind != 1)
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return remaining[a] > remaining[b]; });
    std::vector<int> used(ntx * nty, 0);
    std::vector<TileItem> items;
    for (int o : order) {
      // ready visits, nearest the upwind corner first
      for (int d = 0; d <= ntx + nty - 2; ++d)
        for (int i = 0; i < ntx; ++i) {
          const int j = d - i;
          if (j < 0 || j >= nty) continue;
          if (done[o][i + ntx * j]) continue;
          if (i > 0 && done[o][(i - 1) + ntx * j] != 1) continue;
          if (j > 0 && done[o][i + ntx * (j - 1)] != 1) continue;
          const int gx = (o & 1) ? ntx - 1 - i : i;
          const int gy = (o & 2) ? nty - 1 - j : j;
          const int g = gx + ntx * gy;
          if (used[g]) continue;
          used[g] = 1;
          TileItem t;
          t.octant = o;
          t.tx = i;
          t.ty = j;
          t.flags = (touched[g] == 0 ? kVisitFirst : 0) | (touched[g] == kNoct - 1 ? kVisitLast : 0);
          items.push_back(t);
        }
    }
    for (const TileItem& t : items) {
      done[t.octant][t.tx + ntx * t.ty] = 2;  // finished in this launch
      const int gx = (t.octant & 1) ? ntx - 1 - t.tx : t.tx;
      const int gy = (t.octant & 2) ? nty - 1 - t.ty : t.ty;
      touched[gx + ntx * gy]++;
      remaining[t.octant]--;
      left--;
    }
    for (int o = 0; o < kNoct; ++o)
      for (auto& v : done[o])
        if (v == 2) v = 1;
    launches.push_back(items);
  }
  return launches;
}

}  // namespace msw

using namespace msw;

extern "C" {

const char* msw_last_error(void) { return g_err.c_str(); }

int msw_init_matrices_host(int nm, int na, double* A, double* M) {
  if (nm <= 0 || na <= 0 || !A || !M)
    return set_err(MSW_ERR_INVALID, "bad matrix init arguments");
  // Quantities_init_am_matrices_ (src/3_sweeper/quantities_testing_c.h:40-201):
  // integer rows that send an affine vector in moments to an affine vector in
  // angles and back, plus rows that annihilate affine vectors.
  const size_t n = (size_t)kNoct * nm * na;
  std::fill(A, A + n, 0.0);
  std::fill(M, M + n, 0.0);
  auto afm = [&](int o, int m, int a) -> double& { return A[a + (size_t)na * (m + (size_t)nm * o)]; };
  auto mfa = [&](int o, int m, int a) -> double& { return M[m + (size_t)nm * (a + (size_t)na * o)]; };
  for (int o = 0; o < kNoct; ++o) {
    for (int i = 0; i < na; ++i) {
      const int q = (i + 1) / nm, r = (i + 1) % nm;
      afm(o, nm - 1, i) += q;
      if (r) {
        afm(o, 0, i) -= 1.0;
        afm(o, r, i) += 1.0;
      }
    }
    for (int m = 0; m < nm - 2; ++m)
      for (int a = 0; a < na; ++a) {
        const int rv = 21 + (m + nm * a) % 17;
        afm(o, m, a) += -1.0 * rv;
        afm(o, m + 1, a) += 2.0 * rv;
        afm(o, m + 2, a) += -1.0 * rv;
      }
    for (int i = 0; i < nm; ++i) {
      const int q = (i + 1) / na, r = (i + 1) % na;
      mfa(o, i, na - 1) += q;
      if (r) {
        mfa(o, i, 0) -= 1.0;
        mfa(o, i, r) += 1.0;
      }
    }
    for (int m = 0; m < nm; ++m)
      for (int a = 0; a < na - 2; ++a) {
        const int rv = 37 + (m + nm * a) % 19;
        mfa(o, m, a) += -1.0 * rv;
        mfa(o, m, a + 1) += 2.0 * rv;
        mfa(o, m, a + 2) += -1.0 * rv;
      }
    for (int m = 0; m < nm; ++m)
      for (int a = 0; a < na; ++a) {
        mfa(o, m, a) /= kNoct;
        mfa(o, m, a) /= (double)(1 << (a & 7));
      }
  }
  return MSW_OK;
}

}  // extern "C"
