// Host-only stand-in for the HIP runtime, for ASan runs of the batch host code.
// Memory is host heap; launches are checked (grid, segment tables, every buffer
// extent a tile kernel would touch) instead of run.  Captured work is stored in
// the graph and checked at each replay.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <random>
#include <vector>
#include "msw_device.h"
#include "msw_internal.h"
#include "msw_batch.h"

struct ihipStream_t { int id; };
struct ihipEvent_t { int id; };
struct ihipGraph { std::vector<std::function<void()>> ops; };
struct hipGraphExec { std::vector<std::function<void()>> ops; };

static std::map<uintptr_t, size_t> g_alloc;
static ihipGraph* g_cap = nullptr;
static ihipStream_t* g_cap_stream = nullptr;
std::mt19937 g_rng(1);
long g_launches = 0, g_checked_wg = 0;

#define DIE(...) do { fprintf(stderr, "FAKE-HIP CHECK FAILED: " __VA_ARGS__); fprintf(stderr, "\n"); abort(); } while (0)

static void check_range(const void* p, size_t bytes, const char* what) {
  if (!p) DIE("%s: null", what);
  uintptr_t a = (uintptr_t)p;
  auto it = g_alloc.upper_bound(a);
  if (it == g_alloc.begin()) DIE("%s: %p not in any allocation", what, p);
  --it;
  if (a + bytes > it->first + it->second)
    DIE("%s: [%p, +%zu) past allocation [%p, +%zu)", what, p, bytes, (void*)it->first, it->second);
}
static void enqueue(std::function<void()> f) {
  if (g_cap) g_cap->ops.push_back(std::move(f)); else f();
}

extern "C" {
hipError_t hipMalloc(void** p, size_t n) {
  if (n == 0) n = 1;
  void* q = aligned_alloc(256, (n + 255) / 256 * 256);
  if (!q) return hipErrorOutOfMemory;
  g_alloc[(uintptr_t)q] = n;
  *p = q;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  auto it = g_alloc.find((uintptr_t)p);
  if (it == g_alloc.end()) DIE("hipFree of unknown %p", p);
  g_alloc.erase(it);
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
static bool dev_dst(hipMemcpyKind k) { return k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice; }
static bool dev_src(hipMemcpyKind k) { return k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) {
  if (g_cap) DIE("synchronous hipMemcpy during capture");
  if (k == hipMemcpyDefault) DIE("hipMemcpyDefault not modelled");
  if (dev_dst(k)) check_range(d, n, "memcpy dst");
  if (dev_src(k)) check_range(s, n, "memcpy src");
  memcpy(d, s, n); return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  if (k == hipMemcpyDefault) DIE("hipMemcpyDefault not modelled");
  if (dev_dst(k)) check_range(d, n, "memcpyasync dst");
  if (dev_src(k)) check_range(s, n, "memcpyasync src");
  enqueue([=] { if (dev_dst(k)) check_range(d, n, "memcpyasync dst@run"); if (dev_src(k)) check_range(s, n, "memcpyasync src@run"); memcpy(d, s, n); });
  return hipSuccess;
}
hipError_t hipMemset(void* d, int v, size_t n) {
  if (g_cap) DIE("synchronous hipMemset during capture");
  check_range(d, n, "memset"); memset(d, v, n); return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { static int id; *s = new ihipStream_t{++id}; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned) { if (!e) DIE("wait on null event"); return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { static int id; *e = new ihipEvent_t{++id}; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { if (!e) DIE("record null event"); return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t e) { if (g_cap) DIE("event sync during capture"); if (!e) DIE("sync null event"); return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 1.0f + (float)(g_rng() % 1000) / 100.0f; return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "fake"; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (g_cap) DIE("nested capture");
  g_cap = new ihipGraph(); g_cap_stream = s; return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  if (!g_cap || s != g_cap_stream) DIE("end capture on a stream not capturing");
  *g = g_cap; g_cap = nullptr; g_cap_stream = nullptr; return hipSuccess;
}
hipError_t hipGraphInstantiateWithFlags(hipGraphExec_t* x, hipGraph_t g, unsigned long long) {
  if (!g) DIE("instantiate null graph");
  *x = new hipGraphExec{g->ops}; return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) {
  if (!x) DIE("launch null exec");
  for (auto& f : x->ops) f();
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t x) { delete x; return hipSuccess; }
}  // extern "C"

namespace msw {

static void check_params(const SweepParams& p, int precision, int nm, int nac, int nwa, int ny) {
  const size_t ns = (size_t)p.nx * p.ny * p.nz * p.ne * nm * p.nu;
  check_range(p.vi, ns * 8, "vi");
  check_range(p.vo, ns * 8, "vo");
  check_range(p.at, (size_t)8 * p.na_pad * nm * 8, "at");
  check_range(p.mf, (size_t)8 * p.na_pad * nm * 8, "mf");
  const size_t fe = precision == MSW_PREC_FP64 ? 8 : 4;
  check_range(p.fyz, (size_t)8 * p.nu * p.ny * p.ne * p.nz * p.na_pad * fe, "fyz");
  check_range(p.fxz, (size_t)8 * p.nu * p.nx * p.ne * p.nz * p.na_pad * fe, "fxz");
  if (precision != MSW_PREC_FP64) {
    check_range(p.vo32, ns * 4, "vo32");
    check_range(p.vlo, ns * 2, "vlo");
    if (precision == MSW_PREC_MIXED_FP32) check_range(p.vhi, ns * 4, "vhi");
  }
  if (p.na_pad != nwa * nac) DIE("na_pad %d != nwa %d * nac %d", p.na_pad, nwa, nac);
  if (p.na > p.na_pad) DIE("na %d > na_pad %d", p.na, p.na_pad);
  if (p.ntx != (p.nx + 15) / 16) DIE("ntx");
  if (p.nty != (p.ny + 4 * ny - 1) / (4 * ny)) DIE("nty %d ny %d rows %d", p.nty, p.ny, 4 * ny);
  if (!mfma_face_offsets_fit(MswDims{p.nx, p.ny, p.nz, p.ne, nm, p.na, p.nu}, p.na_pad)) DIE("offsets");
}

void launch_mfma_multi(int precision, int nm, int nac, int nwa, int ny, unsigned grid,
                       const SweepParams* probs, const MultiLaunch& ml, hipStream_t st) {
  if (!mfma_shape_supported(nwa, ny)) DIE("unsupported shape nwa %d ny %d", nwa, ny);
  if (nac != 16 && nac != 32 && nac != 64) DIE("nac %d", nac);
  if (nm != 4 && nm != 8 && nm != 16) DIE("nm %d", nm);
  if (mfma_lds_bytes(precision, nac, nwa, ny) > 160 * 1024) DIE("lds");
  if (ml.nseg < 1 || ml.nseg > kMaxSegments) DIE("nseg %d", ml.nseg);
  MultiLaunch m = ml;  // kernel argument: by value
  ++g_launches;
  enqueue([=] {
    for (unsigned blk = 0; blk < grid; ++blk) {
      int k = 0;
      for (int step = kMaxSegments / 2; step > 0; step >>= 1)
        if (k + step < m.nseg && m.start[k + step] <= (int)blk) k += step;
      if (m.start[0] != 0) DIE("segment 0 starts at %d", m.start[0]);
      check_range(probs + m.prob[k], sizeof(SweepParams), "probs[prob[k]]");
      const SweepParams p = probs[m.prob[k]];
      const int epw = m.epw_ng[k] & kEpwMask;
      const bool xmap = (m.epw_ng[k] & kXcdFlag) != 0;
      const int ng = m.epw_ng[k] >> (kEpwBits + 1);
      if (epw < 1 || epw > kMaxStreamGroups) DIE("epw %d", epw);
      const long long seg_end = k + 1 < m.nseg ? m.start[k + 1] : (long long)grid;
      const long long want = xmap ? xcd_round_ids(ng, p.nu) : (long long)ng * p.nu;
      if (seg_end - m.start[k] != want) DIE("segment %d: %lld workgroups for %d groups", k, seg_end - m.start[k], ng);
      if (xmap && m.start[k] % 8) DIE("segment %d starts off an XCD round", k);
      int bid = (int)blk - m.start[k];
      if (xmap) {  // the kernel's XCD mapping: id -> (group, unknown), ids past the groups exit
        const int kk = bid / 8, g = (kk / p.nu) * 8 + bid % 8;
        if (g >= ng) { ++g_checked_wg; continue; }
        bid = g * p.nu + kk % p.nu;
      }
      const int iu = bid % p.nu; bid /= p.nu;
      const int nec = (p.ne + epw - 1) / epw;
      const int ec = bid % nec; bid /= nec;
      if (iu >= p.nu || ec * epw >= p.ne) DIE("iu/ec");
      check_range(p.items + m.item_base[k] + bid, sizeof(TileItem), "item");
      const TileItem it = p.items[m.item_base[k] + bid];
      if (it.octant < 0 || it.octant > 7 || it.tx < 0 || it.tx >= p.ntx || it.ty < 0 || it.ty >= p.nty)
        DIE("item (%d %d %d) outside %dx%d", it.octant, it.tx, it.ty, p.ntx, p.nty);
      if (blk == 0 || (int)blk == m.start[k]) check_params(p, precision, nm, nac, nwa, ny);
      ++g_checked_wg;
    }
  });
  (void)st;
}

void launch_pack(const double* afm, const double* mfa, double* at, double* mf, int nm, int na,
                 int na_pad, hipStream_t) {
  check_range(afm, (size_t)8 * nm * na * 8, "afm");
  check_range(mfa, (size_t)8 * nm * na * 8, "mfa");
  check_range(at, (size_t)8 * na_pad * nm * 8, "at out");
  check_range(mf, (size_t)8 * na_pad * nm * 8, "mf out");
  enqueue([=] { check_range(at, (size_t)8 * na_pad * nm * 8, "at out@run"); check_range(afm, (size_t)8 * nm * na * 8, "afm@run"); });
}
void launch_operand_copy(int precision, const MswDims& d, const double* vi, float* hi,
                         unsigned short* lo, int, hipStream_t) {
  const size_t n = (size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.nm * d.nu;
  enqueue([=] {
    check_range(vi, n * 8, "copy vi");
    check_range(lo, n * 2, "copy lo");
    if (precision == MSW_PREC_MIXED_FP32) check_range(hi, n * 4, "copy hi");
  });
}
void launch_init_state(const MswDims& d, double* v, hipStream_t) {
  const size_t n = (size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.nm * d.nu;
  enqueue([=] { check_range(v, n * 8, "init_state"); });
}
size_t norms_scratch_doubles() { return 2 * 1024; }
void launch_norms(const double* vi, const double* vo, size_t n, double* part, double* out2,
                  hipStream_t) {
  enqueue([=] {
    check_range(vi, n * 8, "norms vi"); check_range(vo, n * 8, "norms vo");
    check_range(part, norms_scratch_doubles() * 8, "norms part"); check_range(out2, 16, "norms out");
    out2[0] = 1; out2[1] = 0;
  });
}
}  // namespace msw
