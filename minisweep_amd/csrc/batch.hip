// batch.hip -- many independent sweeps on one GPU: the batched form (one
// hipGraph replay per run, autotuned) and the served form (a request loop with
// continuous batching and admission by HBM footprint).
//
// Both run on one mechanism, the combined tick: every problem in flight
// advances by one launch of its own tile schedule (the same list schedule and
// kernels as a single msw_plan_sweep), and the tick issues one
// k_sweep_mfma_multi launch per kernel variant (nm, angles per wave, waves per
// tile) covering all of them.  Small problems that would leave most of the
// 256 CUs idle on their own fill the chip together, and each problem's
// arithmetic, order of operations and result bits are exactly those of its
// single sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "msw_device.h"
#include "msw_internal.h"
#include "../../include/msw_batch.h"

namespace msw {

namespace {

int hip_fail(hipError_t e, const char* what) {
  return set_error(e == hipErrorOutOfMemory ? MSW_ERR_NOMEM : MSW_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_OK(call, what)                  \
  do {                                      \
    hipError_t e_ = (call);                 \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Tile schedule of one (ntx, nty, kind), items on the device.
struct DevSchedule {
  std::vector<int> off, cnt;
  TileItem* items = nullptr;
  int nlaunch() const { return (int)off.size(); }
};

int make_schedule(int ntx, int nty, int kind, DevSchedule* s) {
  auto launches = build_schedule(ntx, nty, kind);
  std::vector<TileItem> flat;
  for (auto& l : launches) {
    s->off.push_back((int)flat.size());
    s->cnt.push_back((int)l.size());
    flat.insert(flat.end(), l.begin(), l.end());
  }
  HIP_OK(hipMalloc(&s->items, sizeof(TileItem) * flat.size()), "schedule allocation");
  HIP_OK(hipMemcpy(s->items, flat.data(), sizeof(TileItem) * flat.size(), hipMemcpyHostToDevice),
         "schedule upload");
  return MSW_OK;
}

typedef std::map<std::tuple<int, int, int>, DevSchedule> ScheduleCache;

int get_schedule(ScheduleCache& c, int ntx, int nty, int kind, const DevSchedule** out) {
  auto key = std::make_tuple(ntx, nty, kind);
  auto it = c.find(key);
  if (it == c.end()) {
    DevSchedule s;
    const int rc = make_schedule(ntx, nty, kind, &s);
    if (rc) return rc;
    it = c.emplace(key, std::move(s)).first;
  }
  *out = &it->second;
  return MSW_OK;
}

void free_schedules(ScheduleCache& c) {
  for (auto& kv : c)
    if (kv.second.items) (void)hipFree(kv.second.items);
  c.clear();
}

// Geometry of one problem under the MFMA kernels.
struct Geom {
  MswDims d{};
  int nac = 16, nwa = 1, nap = 16, ntx = 1, nty = 1, epw = 1, ny = 1;
  size_t mat_doubles() const { return ((size_t)kNoct * nap + nac) * d.nm; }
  size_t face_elems() const {
    return (size_t)kNoct * d.ne * d.nu * (d.ncell_x + d.ncell_y) * d.ncell_z * nap;
  }
  size_t state_doubles() const {
    return (size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.nm * d.nu;
  }
  std::tuple<int, int, int, int> variant() const { return std::make_tuple(d.nm, nac, nwa, ny); }
};

int check_dims_mfma(const MswDims& d) {
  if (d.ncell_x <= 0 || d.ncell_y <= 0 || d.ncell_z <= 0 || d.ne <= 0 || d.na <= 0 ||
      d.nu <= 0)
    return set_error(MSW_ERR_INVALID, "all dimensions must be positive");
  if (d.nm != 4 && d.nm != 8 && d.nm != 16)
    return set_error(MSW_ERR_UNSUPPORTED, "nm must be 4, 8 or 16");
  if ((size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.nu >= (1ull << 32))
    return set_error(MSW_ERR_UNSUPPORTED, "state of 2^32 or more cells x unknowns");
  return MSW_OK;
}

int make_geom(const MswDims& d, int nac, Geom* g, int ny = 0, int precision = MSW_PREC_FP64) {
  int rc = check_dims_mfma(d);
  if (rc) return rc;
  if (nac != 16 && nac != 32 && nac != 64)
    return set_error(MSW_ERR_INVALID, "angles per wave must be 16, 32 or 64");
  g->d = d;
  g->nac = nac;
  g->nwa = (d.na + nac - 1) / nac;
  if (g->nwa > 4) return set_error(MSW_ERR_UNSUPPORTED, "na above 4 waves of angles");
  g->nap = g->nwa * nac;
  if (!mfma_face_offsets_fit(d, g->nap))
    return set_error(MSW_ERR_UNSUPPORTED, "face planes too large for 32-bit offsets");
  g->ntx = (d.ncell_x + mfma_tile_x() - 1) / mfma_tile_x();
  // requested row groups, or fewer where that workgroup shape is not built
  g->ny = ny > 0 ? ny : default_mfma_ny(precision, g->nwa, d.ncell_y);
  while (g->ny > 1 && !mfma_shape_supported(g->nwa, g->ny)) g->ny /= 2;
  const int rows = mfma_tile_rows() * g->ny;
  g->nty = (d.ncell_y + rows - 1) / rows;
  g->epw = default_mfma_epw(d.ncell_z, d.ne);  // preferred; a tick may use fewer
  return MSW_OK;
}

// Mixed forms (none in FP64): the FP32 running vo sum, then the operand
// copies of vi -- the BF16 part (both forms), then the FP32 part (FP32 form).
size_t vo32_bytes(const Geom& g, int precision) {
  if (precision == MSW_PREC_FP64) return 0;
  const size_t n = g.state_doubles();
  return align_up(sizeof(float) * n) + align_up(sizeof(unsigned short) * n) +
         (precision == MSW_PREC_MIXED_FP32 ? align_up(sizeof(float) * n) : 0);
}
unsigned short* vlo_of(const Geom& g, float* vo32) {
  return vo32 ? (unsigned short*)((char*)vo32 + align_up(sizeof(float) * g.state_doubles()))
              : nullptr;
}
float* vhi_of(const Geom& g, float* vo32, int precision) {
  return vo32 && precision == MSW_PREC_MIXED_FP32
             ? (float*)((char*)vlo_of(g, vo32) + align_up(sizeof(unsigned short) * g.state_doubles()))
             : nullptr;
}
// The mixed forms' operand copies of one problem's vi (nothing in FP64).
void enqueue_operand_copy(const Geom& g, const double* vi, float* vo32, int precision,
                          hipStream_t st) {
  if (vo32)
    launch_operand_copy(precision, g.d, vi, vhi_of(g, vo32, precision), vlo_of(g, vo32), 0, st);
}

SweepParams make_params(const Geom& g, const double* vi, double* vo, const double* at,
                        const double* mf, void* fyz, void* fxz, float* vo32,
                        const DevSchedule* s, int precision) {
  SweepParams sp;
  sp.vo32 = vo32;
  sp.vhi = vhi_of(g, vo32, precision);
  sp.vlo = vlo_of(g, vo32);
  sp.vib = sp.vlo;
  sp.vi = vi;
  sp.vo = vo;
  sp.at = at;
  sp.mf = mf;
  sp.fyz = fyz;
  sp.fxz = fxz;
  sp.items = s->items;
  sp.nx = g.d.ncell_x;
  sp.ny = g.d.ncell_y;
  sp.nz = g.d.ncell_z;
  sp.ne = g.d.ne;
  sp.na = g.d.na;
  sp.nu = g.d.nu;
  sp.na_pad = g.nap;
  sp.ntx = g.ntx;
  sp.nty = g.nty;
  return sp;
}

// One problem's share of a combined tick.
struct TickWork {
  int slot;                 // index into the device SweepParams array
  const Geom* g;
  const DevSchedule* s;
  int launch;               // which launch of its schedule
  int epw = 1;              // energy groups per workgroup in this tick
};

// Energy groups per workgroup for a tick: each problem's preferred stream
// length, halved for all while the tick would have < 4 waves per SIMD.
void choose_tick_epw(std::vector<TickWork>& work, int precision) {
  const long long target = precision == MSW_PREC_FP64 ? 4096 : 2048;
  for (TickWork& w : work) w.epw = w.g->epw;
  for (;;) {
    long long waves = 0;
    int emax = 1;
    for (const TickWork& w : work) {
      waves += (long long)w.s->cnt[w.launch] * ((w.g->d.ne + w.epw - 1) / w.epw) * w.g->d.nu *
               w.g->nwa * w.g->ny;
      emax = std::max(emax, w.epw);
    }
    if (waves >= target || emax == 1) return;
    for (TickWork& w : work) w.epw = std::max(1, w.epw / 2);
  }
}

// Host-side checks of one problem's share of a tick before it goes into a
// MultiLaunch: the launch exists in its schedule, the stream length fits its
// bits, and its workgroups fit the kernel's int indexing.  Returns the
// (item, group chunk) count, or -1 (and sets the error text); its workgroups
// are xcd_round_ids(count, nu).
long long tick_groups(const TickWork& w) {
  if (!w.g || !w.s || w.launch < 0 || w.launch >= w.s->nlaunch() || w.epw < 1 ||
      w.epw > kEpwMask || w.s->cnt[w.launch] < 1 || w.s->off[w.launch] < 0) {
    set_error(MSW_ERR_INVALID, "internal: tick work outside its schedule");
    return -1;
  }
  const long long ng = (long long)w.s->cnt[w.launch] * ((w.g->d.ne + w.epw - 1) / w.epw);
  if (ng >= (1ll << (30 - kEpwBits))) {
    set_error(MSW_ERR_UNSUPPORTED, "internal: tick segment of too many workgroups");
    return -1;
  }
  return ng;
}

// Enqueue one combined tick: per kernel variant, k_sweep_mfma_multi launches
// of up to kMaxSegments problems each.  Returns the number of launches, or a
// negated MSW_ERR_* code when the assembly checks fail (nothing of the
// offending launch is enqueued).
int enqueue_tick(int precision, std::vector<TickWork>& work, const SweepParams* d_params,
                 hipStream_t st, bool xcd_map) {
  choose_tick_epw(work, precision);
  std::map<std::tuple<int, int, int, int>, std::vector<const TickWork*>> by_variant;
  for (const TickWork& w : work) by_variant[w.g->variant()].push_back(&w);
  int nl = 0;
  for (auto& kv : by_variant) {
    const auto& ws = kv.second;
    for (size_t b = 0; b < ws.size(); b += kMaxSegments) {
      MultiLaunch ml;
      ml.nseg = 0;
      long long grid = 0;
      for (size_t k = b; k < ws.size() && k < b + kMaxSegments; ++k) {
        const TickWork& w = *ws[k];
        const long long ng = tick_groups(w);
        if (ng < 0) return -MSW_ERR_INVALID;
        const long long nwg = xcd_map ? xcd_round_ids(ng, w.g->d.nu) : ng * w.g->d.nu;
        if (ml.nseg >= kMaxSegments || w.slot < 0) {
          set_error(MSW_ERR_INVALID, "internal: segment table overflow");
          return -MSW_ERR_INVALID;
        }
        ml.start[ml.nseg] = (int)grid;
        ml.prob[ml.nseg] = w.slot;
        ml.item_base[ml.nseg] = w.s->off[w.launch];
        ml.epw_ng[ml.nseg] = w.epw | (xcd_map ? kXcdFlag : 0) | (int)ng << (kEpwBits + 1);
        ml.nseg++;
        grid += nwg;
      }
      if (grid < 1 || grid > 0x7fffffffll) {
        set_error(MSW_ERR_INVALID, "internal: combined launch of " + std::to_string(grid) +
                                       " workgroups");
        return -MSW_ERR_INVALID;
      }
      const Geom& g = *ws[b]->g;
      launch_mfma_multi(precision, g.d.nm, g.nac, g.nwa, g.ny, (unsigned)grid, d_params, ml, st);
      ++nl;
    }
  }
  return nl;
}

// One tick with each kernel variant's launches on a stream of its own
// (round-robin over ncs streams), forked from st and joined back to it.
int enqueue_tick_forked(int precision, std::vector<TickWork>& work, const SweepParams* d_params,
                        hipStream_t st, hipStream_t* cs, hipEvent_t fork, hipEvent_t* join,
                        int ncs, bool xcd_map) {
  std::map<std::tuple<int, int, int, int>, std::vector<TickWork>> by_variant;
  for (const TickWork& w : work) by_variant[w.g->variant()].push_back(w);
  const int n = std::min<int>((int)by_variant.size(), ncs);
  if (n < 2) return enqueue_tick(precision, work, d_params, st, xcd_map);
  (void)hipEventRecord(fork, st);
  for (int c = 0; c < n; ++c) (void)hipStreamWaitEvent(cs[c], fork, 0);
  int nl = 0, c = 0, err = 0;
  for (auto& kv : by_variant) {
    const int k = enqueue_tick(precision, kv.second, d_params, cs[c++ % n], xcd_map);
    if (k < 0) err = k;
    else nl += k;
  }
  // joined also after a failed assembly, so that no stream is left forked
  for (int k = 0; k < n; ++k) {
    (void)hipEventRecord(join[k], cs[k]);
    (void)hipStreamWaitEvent(st, join[k], 0);
  }
  return err ? err : nl;
}

}  // namespace

}  // namespace msw

using namespace msw;

// ============================================================================
// Batched sweeps
// ============================================================================

namespace {

struct BatchProb {
  Geom g;
  const DevSchedule* s = nullptr;
  const double* vi = nullptr;
  double* vo = nullptr;
  const double* afm = nullptr;
  const double* mfa = nullptr;
  double* at = nullptr;
  double* mf = nullptr;
  void* fyz = nullptr;
  void* fxz = nullptr;
  float* vo32 = nullptr;
  int start_tick = 0;
};

}  // namespace

// Streams the tick chains of different kernel variants run on concurrently.
constexpr int kChainStreams = 4;

struct MswBatch {
  MswBatchOptions opt{};
  std::vector<MswDims> dims;
  std::vector<BatchProb> probs;
  std::map<std::pair<int, int>, int> class_nac;  // (nm, na) -> angles per wave
  ScheduleCache scheds;
  char* arena = nullptr;
  size_t arena_bytes = 0;
  SweepParams* d_params = nullptr;
  bool built = false;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nkernels = 0, nticks = 0;
  int stream_target = 128;  // z-stream steps per workgroup (energy groups streamed)
  int row_groups = 0;       // row groups of 4 rows per tile (0: the plan default)
  std::string log;
  // the kernel variants' tick chains run on their own streams, forked from
  // and joined back to the run's stream (also inside the captured graph)
  hipStream_t chain_st[kChainStreams] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kChainStreams] = {};
  // recorded after every run on the run's stream: the graph, the arena and
  // the parameters of the last build are released only after it
  hipEvent_t done_ev = nullptr;
  bool ran = false;
};

namespace {

// Drop the built run.  Waits for the last run first: a graph exec or the
// memory its kernels use must not go away while a replay may still run.
void batch_release(MswBatch* b) {
  if (b->ran && b->done_ev) (void)hipEventSynchronize(b->done_ev);
  b->ran = false;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (b->arena) (void)hipFree(b->arena);
  if (b->d_params) (void)hipFree(b->d_params);
  b->arena = nullptr;
  b->d_params = nullptr;
  b->built = false;
}

int nac_for(const MswBatch* b, const MswDims& d) {
  auto it = b->class_nac.find(std::make_pair(d.nm, d.na));
  return it != b->class_nac.end() && it->second ? it->second
                                                : default_mfma_nac(b->opt.precision, d, false);
}

// Start ticks that even out the combined launches: problems with the most
// work first, each at the offset that keeps the busiest tick it touches
// least loaded.  Loads are in workgroups.
void stagger_starts(MswBatch* b) {
  int horizon = 0;
  for (auto& p : b->probs) horizon = std::max(horizon, p.s->nlaunch());
  horizon += horizon / 2;
  std::vector<long long> load(horizon, 0);
  std::vector<int> order(b->probs.size());
  std::vector<long long> total(b->probs.size(), 0);
  for (size_t i = 0; i < b->probs.size(); ++i) {
    order[i] = (int)i;
    const BatchProb& p = b->probs[i];
    for (int l = 0; l < p.s->nlaunch(); ++l) total[i] += (long long)p.s->cnt[l] * ((p.g.d.ne + p.g.epw - 1) / p.g.epw) * p.g.d.nu;
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return total[x] > total[y]; });
  for (int i : order) {
    BatchProb& p = b->probs[i];
    const int L = p.s->nlaunch();
    int best = 0;
    long long best_peak = -1;
    for (int s0 = 0; s0 + L <= horizon; ++s0) {
      long long peak = 0;
      for (int l = 0; l < L; ++l)
        peak = std::max(peak, load[s0 + l] + (long long)p.s->cnt[l] * ((p.g.d.ne + p.g.epw - 1) / p.g.epw) * p.g.d.nu);
      if (best_peak < 0 || peak < best_peak) {
        best_peak = peak;
        best = s0;
      }
    }
    p.start_tick = best;
    for (int l = 0; l < L; ++l) load[best + l] += (long long)p.s->cnt[l] * ((p.g.d.ne + p.g.epw - 1) / p.g.epw) * p.g.d.nu;
  }
}

// Everything one run does, enqueued on st.
int batch_enqueue(MswBatch* b, hipStream_t st) {
  const int prec = b->opt.precision;
  int nk = 0;
  for (auto& p : b->probs) {
    launch_pack(p.afm, p.mfa, p.at, p.mf, p.g.d.nm, p.g.d.na, p.g.nap, st);
    ++nk;
    if (p.vo32) {
      enqueue_operand_copy(p.g, p.vi, p.vo32, prec, st);
      ++nk;
    }
  }
  int nt = 0;
  for (auto& p : b->probs) nt = std::max(nt, p.start_tick + p.s->nlaunch());
  // One chain of ticks per kernel variant: problems of different variants
  // share no launch, so the chains run concurrently on their own streams
  // (round-robin over kChainStreams), forked after the matrix packing and
  // operand copies and joined at the end.  Within a chain, tick t+1 follows
  // tick t.
  std::map<std::tuple<int, int, int, int>, std::vector<int>> chains;
  for (size_t i = 0; i < b->probs.size(); ++i) chains[b->probs[i].g.variant()].push_back((int)i);
  const int nch = std::min<int>((int)chains.size(), kChainStreams);
  if (nch > 1) {
    // the chain streams and events exist from msw_batch_create on: nothing
    // is created while a capture is open
    HIP_OK(hipEventRecord(b->fork_ev, st), "fork");
    for (int c = 0; c < nch; ++c) HIP_OK(hipStreamWaitEvent(b->chain_st[c], b->fork_ev, 0), "fork");
  }
  int ci = 0, err = 0;
  for (auto& kv : chains) {
    hipStream_t cs = nch > 1 ? b->chain_st[ci % nch] : st;
    ++ci;
    for (int t = 0; t < nt && !err; ++t) {
      std::vector<TickWork> work;
      for (int i : kv.second) {
        const BatchProb& p = b->probs[i];
        const int l = t - p.start_tick;
        if (l >= 0 && l < p.s->nlaunch()) work.push_back(TickWork{i, &p.g, p.s, l});
      }
      if (work.empty()) continue;
      const int k = enqueue_tick(prec, work, b->d_params, cs, b->opt.xcd_map != 0);
      if (k < 0) err = -k;
      else nk += k;
    }
  }
  if (nch > 1) {
    for (int c = 0; c < nch; ++c) {
      HIP_OK(hipEventRecord(b->join_ev[c], b->chain_st[c]), "join");
      HIP_OK(hipStreamWaitEvent(st, b->join_ev[c], 0), "join");
    }
  }
  if (err) return err;  // the chains are joined: a capture can still end
  b->nticks = nt;
  b->nkernels = nk;
  HIP_OK(hipGetLastError(), "batch launch");
  return MSW_OK;
}

int batch_build(MswBatch* b) {
  batch_release(b);
  for (size_t i = 0; i < b->probs.size(); ++i)
    if (!b->probs[i].vi || !b->probs[i].vo || !b->probs[i].afm || !b->probs[i].mfa)
      return set_error(MSW_ERR_INVALID, "problem " + std::to_string(i) + " has no buffers");
  const size_t fsz = mfma_face_elem_bytes(b->opt.precision);
  size_t total = 0;
  std::vector<size_t> offs(b->probs.size());
  for (size_t i = 0; i < b->probs.size(); ++i) {
    BatchProb& p = b->probs[i];
    int rc = make_geom(b->dims[i], nac_for(b, b->dims[i]), &p.g, b->row_groups, b->opt.precision);
    if (rc) return rc;
    const int nz = p.g.d.ncell_z;
    p.g.epw = std::max(1, std::min({p.g.d.ne, (b->stream_target + nz / 2) / nz, kMaxStreamGroups}));
    rc = get_schedule(b->scheds, p.g.ntx, p.g.nty, b->opt.schedule_kind, &p.s);
    if (rc) return rc;
    offs[i] = total;
    const size_t nfx = (size_t)kNoct * p.g.d.ne * p.g.d.nu * p.g.d.ncell_y * p.g.d.ncell_z * p.g.nap;
    const size_t nfy = (size_t)kNoct * p.g.d.ne * p.g.d.nu * p.g.d.ncell_x * p.g.d.ncell_z * p.g.nap;
    total += 2 * align_up(sizeof(double) * p.g.mat_doubles()) + align_up(fsz * nfx) +
             align_up(fsz * nfy) + vo32_bytes(p.g, b->opt.precision);
  }
  HIP_OK(hipMalloc(&b->arena, total), "batch arena");
  b->arena_bytes = total;
  HIP_OK(hipMemset(b->arena, 0, total), "batch arena clear");
  for (size_t i = 0; i < b->probs.size(); ++i) {
    BatchProb& p = b->probs[i];
    char* q = b->arena + offs[i];
    const size_t mb = align_up(sizeof(double) * p.g.mat_doubles());
    p.at = (double*)q;
    p.mf = (double*)(q + mb);
    q += 2 * mb;
    const size_t nfx = (size_t)kNoct * p.g.d.ne * p.g.d.nu * p.g.d.ncell_y * p.g.d.ncell_z * p.g.nap;
    const size_t nfy = (size_t)kNoct * p.g.d.ne * p.g.d.nu * p.g.d.ncell_x * p.g.d.ncell_z * p.g.nap;
    p.fyz = q;
    p.fxz = q + align_up(fsz * nfx);
    p.vo32 = vo32_bytes(p.g, b->opt.precision)
                 ? (float*)(q + align_up(fsz * nfx) + align_up(fsz * nfy))
                 : nullptr;
    p.start_tick = 0;
  }
  if (b->opt.stagger) stagger_starts(b);
  std::vector<SweepParams> hp(b->probs.size());
  for (size_t i = 0; i < b->probs.size(); ++i) {
    const BatchProb& p = b->probs[i];
    hp[i] = make_params(p.g, p.vi, p.vo, p.at, p.mf, p.fyz, p.fxz, p.vo32, p.s,
                        b->opt.precision);
  }
  HIP_OK(hipMalloc(&b->d_params, sizeof(SweepParams) * hp.size()), "batch params");
  HIP_OK(hipMemcpy(b->d_params, hp.data(), sizeof(SweepParams) * hp.size(), hipMemcpyHostToDevice),
         "batch params upload");
  if (b->opt.use_graph) {
    HIP_OK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed), "begin capture");
    const int rc = batch_enqueue(b, b->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIP_OK(e, "end capture");
    b->graph = g;
    HIP_OK(hipGraphInstantiateWithFlags(&b->exec, b->graph, 0), "graph instantiate");
  }
  b->built = true;
  return MSW_OK;
}

int batch_run(MswBatch* b, hipStream_t st) {
  if (!b->built) {
    const int rc = batch_build(b);
    if (rc) return rc;
  }
  if (b->opt.use_graph) {
    HIP_OK(hipGraphLaunch(b->exec, st), "graph launch");
  } else {
    const int rc = batch_enqueue(b, st);
    if (rc) return rc;
  }
  HIP_OK(hipEventRecord(b->done_ev, st), "run event");
  b->ran = true;
  return MSW_OK;
}

// Mean milliseconds of one run over reps (after one untimed run).
int batch_time(MswBatch* b, hipStream_t st, int reps, double* ms) {
  int rc = batch_run(b, st);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0), "event");
  HIP_OK(hipEventCreate(&e1), "event");
  HIP_OK(hipEventRecord(e0, st), "event record");
  for (int r = 0; r < reps && !rc; ++r) rc = batch_run(b, st);
  HIP_OK(hipEventRecord(e1, st), "event record");
  HIP_OK(hipEventSynchronize(e1), "event sync");
  float t = 0;
  HIP_OK(hipEventElapsedTime(&t, e0, e1), "event time");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / std::max(1, reps);
  return rc;
}

}  // namespace

extern "C" {

void msw_batch_default_options(MswBatchOptions* o) {
  if (!o) return;
  o->precision = MSW_PREC_FP64;
  o->use_graph = 1;
  o->schedule_kind = 0;
  o->stagger = 1;
  o->row_groups = 0;
  o->stream_target = 0;
  o->xcd_map = 1;
}

int msw_batch_create(int nprob, const MswDims* dims, const MswBatchOptions* opt, MswBatch** out) {
  if (!out || !dims || nprob <= 0) return set_error(MSW_ERR_INVALID, "bad batch arguments");
  *out = nullptr;
  MswBatchOptions o;
  msw_batch_default_options(&o);
  if (opt) o = *opt;
  if (o.precision < MSW_PREC_FP64 || o.precision > MSW_PREC_MIXED_BF16)
    return set_error(MSW_ERR_INVALID, "unknown precision");
  if (o.schedule_kind != 0 && o.schedule_kind != 1)
    return set_error(MSW_ERR_INVALID, "schedule_kind must be 0 or 1");
  if (o.row_groups != 0 && o.row_groups != 1 && o.row_groups != 2 && o.row_groups != 4)
    return set_error(MSW_ERR_INVALID, "row_groups must be 0, 1, 2 or 4");
  if (o.stream_target < 0 || o.stream_target > 4096)
    return set_error(MSW_ERR_INVALID, "stream_target must be 0..4096");
  for (int i = 0; i < nprob; ++i) {
    Geom g;
    const int rc = make_geom(dims[i], default_mfma_nac(o.precision, dims[i], false), &g);
    if (rc) return rc;
  }
  MswBatch* b = new MswBatch();
  b->opt = o;
  b->row_groups = o.row_groups;
  b->stream_target = o.stream_target ? o.stream_target : 128;
  b->dims.assign(dims, dims + nprob);
  b->probs.resize(nprob);
  hipError_t e = hipEventCreateWithFlags(&b->done_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming);
  for (int c = 0; c < kChainStreams && e == hipSuccess; ++c) {
    e = hipStreamCreateWithFlags(&b->chain_st[c], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join_ev[c], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    msw_batch_destroy(b);
    return hip_fail(e, "batch streams and events");
  }
  *out = b;
  return MSW_OK;
}

int msw_batch_set_problem(MswBatch* b, int i, const double* vi, double* vo,
                          const double* a_from_m, const double* m_from_a) {
  if (!b || i < 0 || i >= (int)b->probs.size() || !vi || !vo || !a_from_m || !m_from_a)
    return set_error(MSW_ERR_INVALID, "bad msw_batch_set_problem arguments");
  if (vi == vo) return set_error(MSW_ERR_INVALID, "vi and vo must differ");
  BatchProb& p = b->probs[i];
  if (p.vi != vi || p.vo != vo || p.afm != a_from_m || p.mfa != m_from_a) b->built = false;
  p.vi = vi;
  p.vo = vo;
  p.afm = a_from_m;
  p.mfa = m_from_a;
  return MSW_OK;
}

int msw_batch_run(MswBatch* b, void* stream) {
  if (!b) return set_error(MSW_ERR_INVALID, "null batch");
  return batch_run(b, (hipStream_t)stream);
}

int msw_batch_set_class_nac(MswBatch* b, int nm, int na, int nac) {
  if (!b) return set_error(MSW_ERR_INVALID, "null batch");
  if (nac != 0 && nac != 16 && nac != 32 && nac != 64)
    return set_error(MSW_ERR_INVALID, "angles per wave must be 0, 16, 32 or 64");
  if (nac && (na + nac - 1) / nac > 4)
    return set_error(MSW_ERR_UNSUPPORTED, "na above 4 waves of angles");
  b->class_nac[std::make_pair(nm, na)] = nac;
  b->built = false;
  return MSW_OK;
}

int msw_batch_autotune(MswBatch* b, void* stream, int reps) {
  if (!b) return set_error(MSW_ERR_INVALID, "null batch");
  // Off unless MSW_BATCH_AUTOTUNE=1: a GPU run of the tuner ended with a host
  // segmentation fault whose cause is not known (a CPU build of this code
  // under AddressSanitizer, on a stand-in runtime that checks every launch,
  // finds nothing).  Without it a batch runs its default configuration.
  const char* on = getenv("MSW_BATCH_AUTOTUNE");
  if (!on || strcmp(on, "1") != 0)
    return set_error(MSW_ERR_UNSUPPORTED,
                     "msw_batch_autotune is disabled (set MSW_BATCH_AUTOTUNE=1 to enable)");
  hipStream_t st = (hipStream_t)stream;
  reps = std::max(1, reps);
  std::string log = "[";
  auto note = [&](const std::string& what, double ms) {
    char buf[160];
    snprintf(buf, sizeof(buf), "%s{\"try\": \"%s\", \"ms\": %.3f}", log.size() > 1 ? ", " : "",
             what.c_str(), ms);
    log += buf;
  };
  double best = 0;
  int rc = batch_time(b, st, reps, &best);
  if (rc) return rc;
  note("start", best);
  std::vector<std::pair<int, int>> classes;
  for (auto& d : b->dims)
    if (std::find(classes.begin(), classes.end(), std::make_pair(d.nm, d.na)) == classes.end())
      classes.emplace_back(d.nm, d.na);
  // angles per wave of each class, one class at a time (run again after the
  // tile height and stream length have moved)
  auto class_pass = [&]() -> int {
    for (auto& c : classes) {
      const int cur = nac_for(b, MswDims{1, 1, 1, 1, c.first, c.second, 1});
      int keep = cur;
      for (int nac : {16, 32, 64}) {
        if (nac == cur || (c.second + nac - 1) / nac > 4 || nac > 16 * ((c.second + 15) / 16))
          continue;
        b->class_nac[c] = nac;
        b->built = false;
        double ms = 0;
        const int e = batch_time(b, st, reps, &ms);
        if (e) return e;
        note("nm" + std::to_string(c.first) + " na" + std::to_string(c.second) + " nac" +
                 std::to_string(nac), ms);
        if (ms < best) {
          best = ms;
          keep = nac;
        }
      }
      b->class_nac[c] = keep;
      b->built = false;
    }
    return MSW_OK;
  };
  if ((rc = class_pass())) return rc;
  for (int rg : {1, 2, 4}) {
    if (rg == b->row_groups) continue;
    const int save = b->row_groups;
    b->row_groups = rg;
    b->built = false;
    double ms = 0;
    rc = batch_time(b, st, reps, &ms);
    if (rc) return rc;
    note("row_groups" + std::to_string(rg), ms);
    if (ms < best)
      best = ms;
    else
      b->row_groups = save;
    b->built = false;
  }
  for (int target : {32, 64, 256}) {
    if (target == b->stream_target) continue;
    const int save = b->stream_target;
    b->stream_target = target;
    b->built = false;
    double ms = 0;
    rc = batch_time(b, st, reps, &ms);
    if (rc) return rc;
    note("stream_target" + std::to_string(target), ms);
    if (ms < best)
      best = ms;
    else
      b->stream_target = save;
    b->built = false;
  }
  if ((rc = class_pass())) return rc;
  for (int kind : {0, 1})
    for (int stag : {0, 1}) {
      if (kind == b->opt.schedule_kind && stag == b->opt.stagger) continue;
      const MswBatchOptions save = b->opt;
      b->opt.schedule_kind = kind;
      b->opt.stagger = stag;
      b->built = false;
      double ms = 0;
      rc = batch_time(b, st, reps, &ms);
      if (rc) return rc;
      note("schedule_kind" + std::to_string(kind) + " stagger" + std::to_string(stag), ms);
      if (ms < best)
        best = ms;
      else
        b->opt = save;
      b->built = false;
    }
  log += "]";
  b->log = log;
  return batch_build(b);
}

int msw_batch_num_kernel_launches(const MswBatch* b) { return b ? b->nkernels : 0; }

size_t msw_batch_workspace_bytes(const MswBatch* b) { return b ? b->arena_bytes : 0; }

int msw_batch_describe(const MswBatch* b, char* buf, int len) {
  if (!b || !buf || len <= 0) return set_error(MSW_ERR_INVALID, "bad describe arguments");
  std::string s = "{\"precision\": " + std::to_string(b->opt.precision) +
                  ", \"schedule_kind\": " + std::to_string(b->opt.schedule_kind) +
                  ", \"stagger\": " + std::to_string(b->opt.stagger) +
                  ", \"graph\": " + std::to_string(b->opt.use_graph) +
                  ", \"stream_target\": " + std::to_string(b->stream_target) +
                  ", \"row_groups\": " + std::to_string(b->row_groups) +
                  ", \"ticks\": " + std::to_string(b->nticks) +
                  ", \"kernels_per_run\": " + std::to_string(b->nkernels) + ", \"class_nac\": {";
  bool first = true;
  std::vector<std::pair<int, int>> classes;
  for (auto& d : b->dims)
    if (std::find(classes.begin(), classes.end(), std::make_pair(d.nm, d.na)) == classes.end())
      classes.emplace_back(d.nm, d.na);
  for (auto& c : classes) {
    s += (first ? "" : ", ") + std::string("\"nm") + std::to_string(c.first) + "_na" +
         std::to_string(c.second) + "\": " +
         std::to_string(nac_for(b, MswDims{1, 1, 1, 1, c.first, c.second, 1}));
    first = false;
  }
  s += "}, \"tuning\": " + (b->log.empty() ? std::string("null") : b->log) + "}";
  snprintf(buf, (size_t)len, "%s", s.c_str());
  return MSW_OK;
}

int msw_batch_destroy(MswBatch* b) {
  if (!b) return MSW_OK;
  batch_release(b);
  free_schedules(b->scheds);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  for (int c = 0; c < kChainStreams; ++c) {
    if (b->chain_st[c]) (void)hipStreamDestroy(b->chain_st[c]);
    if (b->join_ev[c]) (void)hipEventDestroy(b->join_ev[c]);
  }
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  if (b->done_ev) (void)hipEventDestroy(b->done_ev);
  delete b;
  return MSW_OK;
}

}  // extern "C"

// ============================================================================
// Served sweeps
// ============================================================================

namespace {

// First-fit allocator over [0, size) in kAlign units.
class Arena {
 public:
  void reset(size_t size) {
    free_.clear();
    free_[0] = size;
    used_ = peak_ = 0;
  }
  size_t alloc(size_t n) {
    n = align_up(n);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        const size_t off = it->first, rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        used_ += n;
        peak_ = std::max(peak_, used_);
        return off;
      }
    return SIZE_MAX;
  }
  void release(size_t off, size_t n) {
    n = align_up(n);
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;
  size_t used_ = 0, peak_ = 0;
};

struct Request {
  long long tag;
  MswDims d;
  double* vi_pinned;  // host copy of vi, or null for the manufactured state
  double t_submit;
};

struct Active {
  Request req;
  Geom g;
  const DevSchedule* s = nullptr;
  int slot = -1;
  size_t off = 0, bytes = 0;
  double* vi = nullptr;
  double* vo = nullptr;
  double* part = nullptr;
  void* fyz = nullptr;
  void* fxz = nullptr;
  float* vo32 = nullptr;
  const double* at = nullptr;
  const double* mf = nullptr;
  int next = 0;  // next launch of its schedule
  bool started = false;
  int err = MSW_OK;  // a tick of this request failed to assemble
  double t_admit = 0;
};

struct Kept {
  double* vo;
  size_t off, bytes, n;
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

struct MswServer {
  MswServerOptions opt{};
  int device = 0;
  hipStream_t st = nullptr;
  char* arena_base = nullptr;
  Arena arena;
  SweepParams* d_params = nullptr;
  SweepParams* h_params = nullptr;  // pinned mirror
  double* d_res = nullptr;
  double* h_res = nullptr;          // pinned
  std::vector<int> free_slots;
  ScheduleCache scheds;
  std::map<std::tuple<int, int, int>, std::pair<double*, double*>> mats;  // (nm, na, nap)
  std::thread thr;
  std::mutex mu;
  std::condition_variable cv;
  bool stop = false;
  std::deque<Request> queue;
  std::vector<Active> active;
  std::vector<MswResult> done;
  std::map<long long, Kept> kept;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double t0 = 0;
  long long ndone = 0, nlaunch = 0, nticks = 0;
  int in_flight = 0;  // requests admitted and not yet reported
  // kernel configuration of admitted problems (msw_server_tune_from_batch):
  // angles per wave per (nm, na) class (absent: the plan default), row groups
  // (0: the plan default), z-stream steps per workgroup, schedule heuristic
  std::map<std::pair<int, int>, int> class_nac;
  int row_groups = 0, stream_target = 128, schedule_kind = 0;
  // a tick's kernel variants run on their own streams, forked from st
  hipStream_t chain_st[kChainStreams] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kChainStreams] = {};
};

namespace {

// Packed standard matrices of a shape, cached on the device.
int server_matrices(MswServer* s, const Geom& g, const double** at, const double** mf) {
  auto key = std::make_tuple(g.d.nm, g.d.na, g.nap);
  auto it = s->mats.find(key);
  if (it == s->mats.end()) {
    const size_t n = (size_t)kNoct * g.d.nm * g.d.na;
    std::vector<double> A(n), M(n);
    msw_init_matrices_host(g.d.nm, g.d.na, A.data(), M.data());
    double *dA = nullptr, *dM = nullptr, *pat = nullptr, *pmf = nullptr;
    HIP_OK(hipMalloc(&dA, sizeof(double) * n), "matrix staging");
    HIP_OK(hipMalloc(&dM, sizeof(double) * n), "matrix staging");
    HIP_OK(hipMalloc(&pat, sizeof(double) * g.mat_doubles()), "matrix cache");
    HIP_OK(hipMalloc(&pmf, sizeof(double) * g.mat_doubles()), "matrix cache");
    HIP_OK(hipMemset(pat, 0, sizeof(double) * g.mat_doubles()), "matrix cache");
    HIP_OK(hipMemset(pmf, 0, sizeof(double) * g.mat_doubles()), "matrix cache");
    HIP_OK(hipMemcpy(dA, A.data(), sizeof(double) * n, hipMemcpyHostToDevice), "matrix upload");
    HIP_OK(hipMemcpy(dM, M.data(), sizeof(double) * n, hipMemcpyHostToDevice), "matrix upload");
    launch_pack(dA, dM, pat, pmf, g.d.nm, g.d.na, g.nap, s->st);
    HIP_OK(hipStreamSynchronize(s->st), "matrix pack");
    (void)hipFree(dA);
    (void)hipFree(dM);
    it = s->mats.emplace(key, std::make_pair(pat, pmf)).first;
  }
  *at = it->second.first;
  *mf = it->second.second;
  return MSW_OK;
}

size_t footprint(const Geom& g, int precision) {
  const size_t fsz = mfma_face_elem_bytes(precision);
  const size_t nfx = (size_t)kNoct * g.d.ne * g.d.nu * g.d.ncell_y * g.d.ncell_z * g.nap;
  const size_t nfy = (size_t)kNoct * g.d.ne * g.d.nu * g.d.ncell_x * g.d.ncell_z * g.nap;
  return 2 * align_up(sizeof(double) * g.state_doubles()) + align_up(fsz * nfx) +
         align_up(fsz * nfy) + align_up(sizeof(double) * norms_scratch_doubles()) +
         vo32_bytes(g, precision);
}

void finish(MswServer* s, Active& a, int status) {
  MswResult r;
  r.tag = a.req.tag;
  r.normsq = status == MSW_OK ? s->h_res[2 * a.slot] : 0.0;
  r.normsqdiff = status == MSW_OK ? s->h_res[2 * a.slot + 1] : 0.0;
  r.t_submit = a.req.t_submit - s->t0;
  r.t_admit = a.t_admit - s->t0;
  r.t_done = now_s() - s->t0;
  r.status = status;
  if (a.req.vi_pinned) (void)hipHostFree(a.req.vi_pinned);
  if (s->opt.keep_outputs && status == MSW_OK) {
    // a tag kept and never fetched is replaced by the newer result
    const auto old = s->kept.find(a.req.tag);
    if (old != s->kept.end()) s->arena.release(old->second.off, old->second.bytes);
    s->kept[a.req.tag] = Kept{a.vo, a.off, a.bytes, a.g.state_doubles()};
  } else if (a.bytes) {
    s->arena.release(a.off, a.bytes);
  }
  s->free_slots.push_back(a.slot);
  s->done.push_back(r);
  s->ndone++;
  s->in_flight--;
}

// The request loop: admit, enqueue one combined tick, retire the previous one.
void server_loop(MswServer* s) {
  (void)hipSetDevice(s->device);
  std::vector<int> pending;  // active indices that finish in the tick in flight
  bool tick_in_flight = false;
  int cur = 0;
  for (;;) {
    std::vector<size_t> fresh;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->stop) break;
      // admission: FIFO, while a slot is free and the footprint fits
      while (!s->queue.empty() && !s->free_slots.empty() &&
             (int)s->active.size() < s->opt.max_active) {
        Request& q = s->queue.front();
        Active a;
        a.req = q;
        const auto cn = s->class_nac.find(std::make_pair(q.d.nm, q.d.na));
        int rc = make_geom(q.d,
                           cn != s->class_nac.end() ? cn->second
                                                    : default_mfma_nac(s->opt.precision, q.d,
                                                                       false),
                           &a.g, s->row_groups, s->opt.precision);
        if (!rc) {
          const int nz = a.g.d.ncell_z;
          a.g.epw =
              std::max(1, std::min({a.g.d.ne, (s->stream_target + nz / 2) / nz, kMaxStreamGroups}));
          rc = get_schedule(s->scheds, a.g.ntx, a.g.nty, s->schedule_kind, &a.s);
        }
        if (!rc) rc = server_matrices(s, a.g, &a.at, &a.mf);
        if (rc) {
          a.slot = s->free_slots.back();
          s->free_slots.pop_back();
          finish(s, a, rc);
          s->queue.pop_front();
          continue;
        }
        a.bytes = footprint(a.g, s->opt.precision);
        a.off = s->arena.alloc(a.bytes);
        if (a.off == SIZE_MAX) {
          a.off = 0;
          if (a.bytes > s->opt.hbm_budget_bytes) {  // can never fit
            a.bytes = 0;
            a.slot = s->free_slots.back();
            s->free_slots.pop_back();
            finish(s, a, MSW_ERR_NOMEM);
            s->queue.pop_front();
            continue;
          }
          break;
        }
        a.slot = s->free_slots.back();
        s->free_slots.pop_back();
        char* q0 = s->arena_base + a.off;
        const size_t sb = align_up(sizeof(double) * a.g.state_doubles());
        const size_t fsz = mfma_face_elem_bytes(s->opt.precision);
        const size_t nfx =
            (size_t)kNoct * a.g.d.ne * a.g.d.nu * a.g.d.ncell_y * a.g.d.ncell_z * a.g.nap;
        const size_t nfy =
            (size_t)kNoct * a.g.d.ne * a.g.d.nu * a.g.d.ncell_x * a.g.d.ncell_z * a.g.nap;
        a.vi = (double*)q0;
        a.vo = (double*)(q0 + sb);
        a.fyz = q0 + 2 * sb;
        a.fxz = q0 + 2 * sb + align_up(fsz * nfx);
        a.part = (double*)(q0 + 2 * sb + align_up(fsz * nfx) + align_up(fsz * nfy));
        a.vo32 = vo32_bytes(a.g, s->opt.precision)
                     ? (float*)((char*)a.part + align_up(sizeof(double) * norms_scratch_doubles()))
                     : nullptr;
        a.t_admit = now_s();
        s->active.push_back(a);
        fresh.push_back(s->active.size() - 1);
        s->queue.pop_front();
      }
      if (s->active.empty() && !tick_in_flight) {
        s->cv.wait_for(lk, std::chrono::milliseconds(1));
        continue;
      }
    }
    // ---- enqueue this tick (the active list is only changed by this thread)
    std::vector<int> finishing;
    if (!s->active.empty()) {
      for (size_t k : fresh) {
        Active& a = s->active[k];
        s->h_params[a.slot] =
            make_params(a.g, a.vi, a.vo, a.at, a.mf, a.fyz, a.fxz, a.vo32, a.s, s->opt.precision);
        (void)hipMemcpyAsync(s->d_params + a.slot, s->h_params + a.slot, sizeof(SweepParams),
                             hipMemcpyHostToDevice, s->st);
        if (a.req.vi_pinned)
          (void)hipMemcpyAsync(a.vi, a.req.vi_pinned, sizeof(double) * a.g.state_doubles(),
                               hipMemcpyHostToDevice, s->st);
        else
          launch_init_state(a.g.d, a.vi, s->st);
        enqueue_operand_copy(a.g, a.vi, a.vo32, s->opt.precision, s->st);
        a.started = true;
      }
      std::vector<TickWork> work;
      for (size_t k = 0; k < s->active.size(); ++k) {
        Active& a = s->active[k];
        if (a.next >= a.s->nlaunch()) continue;  // already finishing
        work.push_back(TickWork{a.slot, &a.g, a.s, a.next});
        a.next++;
        if (a.next == a.s->nlaunch()) finishing.push_back((int)k);
      }
      if (!work.empty()) {
        const int k = enqueue_tick_forked(s->opt.precision, work, s->d_params, s->st,
                                          s->chain_st, s->fork_ev, s->join_ev, kChainStreams, s->opt.xcd_map != 0);
        if (k < 0) {  // reported with each request of the tick when it retires
          for (const TickWork& w : work)
            for (Active& a : s->active)
              if (a.slot == w.slot) a.err = -k;
        } else {
          s->nlaunch += k;
        }
        s->nticks++;
      }
      // norms of the problems whose last launch was in this tick, after it
      for (int& k : finishing) {
        Active& a = s->active[k];
        launch_norms(a.vi, a.vo, a.g.state_doubles(), a.part, s->d_res + 2 * a.slot, s->st);
        (void)hipMemcpyAsync(s->h_res + 2 * a.slot, s->d_res + 2 * a.slot, 2 * sizeof(double),
                             hipMemcpyDeviceToHost, s->st);
        k = a.slot;
      }
      (void)hipEventRecord(s->ev[cur], s->st);
    }
    // ---- retire the previous tick: its finishing problems are complete
    if (tick_in_flight) {
      const hipError_t e = hipEventSynchronize(s->ev[cur ^ 1]);
      std::lock_guard<std::mutex> lk(s->mu);
      for (int slot : pending)
        for (size_t k = 0; k < s->active.size(); ++k)
          if (s->active[k].slot == slot) {
            finish(s, s->active[k],
                   s->active[k].err ? s->active[k].err : e == hipSuccess ? MSW_OK : MSW_ERR_HIP);
            s->active.erase(s->active.begin() + k);
            break;
          }
    }
    pending = finishing;
    tick_in_flight = !finishing.empty() || !s->active.empty();
    if (!tick_in_flight) {
      // nothing was enqueued: make sure the last tick has retired
      (void)hipStreamSynchronize(s->st);
    }
    cur ^= 1;
    s->cv.notify_all();
  }
  (void)hipStreamSynchronize(s->st);
}

}  // namespace

extern "C" {

void msw_server_default_options(MswServerOptions* o) {
  if (!o) return;
  o->precision = MSW_PREC_FP64;
  o->hbm_budget_bytes = (size_t)32 << 30;
  o->max_active = 512;
  o->keep_outputs = 0;
  o->schedule_kind = 0;
  o->xcd_map = 1;
}

int msw_server_tune_from_batch(MswServer* s, const MswBatch* b) {
  if (!s || !b) return set_error(MSW_ERR_INVALID, "null server or batch");
  if (b->opt.precision != s->opt.precision)
    return set_error(MSW_ERR_INVALID, "batch and server precisions differ");
  std::lock_guard<std::mutex> lk(s->mu);
  s->class_nac.clear();
  for (const auto& kv : b->class_nac)
    if (kv.second) s->class_nac[kv.first] = kv.second;
  s->row_groups = b->row_groups;
  s->stream_target = b->stream_target;
  s->schedule_kind = b->opt.schedule_kind;
  return MSW_OK;
}

int msw_server_create(const MswServerOptions* opt, MswServer** out) {
  if (!out) return set_error(MSW_ERR_INVALID, "null server pointer");
  *out = nullptr;
  MswServerOptions o;
  msw_server_default_options(&o);
  if (opt) o = *opt;
  if (o.precision < MSW_PREC_FP64 || o.precision > MSW_PREC_MIXED_BF16)
    return set_error(MSW_ERR_INVALID, "unknown precision");
  if (o.max_active < 1 || o.hbm_budget_bytes < kAlign)
    return set_error(MSW_ERR_INVALID, "max_active >= 1 and a nonzero budget are required");
  if (o.schedule_kind != 0 && o.schedule_kind != 1)
    return set_error(MSW_ERR_INVALID, "schedule_kind must be 0 or 1");
  MswServer* s = new MswServer();
  s->opt = o;
  s->schedule_kind = o.schedule_kind;
  auto fail = [&](hipError_t e, const char* what) {
    msw_server_destroy(s);
    return hip_fail(e, what);
  };
  hipError_t e = hipGetDevice(&s->device);
  if (e != hipSuccess) return fail(e, "hipGetDevice");
  if ((e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking)) != hipSuccess)
    return fail(e, "server stream");
  for (int c = 0; c < kChainStreams; ++c) {
    if ((e = hipStreamCreateWithFlags(&s->chain_st[c], hipStreamNonBlocking)) != hipSuccess)
      return fail(e, "server stream");
    if ((e = hipEventCreateWithFlags(&s->join_ev[c], hipEventDisableTiming)) != hipSuccess)
      return fail(e, "server event");
  }
  if ((e = hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming)) != hipSuccess)
    return fail(e, "server event");
  if ((e = hipMalloc(&s->arena_base, o.hbm_budget_bytes)) != hipSuccess)
    return fail(e, "server arena");
  s->arena.reset(o.hbm_budget_bytes);
  const int slots = o.max_active + 2;
  if ((e = hipMalloc(&s->d_params, sizeof(SweepParams) * slots)) != hipSuccess)
    return fail(e, "server params");
  if ((e = hipHostMalloc(&s->h_params, sizeof(SweepParams) * slots, 0)) != hipSuccess)
    return fail(e, "server params");
  if ((e = hipMalloc(&s->d_res, 2 * sizeof(double) * slots)) != hipSuccess)
    return fail(e, "server results");
  if ((e = hipHostMalloc(&s->h_res, 2 * sizeof(double) * slots, 0)) != hipSuccess)
    return fail(e, "server results");
  for (int i = slots - 1; i >= 0; --i) s->free_slots.push_back(i);
  if ((e = hipEventCreateWithFlags(&s->ev[0], hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&s->ev[1], hipEventDisableTiming)) != hipSuccess)
    return fail(e, "server events");
  s->t0 = now_s();
  s->thr = std::thread(server_loop, s);
  *out = s;
  return MSW_OK;
}

int msw_server_submit(MswServer* s, const MswDims* d, long long tag, const double* vi_host) {
  if (!s || !d) return set_error(MSW_ERR_INVALID, "null submit argument");
  int rc = check_dims_mfma(*d);
  if (rc) return rc;
  Request q;
  q.tag = tag;
  q.d = *d;
  q.vi_pinned = nullptr;
  if (vi_host) {
    const size_t n = (size_t)d->ncell_x * d->ncell_y * d->ncell_z * d->ne * d->nm * d->nu;
    HIP_OK(hipHostMalloc(&q.vi_pinned, sizeof(double) * n, 0), "request staging");
    memcpy(q.vi_pinned, vi_host, sizeof(double) * n);
  }
  q.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->queue.push_back(q);
    s->in_flight++;
  }
  s->cv.notify_all();
  return MSW_OK;
}

int msw_server_poll(MswServer* s, MswResult* out, int max) {
  if (!s || (!out && max > 0)) return -set_error(MSW_ERR_INVALID, "bad poll arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  const int n = std::min<int>(max, (int)s->done.size());
  for (int i = 0; i < n; ++i) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + n);
  return n;
}

int msw_server_fetch(MswServer* s, long long tag, double* vo_host) {
  if (!s || !vo_host) return set_error(MSW_ERR_INVALID, "bad fetch arguments");
  Kept k;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->kept.find(tag);
    if (it == s->kept.end()) return set_error(MSW_ERR_INVALID, "no kept output for this tag");
    k = it->second;
    s->kept.erase(it);
  }
  const hipError_t e = hipMemcpy(vo_host, k.vo, sizeof(double) * k.n, hipMemcpyDeviceToHost);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->arena.release(k.off, k.bytes);
  }
  s->cv.notify_all();
  HIP_OK(e, "fetch copy");
  return MSW_OK;
}

int msw_server_drain(MswServer* s, double timeout_s) {
  if (!s) return set_error(MSW_ERR_INVALID, "null server");
  const double t_end = now_s() + timeout_s;
  std::unique_lock<std::mutex> lk(s->mu);
  while (s->in_flight > 0) {
    if (timeout_s > 0 && now_s() > t_end) return set_error(MSW_ERR_INVALID, "drain timed out");
    s->cv.wait_for(lk, std::chrono::milliseconds(2));
  }
  return MSW_OK;
}

double msw_server_now(const MswServer* s) { return s ? now_s() - s->t0 : 0.0; }

int msw_server_stats(const MswServer* s, double* out) {
  if (!s || !out) return set_error(MSW_ERR_INVALID, "bad stats arguments");
  MswServer* m = const_cast<MswServer*>(s);
  std::lock_guard<std::mutex> lk(m->mu);
  out[0] = (double)s->ndone;
  out[1] = (double)s->nlaunch;
  out[2] = (double)s->nticks;
  out[3] = (double)m->arena.peak();
  return MSW_OK;
}

int msw_server_destroy(MswServer* s) {
  if (!s) return MSW_OK;
  if (s->thr.joinable()) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv.notify_all();
    s->thr.join();
  }
  if (s->st) (void)hipStreamSynchronize(s->st);
  for (auto& q : s->queue)
    if (q.vi_pinned) (void)hipHostFree(q.vi_pinned);
  for (auto& a : s->active)
    if (a.req.vi_pinned) (void)hipHostFree(a.req.vi_pinned);
  for (auto& kv : s->mats) {
    (void)hipFree(kv.second.first);
    (void)hipFree(kv.second.second);
  }
  free_schedules(s->scheds);
  for (auto ev : s->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (s->arena_base) (void)hipFree(s->arena_base);
  if (s->d_params) (void)hipFree(s->d_params);
  if (s->h_params) (void)hipHostFree(s->h_params);
  if (s->d_res) (void)hipFree(s->d_res);
  if (s->h_res) (void)hipHostFree(s->h_res);
  if (s->st) (void)hipStreamDestroy(s->st);
  for (int c = 0; c < kChainStreams; ++c) {
    if (s->chain_st[c]) (void)hipStreamDestroy(s->chain_st[c]);
    if (s->join_ev[c]) (void)hipEventDestroy(s->join_ev[c]);
  }
  if (s->fork_ev) (void)hipEventDestroy(s->fork_ev);
  delete s;
  return MSW_OK;
}

}  // extern "C"
