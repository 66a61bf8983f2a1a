// Drives msw_batch_* on the stand-in runtime (fake_hip.cpp) under ASan/UBSan:
// batches of the GPU tests' shapes and of random uneven shapes, in every
// precision, graph and eager, with autotune (random timings), reruns,
// describe, a buffer swap that forces a rebuild, and destroy.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "msw_batch.h"
#include "msw_plan.h"
extern std::mt19937 g_rng;
extern long g_launches, g_checked_wg;

static const int SHAPES[][7] = {
    {9, 7, 3, 2, 16, 64, 4}, {17, 3, 5, 1, 4, 32, 2}, {3, 20, 9, 3, 4, 13, 3},
    {16, 16, 4, 2, 16, 32, 4}, {5, 5, 5, 1, 8, 40, 1}, {20, 9, 6, 2, 16, 64, 4}};

static std::vector<MswDims> random_dims(std::mt19937& r, int n) {
  std::vector<MswDims> v;
  for (int i = 0; i < n; ++i) {
    MswDims d;
    d.ncell_x = 1 + r() % 24; d.ncell_y = 1 + r() % 24; d.ncell_z = 1 + r() % 16;
    d.ne = 1 + r() % 4; const int nms[3] = {4, 8, 16}; d.nm = nms[r() % 3];
    d.na = 1 + r() % 64; d.nu = 1 + r() % 4;
    v.push_back(d);
  }
  return v;
}

static int run_batch(std::vector<MswDims> dims, int prec, int graph, unsigned seed, bool tune) {
  MswBatchOptions o;
  msw_batch_default_options(&o);
  o.precision = prec; o.use_graph = graph; o.stagger = seed & 1; o.schedule_kind = (seed >> 1) & 1;
  const int rgs[4] = {0, 1, 2, 4};
  o.row_groups = rgs[(seed >> 2) % 4]; o.stream_target = seed % 3 == 0 ? 32 : 0;
  o.xcd_map = (seed >> 4) & 1 ? 0 : 1;
  MswBatch* b = nullptr;
  int rc = msw_batch_create((int)dims.size(), dims.data(), &o, &b);
  if (rc) { printf("create rc %d: %s\n", rc, msw_last_error()); return rc; }
  std::vector<void*> bufs;
  for (size_t i = 0; i < dims.size(); ++i) {
    const MswDims& d = dims[i];
    size_t n = (size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.nm * d.nu;
    void *vi, *vo, *a, *m;
    hipMalloc(&vi, n * 8); hipMalloc(&vo, n * 8);
    hipMalloc(&a, (size_t)8 * d.nm * d.na * 8); hipMalloc(&m, (size_t)8 * d.nm * d.na * 8);
    bufs.insert(bufs.end(), {vi, vo, a, m});
    rc = msw_batch_set_problem(b, (int)i, (double*)vi, (double*)vo, (double*)a, (double*)m);
    if (rc) { printf("set rc %d %s\n", rc, msw_last_error()); return rc; }
  }
  hipStream_t st; hipStreamCreateWithFlags(&st, 0);
  rc = tune ? msw_batch_autotune(b, st, 1 + seed % 2) : msw_batch_run(b, st);
  if (rc) { printf("autotune/run rc %d: %s\n", rc, msw_last_error()); }
  for (int r = 0; r < 2 && !rc; ++r) rc = msw_batch_run(b, st);
  char desc[1 << 16];
  msw_batch_describe(b, desc, sizeof(desc));
  // swap a problem's buffers: forces a rebuild while the last run is "in flight"
  if (!rc) {
    const MswDims& d = dims[0];
    size_t n = (size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.nm * d.nu;
    void *vi, *vo;
    hipMalloc(&vi, n * 8); hipMalloc(&vo, n * 8);
    bufs.insert(bufs.end(), {vi, vo});
    rc = msw_batch_set_problem(b, 0, (double*)vi, (double*)vo, (double*)bufs[2], (double*)bufs[3]);
    if (!rc) rc = msw_batch_run(b, st);
  }
  msw_batch_destroy(b);
  for (void* p : bufs) hipFree(p);
  hipStreamDestroy(st);
  return rc;
}

int main(int argc, char** argv) {
  const int iters = argc > 1 ? atoi(argv[1]) : 20;
  std::vector<MswDims> test_dims;
  for (auto& s : SHAPES) test_dims.push_back(MswDims{s[0], s[1], s[2], s[3], s[4], s[5], s[6]});
  int fails = 0;
  for (int it = 0; it < iters; ++it) {
    g_rng.seed(it);
    for (int prec : {MSW_PREC_FP64, MSW_PREC_MIXED_FP32, MSW_PREC_MIXED_BF16})
      for (int graph : {1, 0}) {
        fails += run_batch(test_dims, prec, graph, it, true) != 0;
        std::mt19937 r(1000 + it);
        fails += run_batch(random_dims(r, 1 + r() % 40), prec, graph, it + 7, it % 3 == 0) != 0;
      }
    printf("iter %d: launches %ld workgroups checked %ld fails %d\n", it, g_launches, g_checked_wg, fails);
    fflush(stdout);
  }
  return fails ? 1 : 0;
}
