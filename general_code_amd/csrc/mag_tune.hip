// mag_tune.hip -- launch-shape autotuner of the link-form half-sweep.
//
// A half-sweep kernel can be launched with NT = 64, 128, 256 or 512 threads
// per workgroup and SPT = 1 or 2 tiles of NT lanes per thread.  Its
// partials are per wave (mag_sweep.hip), so every shape gives bit-identical
// results and the choice is purely a matter of speed: small lattices want
// small workgroups (more of them, every CU busy), large ones fewer, bigger
// workgroups or two tiles per thread (more loads in flight per wave).
//
// The first sweep of a job on a (group, precision, lattice) not seen
// before times every shape on a scratch copy of the job's own links (8
// back-to-back half-sweeps per shape between HIP events, after 2 warm-up
// launches; three rounds over the shapes, each shape's fastest round),
// keeps the fastest, and caches it for the process.  The job's
// data are not touched.  Batched launches are tuned on their own (see
// tune_batch).  mag_tune_set switches to a fixed shape or to the default
// shape (256, 1) without tuning.
#include "mag_capi_util.h"
#include "mag_gpu.h"
#include "mag_internal.h"
#include "mag_job_impl.h"

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace mag {
namespace {

const Shape kShapes[] = {{64, 1},  {64, 2},  {128, 1}, {128, 2},
                         {256, 1}, {256, 2}, {512, 1}, {512, 2}};

std::mutex g_mu;
int g_mode = MAG_TUNE_AUTO;
Shape g_fixed;
std::map<std::tuple<int, int, int, int, int, int, int>, Shape> g_cache;

bool valid_shape(int nt, int spt) {
  for (const Shape& s : kShapes)
    if (s.nt == nt && s.spt == spt) return true;
  return false;
}

// control block of the scratch sweeps: not done, no cap in reach
const JobCtl kScratchCtl = [] {
  JobCtl c{};
  c.max_sweeps = 1 << 30;
  return c;
}();

// launch one half-sweep of the job's current kernel with shape sh on the
// scratch arrays (FP32: the red half accumulates the site gauge)
void launch_scratch(const mag_job* j, bool fp32, void* U, void* G, double* part, JobCtl* ctl,
                    int parity, Shape sh) {
  if (j->prm.group == MAG_SU2) {
    if (fp32) launch_su2_half<float>((float*)U, parity == 0 ? (float*)G : nullptr, j->g, parity, j->prm.omega, part, ctl, sh, j->st);
    else launch_su2_half<double>((double*)U, nullptr, j->g, parity, j->prm.omega, part, ctl, sh, j->st);
  } else {
    if (fp32) launch_su3_half<float>((float*)U, parity == 0 ? (float*)G : nullptr, j->g, parity, j->prm.omega, part, ctl, sh, j->st);
    else launch_su3_half<double>((double*)U, nullptr, j->g, parity, j->prm.omega, part, ctl, sh, j->st);
  }
}

int tune(const mag_job* j, bool fp32, Shape* out) {
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  const size_t ubytes = j->nreal * esz;
  const size_t gbytes = (size_t)2 * j->g.hp * j->nr * sizeof(float);
  void *U = nullptr, *G = nullptr;
  double* part = nullptr;
  JobCtl* ctl = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MAG_OK;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(j->st);
    if (U) (void)hipFree(U);
    if (G) (void)hipFree(G);
    if (part) (void)hipFree(part);
    if (ctl) (void)hipFree(ctl);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  hipError_t e = hipMalloc(&U, ubytes);
  if (e == hipSuccess && fp32) e = hipMalloc(&G, gbytes);
  if (e == hipSuccess) e = hipMalloc((void**)&part, (size_t)2 * j->nw * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&ctl, sizeof(JobCtl));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess)
    e = hipMemcpyAsync(U, fp32 ? (const void*)j->W32 : (const void*)j->U64, ubytes,
                       hipMemcpyDeviceToDevice, j->st);
  if (e == hipSuccess && fp32) e = hipMemcpyAsync(G, j->G32, gbytes, hipMemcpyDeviceToDevice, j->st);
  if (e == hipSuccess) e = hipMemcpyAsync(ctl, &kScratchCtl, sizeof(JobCtl), hipMemcpyHostToDevice, j->st);
  if (e != hipSuccess) {
    cleanup();
    return fail(MAG_ERR_NOMEM, "autotune scratch: %s", hipGetErrorString(e));
  }
  // kTuneRounds rounds over all shapes, each shape's fastest round kept:
  // neighbouring shapes differ by a few per cent, about the spread of one
  // timing, and a single round picked a shape 2-3% off the best at 32^4
  constexpr int kNs = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
  constexpr int kTuneRounds = 3;
  float tmin[kNs] = {};
  for (int r = 0; r < kTuneRounds && rc == MAG_OK; r++) {
    for (int k = 0; k < kNs; k++) {
      const Shape sh = kShapes[k];
      for (int i = 0; i < 2; i++) launch_scratch(j, fp32, U, G, part, ctl, i & 1, sh);
      (void)hipEventRecord(e0, j->st);
      for (int i = 0; i < 8; i++) launch_scratch(j, fp32, U, G, part, ctl, i & 1, sh);
      (void)hipEventRecord(e1, j->st);
      e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(MAG_ERR_HIP, "autotune: %s", hipGetErrorString(e));
        break;
      }
      if (r == 0 || ms < tmin[k]) tmin[k] = ms;
    }
  }
  cleanup();
  if (rc == MAG_OK) {
    int b = 0;
    for (int k = 1; k < kNs; k++)
      if (tmin[k] < tmin[b]) b = k;
    *out = kShapes[b];
  }
  return rc;
}

// Batched link-form half-sweeps (mag_batch.hip) concatenate many jobs'
// grids, so the best shape of one job alone is not that of the batch: the
// batch is tuned on its own launch.  Scratch copies of the class's jobs (in
// order, up to kBatchScratch bytes, at least one job) are swept with each
// shape through the batched kernel (2 warm-up + 6 timed launches); the
// fastest is cached per (group, precision, device, log2 of the class's
// total lanes).
constexpr size_t kBatchScratch = (size_t)1 << 30;

int tune_batch(int group, bool fp32, const std::vector<mag_job*>& jobs, Shape* out) {
  const hipStream_t st = jobs[0]->st;
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  std::vector<void*> bufs;
  std::vector<BatchJob> tab;
  BatchJob* d_tab = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(st);
    for (void* b : bufs) (void)hipFree(b);
    if (d_tab) (void)hipFree(d_tab);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  auto alloc = [&](size_t n) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, n) != hipSuccess) return nullptr;
    bufs.push_back(p);
    return p;
  };
  size_t used = 0;
  // scratch stays within a quarter of the free device memory as well
  size_t cap = kBatchScratch, free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b / 4 < cap) cap = free_b / 4;
  hipError_t e = hipSuccess;
  for (const mag_job* j : jobs) {
    const size_t ub = j->nreal * esz, gb = fp32 ? (size_t)2 * j->g.hp * j->nr * sizeof(float) : 0;
    if (!tab.empty() && used + ub + gb > cap) break;
    used += ub + gb;
    BatchJob b{};
    b.U = alloc(ub);
    b.G = fp32 ? alloc(gb) : nullptr;
    b.part = (double*)alloc((size_t)2 * j->nw * 3 * sizeof(double));
    b.ctl = (JobCtl*)alloc(sizeof(JobCtl));
    if (!b.U || (fp32 && !b.G) || !b.part || !b.ctl) {
      cleanup();
      return fail(MAG_ERR_NOMEM, "batch autotune: scratch allocation failed");
    }
    b.g = j->g;
    b.omega = j->prm.omega;
    b.nw = j->nw;
    if (e == hipSuccess)
      e = hipMemcpyAsync(b.U, fp32 ? (const void*)j->W32 : (const void*)j->U64, ub,
                         hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && fp32) e = hipMemcpyAsync(b.G, j->G32, gb, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(b.ctl, &kScratchCtl, sizeof(JobCtl), hipMemcpyHostToDevice, st);
    tab.push_back(b);
  }
  if (e == hipSuccess) e = hipMalloc((void**)&d_tab, tab.size() * sizeof(BatchJob));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    cleanup();
    return fail(MAG_ERR_HIP, "batch autotune: %s", hipGetErrorString(e));
  }
  float best = 0.f;
  Shape bs;
  for (const Shape& sh : kShapes) {
    const size_t tile = (size_t)sh.nt * sh.spt;
    int nblocks = 0;
    for (BatchJob& b : tab) {
      b.block0 = nblocks;
      nblocks += (int)((sweep_lanes(b.g, group) + tile - 1) / tile);
    }
    e = hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(BatchJob), hipMemcpyHostToDevice, st);
    for (int i = 0; i < 2 && e == hipSuccess; i++)
      launch_half_batch(group, fp32, sh, d_tab, (int)tab.size(), nblocks, i & 1, st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int i = 0; i < 6 && e == hipSuccess; i++)
      launch_half_batch(group, fp32, sh, d_tab, (int)tab.size(), nblocks, i & 1, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) break;
    if (best == 0.f || ms < best) {
      best = ms;
      bs = sh;
    }
  }
  cleanup();
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "batch autotune: %s", hipGetErrorString(e));
  *out = bs;
  return MAG_OK;
}

std::map<std::tuple<int, int, int, int>, Shape> g_batch_cache;

}  // namespace

int batch_shape(int group, bool fp32, const std::vector<mag_job*>& jobs, Shape* out) {
  size_t lanes = 0;
  for (const mag_job* j : jobs) lanes += sweep_lanes(j->g, group);
  int bucket = 0;
  while (((size_t)2 << bucket) <= lanes) bucket++;
  std::tuple<int, int, int, int> key{group, fp32 ? 1 : 0, jobs[0]->prm.device, bucket};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_mode == MAG_TUNE_FIXED) {
      *out = g_fixed;
      return MAG_OK;
    }
    if (g_mode == MAG_TUNE_DEFAULT) {
      *out = Shape{};
      return MAG_OK;
    }
    auto it = g_batch_cache.find(key);
    if (it != g_batch_cache.end()) {
      *out = it->second;
      return MAG_OK;
    }
  }
  // a failed tuning run (e.g. no memory for its scratch copies) falls back
  // to the default shape, which gives the same results as any other; its
  // error is cleared so that it is not reported by a later HIP call
  Shape sh;
  if (tune_batch(group, fp32, jobs, &sh) != MAG_OK) {
    (void)hipGetLastError();
    sh = Shape{};
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_batch_cache[key] = sh;
  *out = sh;
  return MAG_OK;
}

int job_shape(const mag_job* j, bool fp32, Shape* out) {
  std::tuple<int, int, int, int, int, int, int> key{j->prm.group, fp32 ? 1 : 0, j->prm.device,
                                                    j->g.L0, j->g.L1, j->g.L2, j->g.L3};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_mode == MAG_TUNE_FIXED) {
      *out = g_fixed;
      return MAG_OK;
    }
    if (g_mode == MAG_TUNE_DEFAULT) {
      *out = Shape{};
      return MAG_OK;
    }
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      *out = it->second;
      return MAG_OK;
    }
  }
  Shape sh;
  if (tune(j, fp32, &sh) != MAG_OK) {  // as for batches: the default shape
    (void)hipGetLastError();
    sh = Shape{};
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_cache[key] = sh;
  *out = sh;
  return MAG_OK;
}

}  // namespace mag

extern "C" {

int mag_tune_set(int mode, int nt, int spt) {
  if (mode != MAG_TUNE_AUTO && mode != MAG_TUNE_FIXED && mode != MAG_TUNE_DEFAULT)
    return mag::fail(MAG_ERR_ARG, "unknown tuning mode %d", mode);
  if (mode == MAG_TUNE_FIXED && !mag::valid_shape(nt, spt))
    return mag::fail(MAG_ERR_ARG, "unsupported shape (%d threads, %d tiles)", nt, spt);
  std::lock_guard<std::mutex> lk(mag::g_mu);
  mag::g_mode = mode;
  mag::g_fixed = mag::Shape{nt, spt};
  return MAG_OK;
}

int mag_tune_clear(void) {
  std::lock_guard<std::mutex> lk(mag::g_mu);
  mag::g_cache.clear();
  mag::g_batch_cache.clear();
  return MAG_OK;
}

}  // extern "C"
