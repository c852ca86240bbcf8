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
// launches), keeps the fastest, and caches it for the process.  The job's
// data are not touched.  mag_tune_set switches to a fixed shape or to the
// default shape (256, 1) without tuning.
#include "mag_capi_util.h"
#include "mag_gpu.h"
#include "mag_internal.h"
#include "mag_job_impl.h"

#include <map>
#include <mutex>
#include <tuple>

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

// launch one half-sweep of the job's current kernel with shape sh on the
// scratch arrays
void launch_scratch(const mag_job* j, bool fp32, void* U, void* G, double* part, JobCtl* ctl,
                    int parity, Shape sh) {
  if (j->prm.group == MAG_SU2) {
    if (fp32) launch_su2_half<float>((float*)U, (float*)G, j->g, parity, j->prm.omega, part, ctl, sh, j->st);
    else launch_su2_half<double>((double*)U, nullptr, j->g, parity, j->prm.omega, part, ctl, sh, j->st);
  } else {
    if (fp32) launch_su3_half<float>((float*)U, (float*)G, j->g, parity, j->prm.omega, part, ctl, sh, j->st);
    else launch_su3_half<double>((double*)U, nullptr, j->g, parity, j->prm.omega, part, ctl, sh, j->st);
  }
}

int tune(const mag_job* j, bool fp32, Shape* out) {
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  const size_t ubytes = j->nreal * esz;
  const size_t gbytes = (size_t)2 * j->g.Vh * j->nr * sizeof(float);
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
  if (e == hipSuccess) e = hipMemsetAsync(ctl, 0, sizeof(JobCtl), j->st);  // done = 0
  if (e != hipSuccess) {
    cleanup();
    return fail(MAG_ERR_NOMEM, "autotune scratch: %s", hipGetErrorString(e));
  }
  float best = 0.f;
  Shape bs;
  for (const Shape& sh : kShapes) {
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
    if (best == 0.f || ms < best) {
      best = ms;
      bs = sh;
    }
  }
  cleanup();
  if (rc == MAG_OK) *out = bs;
  return rc;
}

}  // namespace

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
  CHK(tune(j, fp32, &sh));
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
  return MAG_OK;
}

}  // extern "C"
