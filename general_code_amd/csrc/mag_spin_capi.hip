// mag_spin_capi.hip -- C ABI of the spin-form MAG path: the call surface of
// the reference's lib/cpu/include/mag.h (generate_spins_uniform,
// make_simulated_annealing, make_maximization_final/_approximate,
// gauge_tranformation_spins, MAG_functional_su2[_spin]) on the GPU.
//
// A session keeps one SU(2) configuration and its spins resident in HBM so
// that a caller can chain annealing -> maximisation -> transformation
// without host round trips; the one-shot functions wrap a session.
// Drivers advance on device-resident state (SpinCtl): the host enqueues
// iterations in chunks and polls only between chunks.
#include "mag_capi_util.h"
#include "mag_gpu.h"
#include "mag_internal.h"

#include <cmath>
#include <cstring>
#include <new>

using mag::Geom;
using mag::SpinCtl;

struct mag_spin_session {
  int device = 0;
  int dims[4] = {0, 0, 0, 0};
  Geom g;
  int64_t V = 0;
  int nb = 0;
  hipStream_t st = nullptr;
  double* U = nullptr;  // links, FP64 parity-split SoA
  double* S = nullptr;  // spins, FP64 SoA (parity, component)
  double* part = nullptr;
  double* out3 = nullptr;
  SpinCtl* ctl = nullptr;
  SpinCtl* ctl_h = nullptr;
  double* out3_h = nullptr;
  uint32_t hb_sweep = 0;  // heat-bath sweeps done (Philox counter)
  bool links = false, spins = false;
};

namespace {
using mag::fail;

constexpr int kIterChunk = 8;  // driver iterations enqueued between polls

int check_dims4(const int* d, int64_t* V) {
  if (!d) return fail(MAG_ERR_ARG, "dims is NULL");
  int64_t v = 1;
  for (int i = 0; i < 4; i++) {
    if (d[i] < 2 || (d[i] & 1))
      return fail(MAG_ERR_DIMS, "extent %d of dimension %d must be even and >= 2", d[i], i);
    v *= d[i];
  }
  if (v / 2 > (int64_t)0x7fffffff / 8) return fail(MAG_ERR_DIMS, "lattice too large");
  *V = v;
  return MAG_OK;
}

void free_session(mag_spin_session* s) {
  if (!s) return;
  hipFree(s->U);
  hipFree(s->S);
  hipFree(s->part);
  hipFree(s->out3);
  hipFree(s->ctl);
  if (s->ctl_h) hipHostFree(s->ctl_h);
  if (s->out3_h) hipHostFree(s->out3_h);
  if (s->st) hipStreamDestroy(s->st);
  delete s;
}

int ready(mag_spin_session* s, bool need_spins) {
  if (!s) return fail(MAG_ERR_ARG, "NULL session");
  if (!s->links) return fail(MAG_ERR_STATE, "no links uploaded");
  if (need_spins && !s->spins) return fail(MAG_ERR_STATE, "no spins set");
  HIPCHK(hipSetDevice(s->device));
  return MAG_OK;
}

int read_ctl(mag_spin_session* s) {
  HIPCHK(hipMemcpyAsync(s->ctl_h, s->ctl, sizeof(SpinCtl), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return MAG_OK;
}

int reset_ctl(mag_spin_session* s, int max_iter, double tol_max, double tol_avg, int digits) {
  std::memset(s->ctl_h, 0, sizeof(SpinCtl));
  s->ctl_h->max_iter = max_iter;
  s->ctl_h->tol_max = tol_max;
  s->ctl_h->tol_avg = tol_avg;
  s->ctl_h->digits = digits;
  s->ctl_h->scale = powf(10.f, (float)digits);
  HIPCHK(hipMemcpyAsync(s->ctl, s->ctl_h, sizeof(SpinCtl), hipMemcpyHostToDevice, s->st));
  return MAG_OK;
}

// one checkerboard sweep of a mode: parity 0 then parity 1
void sweep(mag_spin_session* s, int mode, double param, uint64_t seed, uint32_t counter) {
  mag::launch_spin_half(mode, s->U, s->S, s->g, 0, param, seed, counter, s->part, s->ctl, s->st);
  mag::launch_spin_half(mode, s->U, s->S, s->g, 1, param, seed, counter, s->part, s->ctl, s->st);
}

int stage_alloc(double** p, size_t n) {
  HIPCHK(hipMalloc((void**)p, n * sizeof(double)));
  return MAG_OK;
}

}  // namespace

extern "C" {

int mag_spin_session_create(const int* dims, int device, mag_spin_session** out) {
  if (!out) return fail(MAG_ERR_ARG, "out is NULL");
  *out = nullptr;
  int64_t V;
  CHK(check_dims4(dims, &V));
  mag_spin_session* s = new (std::nothrow) mag_spin_session();
  if (!s) return fail(MAG_ERR_NOMEM, "host allocation failed");
  s->device = device;
  for (int i = 0; i < 4; i++) s->dims[i] = dims[i];
  s->V = V;
  s->g = mag::make_geom(dims);
  s->nb = mag::nblocks(s->g);
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&s->U, (size_t)8 * s->g.hp * 4 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&s->S, (size_t)V * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&s->part, (size_t)2 * s->nb * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&s->out3, 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&s->ctl, sizeof(SpinCtl));
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->ctl_h, sizeof(SpinCtl));
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->out3_h, 3 * sizeof(double));
  if (e != hipSuccess) {
    free_session(s);
    return fail(e == hipErrorOutOfMemory ? MAG_ERR_NOMEM : MAG_ERR_HIP, "session setup failed: %s",
                hipGetErrorString(e));
  }
  *out = s;
  return MAG_OK;
}

int mag_spin_session_destroy(mag_spin_session* s) {
  if (!s) return MAG_OK;
  hipSetDevice(s->device);
  hipStreamSynchronize(s->st);
  free_session(s);
  return MAG_OK;
}

int mag_spin_session_upload_links(mag_spin_session* s, const double* links) {
  if (!s || !links) return fail(MAG_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(s->device));
  const size_t n = (size_t)s->V * 16;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(links[i])) return fail(MAG_ERR_NONFINITE, "links hold NaN or Inf");
  double* st = nullptr;
  CHK(stage_alloc(&st, n));
  hipError_t e = hipMemcpyAsync(st, links, n * sizeof(double), hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess) {
    mag::launch_su2_aos_to_soa<double>(st, s->U, s->g, s->st);
    e = hipStreamSynchronize(s->st);
  }
  hipFree(st);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
  s->links = true;
  return MAG_OK;
}

int mag_spin_session_download_links(mag_spin_session* s, double* links) {
  CHK(ready(s, false));
  if (!links) return fail(MAG_ERR_ARG, "links is NULL");
  const size_t n = (size_t)s->V * 16;
  double* st = nullptr;
  CHK(stage_alloc(&st, n));
  mag::launch_su2_soa_to_aos<double>(s->U, st, s->g, s->st);
  hipError_t e = hipMemcpyAsync(links, st, n * sizeof(double), hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  hipFree(st);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "download failed: %s", hipGetErrorString(e));
  return MAG_OK;
}

int mag_spin_session_set_spins(mag_spin_session* s, const double* spins) {
  if (!s || !spins) return fail(MAG_ERR_ARG, "NULL argument");
  HIPCHK(hipSetDevice(s->device));
  const size_t n = (size_t)s->V * 3;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(spins[i])) return fail(MAG_ERR_NONFINITE, "spins hold NaN or Inf");
  double* st = nullptr;
  CHK(stage_alloc(&st, n));
  hipError_t e = hipMemcpyAsync(st, spins, n * sizeof(double), hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess) {
    mag::launch_spins_layout(st, s->S, s->g, 1, s->st);
    e = hipStreamSynchronize(s->st);
  }
  hipFree(st);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "spin upload failed: %s", hipGetErrorString(e));
  s->spins = true;
  return MAG_OK;
}

int mag_spin_session_get_spins(mag_spin_session* s, double* spins) {
  if (!s) return fail(MAG_ERR_ARG, "NULL session");
  if (!s->spins) return fail(MAG_ERR_STATE, "no spins set");
  HIPCHK(hipSetDevice(s->device));
  if (!spins) return fail(MAG_ERR_ARG, "spins is NULL");
  const size_t n = (size_t)s->V * 3;
  double* st = nullptr;
  CHK(stage_alloc(&st, n));
  mag::launch_spins_layout(st, s->S, s->g, 0, s->st);
  hipError_t e = hipMemcpyAsync(spins, st, n * sizeof(double), hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  hipFree(st);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "spin download failed: %s", hipGetErrorString(e));
  return MAG_OK;
}

int mag_spin_session_generate_spins(mag_spin_session* s, uint64_t seed) {
  if (!s) return fail(MAG_ERR_ARG, "NULL session");
  HIPCHK(hipSetDevice(s->device));
  double* st = nullptr;
  CHK(stage_alloc(&st, (size_t)s->V * 3));
  mag::launch_spins_uniform(st, s->g, seed, s->st);
  mag::launch_spins_layout(st, s->S, s->g, 1, s->st);
  hipError_t e = hipStreamSynchronize(s->st);
  hipFree(st);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "spin generation failed: %s", hipGetErrorString(e));
  s->spins = true;
  return MAG_OK;
}

int mag_spin_session_functional_spin(mag_spin_session* s, double* F) {
  CHK(ready(s, true));
  if (!F) return fail(MAG_ERR_ARG, "F is NULL");
  mag::launch_spin_functional(s->U, s->S, s->g, s->part, nullptr, s->st);
  mag::launch_spin_reduce(s->part, 2 * s->nb, mag::SRED_PLAIN, 1.0 / s->V, 1.0 / (4.0 * s->V),
                          s->ctl, s->out3, s->st);
  HIPCHK(hipMemcpyAsync(s->out3_h, s->out3, 3 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  *F = s->out3_h[0];
  return MAG_OK;
}

int mag_spin_session_functional(mag_spin_session* s, double* F) {
  CHK(ready(s, false));
  if (!F) return fail(MAG_ERR_ARG, "F is NULL");
  mag::launch_su2_functional<double>(s->U, s->g, s->part, s->st);
  mag::launch_spin_reduce(s->part, 2 * s->nb, mag::SRED_PLAIN, 1.0 / s->V, 1.0 / (4.0 * s->V),
                          s->ctl, s->out3, s->st);
  HIPCHK(hipMemcpyAsync(s->out3_h, s->out3, 3 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  *F = s->out3_h[0];
  return MAG_OK;
}

int mag_spin_session_sweep(mag_spin_session* s, int mode, double param, uint64_t seed,
                           double* dmax, double* davg) {
  CHK(ready(s, true));
  if (mode < mag::SPIN_OR || mode > mag::SPIN_HEATBATH) return fail(MAG_ERR_ARG, "bad mode %d", mode);
  if (mode == mag::SPIN_HEATBATH && !(param > 0)) return fail(MAG_ERR_ARG, "temperature must be > 0");
  CHK(reset_ctl(s, 1, 0, 0, 0));
  sweep(s, mode, param, seed, s->hb_sweep);
  if (mode == mag::SPIN_HEATBATH) s->hb_sweep++;
  mag::launch_spin_reduce(s->part, 2 * s->nb, mag::SRED_PLAIN, 1.0 / s->V, 1.0, s->ctl, s->out3,
                          s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(s->out3_h, s->out3, 3 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  if (davg) *davg = s->out3_h[1];
  if (dmax) *dmax = s->out3_h[2];
  return MAG_OK;
}

// make_simulated_annealing (mag.cpp:1585-1603): thermalisation at T_init,
// then heat bath + OR_steps reflection sweeps per temperature while
// T > T_final/2, T lowered by T_step/4 inside [1.2, 1.6 + T_step].
int mag_spin_session_simulated_annealing(mag_spin_session* s, double T_init, double T_final,
                                         double T_step, int OR_steps, int thermalization_steps,
                                         uint64_t seed) {
  CHK(ready(s, true));
  if (!(T_init > 0) || !(T_step > 0) || OR_steps < 0 || thermalization_steps < 0)
    return fail(MAG_ERR_ARG, "bad annealing schedule");
  CHK(reset_ctl(s, 1, 0, 0, 0));
  for (int i = 0; i < thermalization_steps; i++) sweep(s, mag::SPIN_HEATBATH, T_init, seed, s->hb_sweep++);
  double T = T_init;
  int enq = 0;
  while (T > T_final / 2) {
    if (!(T > 0)) break;
    sweep(s, mag::SPIN_HEATBATH, T, seed, s->hb_sweep++);
    for (int i = 0; i < OR_steps; i++) sweep(s, mag::SPIN_OR, 0.0, 0, 0);
    if (T <= 1.6 + T_step && T >= 1.2) T -= T_step / 4;
    else T -= T_step;
    if (++enq % 64 == 0) HIPCHK(hipStreamSynchronize(s->st));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->st));
  return MAG_OK;
}

// make_maximization_final (mag.cpp:1800-1816):
// do { relax; converged = max < tol_max && avg < tol_avg; OR x OR_steps }
// while (!converged), capped at max_iter iterations.
int mag_spin_session_maximize_final(mag_spin_session* s, int OR_steps, double tol_max,
                                    double tol_avg, int max_iter, int* iters) {
  CHK(ready(s, true));
  if (OR_steps < 0 || max_iter < 1) return fail(MAG_ERR_ARG, "bad argument");
  CHK(reset_ctl(s, max_iter, tol_max, tol_avg, 0));
  for (;;) {
    for (int k = 0; k < kIterChunk; k++) {
      sweep(s, mag::SPIN_RELAX, 0.0, 0, 0);
      mag::launch_spin_reduce(s->part, 2 * s->nb, mag::SRED_RELAX_FINAL, 1.0 / s->V, 1.0, s->ctl,
                              nullptr, s->st);
      for (int i = 0; i < OR_steps; i++) sweep(s, mag::SPIN_OR, 0.0, 0, 0);
    }
    HIPCHK(hipGetLastError());
    CHK(read_ctl(s));
    if (s->ctl_h->stop_next || s->ctl_h->done) break;
  }
  if (iters) *iters = s->ctl_h->iter;
  if (!s->ctl_h->converged) return fail(MAG_ERR_NOT_CONVERGED, "no convergence in %d iterations", max_iter);
  return MAG_OK;
}

// make_maximization_approximate (mag.cpp:1733-1750):
// do { relax; OR x OR_steps; F = F_spin; equal = trunc(powf(10,d) F) ==
// trunc(powf(10,d) F_old) && same sign } while (!equal)
int mag_spin_session_maximize_approximate(mag_spin_session* s, int OR_steps,
                                          int tolerance_digits, int max_iter, int* iters) {
  CHK(ready(s, true));
  if (OR_steps < 0 || max_iter < 1) return fail(MAG_ERR_ARG, "bad argument");
  CHK(reset_ctl(s, max_iter, 0, 0, tolerance_digits));
  const double invV = 1.0 / s->V, invF = 1.0 / (4.0 * s->V);
  mag::launch_spin_functional(s->U, s->S, s->g, s->part, s->ctl, s->st);
  mag::launch_spin_reduce(s->part, 2 * s->nb, mag::SRED_FUNC_INIT, invV, invF, s->ctl, nullptr, s->st);
  for (;;) {
    for (int k = 0; k < kIterChunk; k++) {
      sweep(s, mag::SPIN_RELAX, 0.0, 0, 0);
      mag::launch_spin_reduce(s->part, 2 * s->nb, mag::SRED_RELAX_PLAIN, invV, invF, s->ctl,
                              nullptr, s->st);
      for (int i = 0; i < OR_steps; i++) sweep(s, mag::SPIN_OR, 0.0, 0, 0);
      mag::launch_spin_functional(s->U, s->S, s->g, s->part, s->ctl, s->st);
      mag::launch_spin_reduce(s->part, 2 * s->nb, mag::SRED_FUNC_APPROX, invV, invF, s->ctl,
                              nullptr, s->st);
    }
    HIPCHK(hipGetLastError());
    CHK(read_ctl(s));
    if (s->ctl_h->stop_next || s->ctl_h->done) break;
  }
  if (iters) *iters = s->ctl_h->iter;
  if (!s->ctl_h->converged) return fail(MAG_ERR_NOT_CONVERGED, "no convergence in %d iterations", max_iter);
  return MAG_OK;
}

// gauge_tranformation_spins (mag.cpp:1045-1068) on the resident links
int mag_spin_session_gauge_transform(mag_spin_session* s) {
  CHK(ready(s, true));
  mag::launch_spin_gauge_transform(s->U, s->S, s->g, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->st));
  return MAG_OK;
}

// ------------------------------------------------------------ one-shots
static int open_session(const int* dims, const double* links, const double* spins,
                        mag_spin_session** s) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  CHK(mag_spin_session_create(dims, dev, s));
  int r = links ? mag_spin_session_upload_links(*s, links) : MAG_OK;
  if (r == MAG_OK && spins) r = mag_spin_session_set_spins(*s, spins);
  if (r != MAG_OK) {
    mag_spin_session_destroy(*s);
    *s = nullptr;
  }
  return r;
}

int mag_su2_generate_spins_uniform(const int* dims, uint64_t seed, double* spins) {
  if (!spins) return fail(MAG_ERR_ARG, "spins is NULL");
  mag_spin_session* s = nullptr;
  CHK(open_session(dims, nullptr, nullptr, &s));
  int r = mag_spin_session_generate_spins(s, seed);
  if (r == MAG_OK) r = mag_spin_session_get_spins(s, spins);
  mag_spin_session_destroy(s);
  return r;
}

int mag_su2_functional_spin(const int* dims, const double* links, const double* spins,
                            double* F) {
  if (!links || !spins) return fail(MAG_ERR_ARG, "NULL argument");
  mag_spin_session* s = nullptr;
  CHK(open_session(dims, links, spins, &s));
  int r = mag_spin_session_functional_spin(s, F);
  mag_spin_session_destroy(s);
  return r;
}

int mag_su2_gauge_transform_spins(const int* dims, double* links, const double* spins) {
  if (!links || !spins) return fail(MAG_ERR_ARG, "NULL argument");
  mag_spin_session* s = nullptr;
  CHK(open_session(dims, links, spins, &s));
  int r = mag_spin_session_gauge_transform(s);
  if (r == MAG_OK) r = mag_spin_session_download_links(s, links);
  mag_spin_session_destroy(s);
  if (r == MAG_OK) {
    int64_t V = 1;
    for (int i = 0; i < 4; i++) V *= dims[i];
    for (int64_t i = 0; i < V * 16; i++)
      if (!std::isfinite(links[i]))
        return fail(MAG_ERR_NONFINITE,
                    "a spin equal to (0,0,1) has no gauge matrix (matrix.h:960-963): links hold NaN");
  }
  return r;
}

int mag_su2_maximize_final(const int* dims, const double* links, double* spins, int OR_steps,
                           double tol_max, double tol_avg, int max_iter, int* iters) {
  if (!links || !spins) return fail(MAG_ERR_ARG, "NULL argument");
  mag_spin_session* s = nullptr;
  CHK(open_session(dims, links, spins, &s));
  int r = mag_spin_session_maximize_final(s, OR_steps, tol_max, tol_avg, max_iter, iters);
  if (r == MAG_OK || r == MAG_ERR_NOT_CONVERGED) {
    const int r2 = mag_spin_session_get_spins(s, spins);
    if (r2 != MAG_OK) r = r2;
  }
  mag_spin_session_destroy(s);
  return r;
}

int mag_su2_maximize_approximate(const int* dims, const double* links, double* spins,
                                 int OR_steps, int tolerance_digits, int max_iter, int* iters) {
  if (!links || !spins) return fail(MAG_ERR_ARG, "NULL argument");
  mag_spin_session* s = nullptr;
  CHK(open_session(dims, links, spins, &s));
  int r = mag_spin_session_maximize_approximate(s, OR_steps, tolerance_digits, max_iter, iters);
  if (r == MAG_OK || r == MAG_ERR_NOT_CONVERGED) {
    const int r2 = mag_spin_session_get_spins(s, spins);
    if (r2 != MAG_OK) r = r2;
  }
  mag_spin_session_destroy(s);
  return r;
}

int mag_su2_simulated_annealing(const int* dims, const double* links, double* spins,
                                double T_init, double T_final, double T_step, int OR_steps,
                                int thermalization_steps, uint64_t seed) {
  if (!links || !spins) return fail(MAG_ERR_ARG, "NULL argument");
  mag_spin_session* s = nullptr;
  CHK(open_session(dims, links, spins, &s));
  int r = mag_spin_session_simulated_annealing(s, T_init, T_final, T_step, OR_steps,
                                               thermalization_steps, seed);
  if (r == MAG_OK) r = mag_spin_session_get_spins(s, spins);
  mag_spin_session_destroy(s);
  return r;
}

}  // extern "C"
