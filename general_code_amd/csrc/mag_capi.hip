// mag_capi.hip -- C ABI of the MAG gauge-fixing path (include/mag_gpu.h).
//
// Host-side job management: device buffers, the per-sweep launch sequence
// (red half-sweep, black half-sweep, one-block reduction), convergence by
// device-resident flags (no host synchronisation per sweep), and the
// mixed-precision schedule (FP32 sweeps with an accumulated site gauge,
// FP64 rebuild from the FP64 master every reunit_every sweeps, FP64 rebuild
// and FP64 polish at the end; optionally after an FP64 warm-up), and the
// spin form with gauge copies (read-only links, one spin field per copy).
#include "mag_gpu.h"
#include "mag_capi_util.h"
#include "mag_internal.h"
#include "mag_job_impl.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using mag::Geom;
using mag::JobCtl;

namespace {
thread_local std::string g_err;
}  // namespace

int mag::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {
using mag::fail;

constexpr int kChunk = 20;  // sweeps enqueued between convergence polls

int check_dims(const int* d, int64_t* V) {
  if (!d) return fail(MAG_ERR_ARG, "dims is NULL");
  int64_t v = 1;
  for (int i = 0; i < 4; i++) {
    if (d[i] < 2 || (d[i] & 1))
      return fail(MAG_ERR_DIMS, "extent %d of dimension %d must be even and >= 2", d[i], i);
    v *= d[i];
  }
  // 8 * Vh * 12 reals must index in size_t and Vh in uint32
  if (v / 2 > (int64_t)0x7fffffff / 8) return fail(MAG_ERR_DIMS, "lattice too large");
  *V = v;
  return MAG_OK;
}

int check_params(const mag_params* p) {
  if (!p) return fail(MAG_ERR_ARG, "params is NULL");
  if (p->group != MAG_SU2 && p->group != MAG_SU3)
    return fail(MAG_ERR_ARG, "group must be 2 or 3, got %d", p->group);
  if (p->precision != MAG_PREC_F64 && p->precision != MAG_PREC_MIXED)
    return fail(MAG_ERR_ARG, "precision must be MAG_PREC_F64 or MAG_PREC_MIXED");
  if (!(p->omega > 0.0 && p->omega <= 2.0))
    return fail(MAG_ERR_ARG, "omega must be in (0, 2], got %g", p->omega);
  if (!(p->tol >= 0.0)) return fail(MAG_ERR_ARG, "tol must be >= 0");
  if (p->max_sweeps < 1) return fail(MAG_ERR_ARG, "max_sweeps must be >= 1");
  if (p->precision == MAG_PREC_MIXED) {
    if (p->reunit_every < 1) return fail(MAG_ERR_ARG, "reunit_every must be >= 1");
    if (!(p->tol_mixed >= 0.0)) return fail(MAG_ERR_ARG, "tol_mixed must be >= 0");
    if (p->fp64_warmup < 0 || p->fp64_warmup % kChunk != 0)
      return fail(MAG_ERR_ARG, "fp64_warmup must be a non-negative multiple of %d", kChunk);
  }
  if (p->track_gauge != 0 && p->track_gauge != 1) return fail(MAG_ERR_ARG, "track_gauge must be 0 or 1");
  if (p->track_gauge && (p->group != MAG_SU2 || p->form != MAG_FORM_LINK))
    return fail(MAG_ERR_ARG, "track_gauge is for the SU(2) link form");
  if (p->form == MAG_FORM_LINK) {
    if (p->copies > 1) return fail(MAG_ERR_ARG, "the link form holds one copy; use MAG_FORM_SPIN");
  } else if (p->form == MAG_FORM_SPIN) {
    if (p->copies < 1 || p->copies > MAG_MAX_COPIES)
      return fail(MAG_ERR_ARG, "copies must be in 1..%d, got %d", MAG_MAX_COPIES, p->copies);
  } else {
    return fail(MAG_ERR_ARG, "form must be MAG_FORM_LINK or MAG_FORM_SPIN");
  }
  int64_t V;
  return check_dims(p->dims, &V);
}

bool all_finite(const double* a, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(a[i])) return false;
  return true;
}

}  // namespace

namespace {

bool mixed32(const mag_job* j) { return j->prm.precision == MAG_PREC_MIXED && j->phase == 0; }

bool su3_site(const mag_job* j) { return j->spin && j->prm.group == MAG_SU3; }
bool su2_spin(const mag_job* j) { return j->spin && j->prm.group == MAG_SU2; }

// SU(2) spin form: the FP64 spins of copy c, refreshed from the FP32 spins
// while the job is in its FP32 phase (which sweeps only those)
double* spins64(mag_job* j, int c) {
  double* S = j->S + (size_t)c * j->sstride;
  if (su2_spin(j) && mixed32(j)) mag::launch_spins_to64(j->S32 + (size_t)c * j->sstride, S, j->g, j->st);
  return S;
}
int set_device(const mag_job* j) {
  HIPCHK(hipSetDevice(j->prm.device));
  return MAG_OK;
}

// the job's launch shape for its FP64 / FP32 link-form kernel (tuned on
// first use; the default shape if tuning fails -- every shape gives the
// same results)
mag::Shape link_shape(mag_job* j, bool fp32) {
  const int k = fp32 ? 1 : 0;
  if (!j->shape_ok[k]) {
    mag::Shape sh;
    if (mag::job_shape(j, fp32, &sh) != MAG_OK) sh = mag::Shape{};
    j->shape[k] = sh;
    j->shape_ok[k] = true;
  }
  return j->shape[k];
}

void launch_half(mag_job* j, int parity, const mag::HalfOpt& o = mag::HalfOpt{}) {
  if (j->spin && j->prm.group == MAG_SU3) {
    if (mixed32(j))  // FP32 phase: partials in sweep slot o.slot
      mag::launch_su3_site_copies<float>(j->W32, j->S32, j->sstride, j->ncopies, j->g, parity,
                                         j->prm.omega, j->part, j->pstride, j->ctl, j->st,
                                         o.slot);
    else
      mag::launch_su3_site_copies<double>(j->U64, j->S, j->sstride, j->ncopies, j->g, parity,
                                          j->prm.omega, j->part, j->pstride, j->ctl, j->st);
  } else if (j->spin) {
    if (mixed32(j))
      mag::launch_su2_copies_half<float>(j->W32, j->S32, j->sstride, j->ncopies, j->g, parity,
                                         j->prm.omega, j->part, j->pstride, j->ctl, j->st);
    else
      mag::launch_su2_copies_half<double>(j->U64, j->S, j->sstride, j->ncopies, j->g, parity,
                                          j->prm.omega, j->part, j->pstride, j->ctl, j->st);
  } else if (mixed32(j)) {
    // red sites accumulate the site gauge; the rebuilding black half reads it
    const mag::Shape sh = link_shape(j, true);
    float* G = (parity == 0 || o.mode == mag::HALF_REBUILD) ? j->G32 : nullptr;
    mag::HalfOpt oo = o;
    if (oo.mode == mag::HALF_REBUILD) {
      oo.U0 = j->U64;
      oo.part_rb = j->part + (size_t)mag::kSlots * j->pstride;
    }
    if (j->prm.group == MAG_SU2)
      mag::launch_su2_half<float>(j->W32, G, j->g, parity, j->prm.omega, j->part, j->ctl, sh, j->st, oo);
    else
      mag::launch_su3_half<float>(j->W32, G, j->g, parity, j->prm.omega, j->part, j->ctl, sh, j->st, oo);
  } else {
    const mag::Shape sh = link_shape(j, false);
    if (j->prm.group == MAG_SU2)  // G64: the tracked gauge (or nullptr)
      mag::launch_su2_half<double>(j->U64, j->G64, j->g, parity, j->prm.omega, j->part, j->ctl, sh, j->st);
    else
      mag::launch_su3_half<double>(j->U64, nullptr, j->g, parity, j->prm.omega, j->part, j->ctl, sh, j->st);
  }
}

// mixed FP32 phase: reduce the slots swept since the last reduction into
// the history and the sweep count; after a rebuilding sweep (SU(3) site
// form: after the F evaluation that ends a period) also test convergence
// on that F
void flush_slots(mag_job* j, bool rebuild) {
  if (j->pending == 0) return;
  if (j->spin) {  // SU(3) site form: every copy of the set
    const size_t hs = (size_t)j->prm.max_sweeps;
    mag::launch_reduce_slots_copies(j->part, j->pstride, 2 * j->nw, j->ncopies, j->pending,
                                    j->invF, j->invV, j->ctl, j->hF, j->hR, j->hA, hs, j->st);
    mag::launch_reduce_period_copies(j->part + (size_t)mag::kSlots * j->ncopies * j->pstride,
                                     j->pstride, j->nw, j->ncopies, j->pending, rebuild ? 1 : 0,
                                     j->invF, j->ctl, j->hF, hs, j->st);
  } else {
    // the black triples (the red ones folded in) of each slot
    mag::launch_reduce_slots(j->part + 3 * (size_t)j->nw, j->pstride, j->nw, j->pending,
                             j->invF, j->invV, j->ctl, j->hF, j->hR, j->hA, j->st);
    mag::launch_reduce_period(j->part + (size_t)mag::kSlots * j->pstride, j->nw, j->pending,
                              rebuild ? 1 : 0, j->invF, j->ctl, j->hF, j->st);
  }
  j->pending = 0;
}

// the FP64 links at the end of the FP32 phase (in place over the master)
void launch_final_rebuild(mag_job* j) {
  if (j->prm.group == MAG_SU2)
    mag::launch_su2_rebuild_final(j->U64, j->G32, j->W32, j->U64, j->g, j->st, j->G64);
  else
    mag::launch_su3_rebuild_final(j->U64, j->G32, j->W32, j->U64, j->g, j->st);
}

// One sweep.  Mixed link form, FP32 phase: periods of reunit_every sweeps;
// the first red half of a period renormalises the red gauges, the last
// black half rebuilds the links from the master; the sweeps' partials go to
// slots reduced at the end of a period or when the slots are full.
void enqueue_sweep(mag_job* j) {
  if (mixed32(j) && !j->spin) {
    const int R = j->prm.reunit_every;
    mag::HalfOpt o;
    o.slot = j->pending;
    o.mode = j->since_rebuild == 0 ? mag::HALF_RENORM : mag::HALF_PLAIN;
    launch_half(j, 0, o);
    const bool rebuild = j->since_rebuild == R - 1;
    o.mode = rebuild ? mag::HALF_REBUILD : mag::HALF_PLAIN;
    launch_half(j, 1, o);
    j->pending++;
    j->since_rebuild = rebuild ? 0 : j->since_rebuild + 1;
    if (rebuild || j->pending == mag::kSlots) flush_slots(j, rebuild);
    return;
  }
  if (su3_site(j) && mixed32(j)) {
    // SU(3) site form, FP32 phase: the sweeps' partials go to slots; every
    // reunit_every sweeps F is evaluated in FP64 from the FP32 state and the
    // period's reduction tests it (the FP32-summed F of the sweeps is too
    // noisy to stop on)
    mag::HalfOpt o;
    o.slot = j->pending;
    launch_half(j, 0, o);
    launch_half(j, 1, o);
    j->pending++;
    const bool check = ++j->since_rebuild == j->prm.reunit_every;
    if (check) {
      j->since_rebuild = 0;
      mag::launch_su3_site_F(j->W32, j->S32, j->sstride, j->ncopies, j->g,
                             j->part + (size_t)mag::kSlots * j->ncopies * j->pstride, j->pstride,
                             j->ctl, j->st);
    }
    if (check || j->pending == mag::kSlots) flush_slots(j, check);
    return;
  }
  launch_half(j, 0);
  launch_half(j, 1);
  if (j->spin) {
    mag::launch_reduce_copies(j->part, j->pstride, 2 * j->nw, j->ncopies, j->invF, j->invV,
                              j->ctl, j->hF, j->hR, j->hA, (size_t)j->prm.max_sweeps, j->st);
    return;
  }
  // link form: the black triples (the red ones folded in)
  mag::launch_reduce(j->part + 3 * (size_t)j->nw, j->nw, mag::RED_SWEEP, j->invF, j->invV, j->ctl, j->hF,
                     j->hR, j->hA, nullptr, j->st);
}

// functional partials of the current working links
void launch_functional(mag_job* j) {
  if (j->spin && j->prm.group == MAG_SU2) {
    mag::launch_spin_functional(j->U64, spins64(j, 0), j->g, j->part, nullptr, j->st);
  } else if (mixed32(j) && !j->spin) {
    if (j->prm.group == MAG_SU2) mag::launch_su2_functional<float>(j->W32, j->g, j->part, j->st);
    else mag::launch_su3_functional<float>(j->W32, j->g, j->part, j->st);
  } else {
    if (j->prm.group == MAG_SU2) mag::launch_su2_functional<double>(j->U64, j->g, j->part, j->st);
    else mag::launch_su3_functional<double>(j->U64, j->g, j->part, j->st);
  }
}

int read_ctl(mag_job* j) {
  HIPCHK(hipMemcpyAsync(j->ctl_h, j->ctl, j->ncopies * sizeof(JobCtl), hipMemcpyDeviceToHost,
                        j->st));
  HIPCHK(hipStreamSynchronize(j->st));
  return MAG_OK;
}

int write_ctl(mag_job* j) {
  HIPCHK(hipMemcpyAsync(j->ctl, j->ctl_h, j->ncopies * sizeof(JobCtl), hipMemcpyHostToDevice,
                        j->st));
  HIPCHK(hipStreamSynchronize(j->st));
  return MAG_OK;
}

bool all_done(const mag_job* j) {
  for (int c = 0; c < j->ncopies; c++)
    if (!j->ctl_h[c].done) return false;
  return true;
}
bool all_converged(const mag_job* j) {
  for (int c = 0; c < j->ncopies; c++)
    if (!j->ctl_h[c].converged) return false;
  return true;
}

// FP64 spins (or SU(3) site gauges) of every copy -> their FP32 working copies
void sites_to32(mag_job* j) {
  for (int c = 0; c < j->ncopies; c++) {
    if (su3_site(j))
      mag::launch_su3_sites_to32(j->S + (size_t)c * j->sstride, j->S32 + (size_t)c * j->sstride,
                                 j->g, j->st);
    else if (su2_spin(j))
      mag::launch_spins_to32(j->S + (size_t)c * j->sstride, j->S32 + (size_t)c * j->sstride, j->g,
                             j->st);
  }
}

// SU(3) site form: the FP64 links of copy c, G_c U G_c^+, into tmp (FP64
// chunked links); in the FP32 phase the FP64 gauges are refreshed first
void su3_copy_links(mag_job* j, int c, double* tmp) {
  double* Gc = j->S + (size_t)c * j->sstride;
  if (j->prm.precision == MAG_PREC_MIXED && j->phase == 0)
    mag::launch_su3_sites_to64(j->S32 + (size_t)c * j->sstride, Gc, j->g, j->st);
  mag::launch_su3_site_links(j->U64, Gc, tmp, j->g, j->st);
}

// FP64 link buffer for the SU(3) site form (nullptr otherwise)
int alloc_tmp(mag_job* j, double** tmp) {
  *tmp = nullptr;
  if (!su3_site(j)) return MAG_OK;
  HIPCHK(hipMalloc((void**)tmp, j->nreal * sizeof(double)));
  return MAG_OK;
}

// initial control block and functional F0 of copy c (site form) or of the
// job (link form); links (and site variables) already in place; tmp: an
// FP64 link buffer for the SU(3) site form
void init_copy(mag_job* j, int c, double* tmp) {
  JobCtl& h = j->ctl_h[c];
  std::memset(&h, 0, sizeof(JobCtl));
  h.max_sweeps = j->prm.max_sweeps;
  const bool fp32 = j->prm.precision == MAG_PREC_MIXED && j->phase == 0;
  h.tol = fp32 ? j->prm.tol_mixed : j->prm.tol;
  // no per-sweep test in the FP32 phase of the link form and of the SU(3)
  // site form (the test runs on FP64 evaluations every reunit_every sweeps)
  h.window = (fp32 && (!j->spin || su3_site(j))) ? -1 : 0;
  h.rb_every = j->prm.reunit_every;
  hipMemcpyAsync(j->ctl + c, &h, sizeof(JobCtl), hipMemcpyHostToDevice, j->st);
  double* part = j->part + (size_t)c * j->pstride;
  if (su3_site(j)) {
    su3_copy_links(j, c, tmp);
    mag::launch_su3_functional<double>(tmp, j->g, part, j->st);
  } else if (j->spin) {
    mag::launch_spin_functional(j->U64, j->S + (size_t)c * j->sstride, j->g, part, nullptr, j->st);
  } else if (j->prm.group == MAG_SU2) mag::launch_su2_functional<double>(j->U64, j->g, part, j->st);
  else mag::launch_su3_functional<double>(j->U64, j->g, part, j->st);
  const size_t hs = (size_t)c * j->prm.max_sweeps;
  mag::launch_reduce(part, 2 * j->nf, mag::RED_INIT, j->invF, j->invV, j->ctl + c, j->hF + hs,
                     j->hR + hs, j->hA + hs, nullptr, j->st);
}

// mixed with FP64 warm-up: FP32 working links from the FP64 links, identity
// gauge, FP32 tolerance (ctl_h current)
int start_fp32_phase(mag_job* j) {
  mag::launch_links_f64_to_f32(j->U64, j->W32, j->g, j->prm.group, j->st);
  if (!j->spin) mag::launch_gauge_identity(j->G32, j->g, j->prm.group, j->st);
  sites_to32(j);
  j->phase = 0;
  j->since_rebuild = 0;
  j->pending = 0;
  for (int c = 0; c < j->ncopies; c++) {
    j->ctl_h[c].tol = j->prm.tol_mixed;
    j->ctl_h[c].window = (j->spin && !su3_site(j)) ? 0 : -1;
    j->ctl_h[c].rb_every = j->prm.reunit_every;
    j->ctl_h[c].F_ref = 0.0;
  }
  return write_ctl(j);
}

int max_sweep(const mag_job* j) {
  int m = 0;
  for (int c = 0; c < j->ncopies; c++) m = j->ctl_h[c].sweep > m ? j->ctl_h[c].sweep : m;
  return m;
}

int maybe_start_fp32(mag_job* j, bool* switched) {
  *switched = false;
  if (j->prm.precision != MAG_PREC_MIXED || j->phase != -1 || all_done(j)) return MAG_OK;
  if (max_sweep(j) < j->prm.fp64_warmup) return MAG_OK;
  *switched = true;
  return start_fp32_phase(j);
}

// after the configuration is in U64 (SoA FP64): reset the job state
int reset_after_upload(mag_job* j) {
  const bool mixed = j->prm.precision == MAG_PREC_MIXED;
  j->phase = (mixed && j->prm.fp64_warmup > 0) ? -1 : 0;
  if (mixed && j->phase == 0) {
    mag::launch_links_f64_to_f32(j->U64, j->W32, j->g, j->prm.group, j->st);
    if (!j->spin) mag::launch_gauge_identity(j->G32, j->g, j->prm.group, j->st);
    sites_to32(j);
  }
  j->fp64_start = mixed ? -1 : 0;
  j->since_rebuild = 0;
  j->pending = 0;
  if (j->G64) mag::launch_gauge_identity64(j->G64, j->g, j->st);
  // initial functional from the FP64 links (and each copy's site variable)
  double* tmp = nullptr;
  CHK(alloc_tmp(j, &tmp));
  for (int c = 0; c < j->ncopies; c++) init_copy(j, c, tmp);
  if (tmp) {
    (void)hipStreamSynchronize(j->st);
    (void)hipFree(tmp);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(j->st));
  j->loaded = true;
  return MAG_OK;
}

// copy_seed: spin form, start the copies from the gauge copies of the
// synthetic configuration of that seed (nullptr: every copy at e3)
int upload_device_aos(mag_job* j, const double* dev, const uint64_t* copy_seed = nullptr) {
  if (j->prm.group == MAG_SU2) mag::launch_su2_aos_to_soa<double>(dev, j->U64, j->g, j->st);
  else mag::launch_su3_aos_to_soa<double>(dev, j->U64, j->g, j->st);
  if (su3_site(j)) {
    if (copy_seed) {
      mag::launch_gen_su3_copy_gauges(j->S, j->sstride, j->ncopies, j->g, *copy_seed, j->st);
    } else {
      for (int c = 0; c < j->ncopies; c++)
        mag::launch_su3_sites_identity(j->S + (size_t)c * j->sstride, j->g, j->st);
    }
  } else if (j->spin) {
    if (copy_seed) mag::launch_gen_copy_spins(j->S, j->sstride, j->ncopies, j->g, *copy_seed, j->st);
    else mag::launch_spins_e3(j->S, j->sstride * j->ncopies, j->g, j->st);
  }
  return reset_after_upload(j);
}

void free_job(mag_job* j) {
  if (!j) return;
  hipFree(j->U64);
  hipFree(j->S);
  hipFree(j->S32);
  hipFree(j->W32);
  hipFree(j->G32);
  hipFree(j->G64);
  hipFree(j->part);
  hipFree(j->ctl);
  hipFree(j->hF);
  hipFree(j->hR);
  hipFree(j->hA);
  hipFree(j->out3);
  if (j->ctl_h) hipHostFree(j->ctl_h);
  if (j->out3_h) hipHostFree(j->out3_h);
  if (j->st) hipStreamDestroy(j->st);
  delete j;
}

// finish the FP32 phase: links <- G U0 G^+ in FP64, F_prev <- F(links)
// spin form: the FP32 links were only read; FP64 sweeps go on from the
// spins with the FP64 links
int finalize_mixed_spin(mag_job* j) {
  flush_slots(j, false);
  CHK(read_ctl(j));
  double* tmp = nullptr;
  CHK(alloc_tmp(j, &tmp));
  for (int c = 0; c < j->ncopies; c++) {
    double* part = j->part + (size_t)c * j->pstride;
    const size_t hs = (size_t)c * j->prm.max_sweeps;
    if (su3_site(j)) {
      su3_copy_links(j, c, tmp);  // FP32 gauges -> FP64 (Gram-Schmidt), then the links
      mag::launch_su3_functional<double>(tmp, j->g, part, j->st);
    } else {
      mag::launch_spin_functional(j->U64, spins64(j, c), j->g, part, nullptr, j->st);
    }
    mag::launch_reduce(part, 2 * j->nf, mag::RED_RESET, j->invF, j->invV, j->ctl + c,
                       j->hF + hs, j->hR + hs, j->hA + hs, nullptr, j->st);
  }
  if (tmp) {
    (void)hipStreamSynchronize(j->st);
    (void)hipFree(tmp);
  }
  j->phase = 1;
  j->fp64_start = max_sweep(j);
  HIPCHK(hipGetLastError());
  CHK(read_ctl(j));
  for (int c = 0; c < j->ncopies; c++) {
    JobCtl& h = j->ctl_h[c];
    h.tol = j->prm.tol;
    h.window = 0;
    h.converged = 0;
    h.done = (h.sweep >= h.max_sweeps) ? 1 : 0;
  }
  return write_ctl(j);
}

int finalize_mixed(mag_job* j) {
  if (!mixed32(j)) return MAG_OK;
  if (j->spin) return finalize_mixed_spin(j);
  flush_slots(j, false);
  CHK(read_ctl(j));
  launch_final_rebuild(j);
  j->phase = 1;
  j->fp64_start = j->ctl_h->sweep;
  if (j->prm.group == MAG_SU2) mag::launch_su2_functional<double>(j->U64, j->g, j->part, j->st);
  else mag::launch_su3_functional<double>(j->U64, j->g, j->part, j->st);
  mag::launch_reduce(j->part, 2 * j->nf, mag::RED_RESET, j->invF, j->invV, j->ctl, j->hF, j->hR,
                     j->hA, nullptr, j->st);
  HIPCHK(hipGetLastError());
  CHK(read_ctl(j));
  j->ctl_h->tol = j->prm.tol;
  j->ctl_h->window = 0;
  j->ctl_h->converged = 0;
  j->ctl_h->done = (j->ctl_h->sweep >= j->ctl_h->max_sweeps) ? 1 : 0;
  return write_ctl(j);
}

int run_until_done(mag_job* j) {
  for (;;) {
    CHK(read_ctl(j));
    if (all_done(j)) return MAG_OK;
    bool sw;
    CHK(maybe_start_fp32(j, &sw));
    for (int i = 0; i < kChunk; i++) enqueue_sweep(j);
    flush_slots(j, false);
    HIPCHK(hipGetLastError());
  }
}

}  // namespace

void mag::job_copy_links64(mag_job* j, int c, double* soa) {
  if (su3_site(j)) su3_copy_links(j, c, soa);
  else mag::launch_spin_links(j->U64, spins64(j, c), soa, j->g, j->st);
}
int mag::job_read_ctl(mag_job* j) { return read_ctl(j); }
int mag::job_finalize_mixed(mag_job* j) { return finalize_mixed(j); }
bool mag::job_mixed32(const mag_job* j) { return mixed32(j); }
bool mag::job_done(const mag_job* j) { return all_done(j); }
int mag::job_maybe_start_fp32(mag_job* j, bool* switched) { return maybe_start_fp32(j, switched); }

// ====================================================================== API
extern "C" {

void mag_params_default(mag_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->group = MAG_SU2;
  p->dims[0] = p->dims[1] = p->dims[2] = p->dims[3] = 8;
  p->precision = MAG_PREC_F64;
  p->device = 0;
  p->omega = 1.7;
  p->tol = 1e-12;
  p->max_sweeps = 1000;
  p->reunit_every = 10;
  p->tol_mixed = 1e-9;
  p->form = MAG_FORM_LINK;
  p->copies = 1;
  p->fp64_warmup = 0;
}

const char* mag_last_error(void) { return g_err.c_str(); }

int mag_device_count(int* n) {
  if (!n) return fail(MAG_ERR_ARG, "n is NULL");
  HIPCHK(hipGetDeviceCount(n));
  return MAG_OK;
}

int mag_job_footprint(const mag_params* p, uint64_t* bytes) {
  CHK(check_params(p));
  if (!bytes) return fail(MAG_ERR_ARG, "bytes is NULL");
  int64_t V = 1;
  for (int i = 0; i < 4; i++) V *= p->dims[i];
  const uint64_t nr = p->group == MAG_SU2 ? 4 : 12;
  const uint64_t nc = p->form == MAG_FORM_SPIN ? (uint64_t)p->copies : 1;
  uint64_t b = (uint64_t)V * 4 * nr * 8;  // FP64 links / master
  if (p->precision == MAG_PREC_MIXED) b += (uint64_t)V * 4 * nr * 4 + (uint64_t)V * nr * 4;
  if (p->form == MAG_FORM_SPIN) {
    const uint64_t sr = p->group == MAG_SU2 ? 3 : 12;  // reals per site per copy
    b += nc * (uint64_t)V * sr * 8;
    if (p->precision == MAG_PREC_MIXED) {
      b -= (uint64_t)V * nr * 4;  // no accumulated link-form gauge
      b += nc * (uint64_t)V * sr * 4;  // FP32 spins / site gauges
    }
    if (p->group == MAG_SU3) b += (uint64_t)V * 4 * nr * 8;  // transient FP64 links
  }
  if (p->track_gauge) b += (uint64_t)V * 4 * 8;  // FP64 site gauges
  const uint64_t lanes = (uint64_t)(V / 2) * (p->group == MAG_SU2 ? 1 : 2);
  const int nw = (int)((lanes + 63) / 64);  // partial triples per half-sweep (at most)
  b += (uint64_t)4 << 20;  // pitch padding of the chunk arrays (4 KB per array)
  // mixed link form and SU(3) site form: kSlots sweep slots and one more
  const uint64_t slots =
      (p->precision == MAG_PREC_MIXED && (p->form == MAG_FORM_LINK || p->group == MAG_SU3))
          ? mag::kSlots + 1
          : 1;
  b += nc * (mag::part_stride(nw) * 8 * slots + (uint64_t)p->max_sweeps * 3 * 8 + sizeof(JobCtl)) + 3 * 8;
  *bytes = b;
  return MAG_OK;
}

int mag_job_create(const mag_params* p, mag_job** out) {
  if (!out) return fail(MAG_ERR_ARG, "out is NULL");
  *out = nullptr;
  CHK(check_params(p));
  mag_job* j = new (std::nothrow) mag_job();
  if (!j) return fail(MAG_ERR_NOMEM, "host allocation failed");
  j->prm = *p;
  int r = set_device(j);
  if (r != MAG_OK) {
    delete j;
    return r;
  }
  check_dims(p->dims, &j->V);
  j->g = mag::make_geom(p->dims);
  j->nr = p->group == MAG_SU2 ? 4 : 12;
  j->nr_in = p->group == MAG_SU2 ? 4 : 18;
  j->nreal = (size_t)8 * j->g.hp * j->nr;  // chunk arrays at the padded pitch
  j->nb = mag::sweep_blocks(j->g, p->group);
  j->nw = mag::sweep_waves(j->g, p->group);
  if (p->form == MAG_FORM_SPIN && p->group == MAG_SU3) {  // site form: lane pairs
    j->nb = (int)((mag::site3_lanes(j->g) + mag::kBlock - 1) / mag::kBlock);
    j->nw = (int)((mag::site3_lanes(j->g) + 63) / 64);
  }
  j->nf = mag::nblocks(j->g);
  j->invF = 1.0 / (double)(j->V * (p->group == MAG_SU2 ? 4 : 12));
  j->invV = 1.0 / (double)j->V;
  j->spin = p->form == MAG_FORM_SPIN;
  j->ncopies = j->spin ? p->copies : 1;
  // SU(2) spins: plain SoA at pitch Vh; SU(3) site gauges: chunk arrays
  j->sstride = p->group == MAG_SU2 ? (size_t)2 * j->g.Vh * 3 : (size_t)2 * j->g.hp * 12;
  j->pstride = mag::part_stride(j->nw);
#define ALLOC(ptr, bytes)                                                          \
  do {                                                                             \
    hipError_t e_ = hipMalloc((void**)&(ptr), (bytes));                            \
    if (e_ != hipSuccess) {                                                        \
      free_job(j);                                                                 \
      return fail(MAG_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", (size_t)(bytes), \
                  hipGetErrorString(e_));                                          \
    }                                                                              \
  } while (0)
  hipError_t e = hipStreamCreateWithFlags(&j->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    free_job(j);
    return fail(MAG_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
  }
  ALLOC(j->U64, j->nreal * sizeof(double));
  if (p->precision == MAG_PREC_MIXED) {
    ALLOC(j->W32, j->nreal * sizeof(float));
    if (!j->spin) ALLOC(j->G32, (size_t)2 * j->g.hp * j->nr * sizeof(float));
    if (j->spin) ALLOC(j->S32, j->sstride * j->ncopies * sizeof(float));
  }
  if (j->spin) ALLOC(j->S, j->sstride * j->ncopies * sizeof(double));
  if (p->track_gauge) ALLOC(j->G64, (size_t)2 * j->g.hp * 4 * sizeof(double));
  // mixed link form: kSlots sweep slots and the rebuild partials; SU(3)
  // site form: kSlots sweep slots and the partials of the FP64 F check
  const size_t nslots =
      (p->precision == MAG_PREC_MIXED && (!j->spin || p->group == MAG_SU3)) ? mag::kSlots + 1 : 1;
  ALLOC(j->part, j->pstride * j->ncopies * nslots * sizeof(double));
  ALLOC(j->ctl, j->ncopies * sizeof(JobCtl));
  ALLOC(j->hF, (size_t)p->max_sweeps * j->ncopies * sizeof(double));
  ALLOC(j->hR, (size_t)p->max_sweeps * j->ncopies * sizeof(double));
  ALLOC(j->hA, (size_t)p->max_sweeps * j->ncopies * sizeof(double));
  ALLOC(j->out3, 3 * sizeof(double));
#undef ALLOC
  if (hipHostMalloc((void**)&j->ctl_h, j->ncopies * sizeof(JobCtl)) != hipSuccess ||
      hipHostMalloc((void**)&j->out3_h, 3 * sizeof(double)) != hipSuccess) {
    free_job(j);
    return fail(MAG_ERR_NOMEM, "pinned host allocation failed");
  }
  std::memset(j->ctl_h, 0, j->ncopies * sizeof(JobCtl));
  for (int c = 0; c < j->ncopies; c++) j->ctl_h[c].done = 1;  // nothing to sweep before an upload
  e = hipMemcpy(j->ctl, j->ctl_h, j->ncopies * sizeof(JobCtl), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_job(j);
    return fail(MAG_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(e));
  }
  *out = j;
  return MAG_OK;
}

int mag_job_destroy(mag_job* job) {
  if (!job) return MAG_OK;
  if (job->batch) mag_batch_remove(job->batch, job);
  hipSetDevice(job->prm.device);
  hipStreamSynchronize(job->st);
  free_job(job);
  return MAG_OK;
}

int mag_job_stream(mag_job* job, void** stream) {
  if (!job || !stream) return fail(MAG_ERR_ARG, "NULL argument");
  *stream = (void*)job->st;
  return MAG_OK;
}

int mag_job_upload_links(mag_job* j, const double* links) {
  if (!j || !links) return fail(MAG_ERR_ARG, "NULL argument");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  CHK(set_device(j));
  const size_t n = (size_t)j->V * 4 * j->nr_in;
  if (!all_finite(links, n)) return fail(MAG_ERR_NONFINITE, "links hold NaN or Inf");
  double* stage = nullptr;
  HIPCHK(hipMalloc((void**)&stage, n * sizeof(double)));
  hipError_t e = hipMemcpyAsync(stage, links, n * sizeof(double), hipMemcpyHostToDevice, j->st);
  int r = (e == hipSuccess) ? upload_device_aos(j, stage)
                            : fail(MAG_ERR_HIP, "hipMemcpy H2D failed: %s", hipGetErrorString(e));
  hipStreamSynchronize(j->st);
  hipFree(stage);
  return r;
}

int mag_job_upload_links_device(mag_job* j, const double* dev_links) {
  if (!j || !dev_links) return fail(MAG_ERR_ARG, "NULL argument");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  CHK(set_device(j));
  return upload_device_aos(j, dev_links);
}

int mag_job_generate(mag_job* j, double eps, uint64_t seed) {
  return mag_job_generate_copy(j, eps, seed, 0);
}

int mag_job_generate_copy(mag_job* j, double eps, uint64_t seed, uint32_t copy) {
  if (!j) return fail(MAG_ERR_ARG, "NULL job");
  if (j->spin && copy != 0)
    return fail(MAG_ERR_ARG, "a spin-form job holds copies 0..copies-1: generate copy 0");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  if (!std::isfinite(eps)) return fail(MAG_ERR_ARG, "eps must be finite");
  CHK(set_device(j));
  const size_t n = (size_t)j->V * 4 * j->nr_in;
  double* stage = nullptr;
  double* gs = nullptr;
  HIPCHK(hipMalloc((void**)&stage, n * sizeof(double)));
  if (j->prm.group == MAG_SU3) {
    hipError_t e = hipMalloc((void**)&gs, (size_t)j->V * 18 * sizeof(double));
    if (e != hipSuccess) {
      hipFree(stage);
      return fail(MAG_ERR_NOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
    }
    mag::launch_gen_su3(stage, gs, j->g, eps, seed, mag::gauge_copy_seed(seed, copy), j->st);
  } else {
    mag::launch_gen_su2(stage, j->g, eps, seed, mag::gauge_copy_seed(seed, copy), j->st);
  }
  int r = upload_device_aos(j, stage, j->spin ? &seed : nullptr);
  hipStreamSynchronize(j->st);
  hipFree(stage);
  hipFree(gs);
  return r;
}

int mag_job_sweeps(mag_job* j, int n) {
  if (!j || n < 0) return fail(MAG_ERR_ARG, "bad argument");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  CHK(set_device(j));
  for (int i = 0; i < n; i++) enqueue_sweep(j);
  flush_slots(j, false);
  HIPCHK(hipGetLastError());
  return MAG_OK;
}

int mag_job_finalize(mag_job* j) {
  if (!j) return fail(MAG_ERR_ARG, "NULL job");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  CHK(set_device(j));
  return finalize_mixed(j);
}

int mag_job_run(mag_job* j) {
  if (!j) return fail(MAG_ERR_ARG, "NULL job");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  CHK(set_device(j));
  CHK(run_until_done(j));
  if (mixed32(j)) {
    CHK(finalize_mixed(j));
    CHK(run_until_done(j));
  }
  CHK(read_ctl(j));
  if (!all_converged(j))
    return fail(MAG_ERR_NOT_CONVERGED, "sweep cap %d reached before tol", j->prm.max_sweeps);
  return MAG_OK;
}

int mag_job_status(mag_job* j, int* sweeps, int* converged, double* F) {
  if (!j) return fail(MAG_ERR_ARG, "NULL job");
  CHK(set_device(j));
  CHK(read_ctl(j));
  if (sweeps) *sweeps = j->ctl_h->sweep;
  if (converged) *converged = all_converged(j) ? 1 : 0;
  if (F) *F = j->ctl_h->F_prev;
  return MAG_OK;
}

int mag_job_initial_functional(mag_job* j, double* F0) {
  if (!j || !F0) return fail(MAG_ERR_ARG, "NULL argument");
  CHK(set_device(j));
  CHK(read_ctl(j));
  *F0 = j->ctl_h->F0;
  return MAG_OK;
}

int mag_job_history(mag_job* j, double* F, double* rmax, double* ravg, int cap, int* n) {
  if (!j || cap < 0) return fail(MAG_ERR_ARG, "bad argument");
  CHK(set_device(j));
  CHK(read_ctl(j));
  int m = j->ctl_h->sweep < cap ? j->ctl_h->sweep : cap;
  if (m > 0) {
    if (F) HIPCHK(hipMemcpy(F, j->hF, m * sizeof(double), hipMemcpyDeviceToHost));
    if (rmax) HIPCHK(hipMemcpy(rmax, j->hR, m * sizeof(double), hipMemcpyDeviceToHost));
    if (ravg) HIPCHK(hipMemcpy(ravg, j->hA, m * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (n) *n = m;
  return MAG_OK;
}

int mag_job_fp64_start(mag_job* j, int* sweep) {
  if (!j || !sweep) return fail(MAG_ERR_ARG, "NULL argument");
  *sweep = j->fp64_start;
  return MAG_OK;
}

int mag_job_download_links(mag_job* j, double* links) {
  if (!j || !links) return fail(MAG_ERR_ARG, "NULL argument");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (j->spin) return mag_job_download_copy_links(j, 0, links);
  CHK(set_device(j));
  const size_t n = (size_t)j->V * 4 * j->nr_in;
  double* stage = nullptr;
  HIPCHK(hipMalloc((void**)&stage, n * sizeof(double)));
  if (mixed32(j)) {
    if (j->prm.group == MAG_SU2) mag::launch_su2_soa_to_aos<float>(j->W32, stage, j->g, j->st);
    else mag::launch_su3_soa_to_aos<float>(j->W32, stage, j->g, j->st);
  } else {
    if (j->prm.group == MAG_SU2) mag::launch_su2_soa_to_aos<double>(j->U64, stage, j->g, j->st);
    else mag::launch_su3_soa_to_aos<double>(j->U64, stage, j->g, j->st);
  }
  hipError_t e = hipMemcpyAsync(links, stage, n * sizeof(double), hipMemcpyDeviceToHost, j->st);
  hipError_t e2 = hipStreamSynchronize(j->st);
  hipFree(stage);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(MAG_ERR_HIP, "download failed: %s",
                hipGetErrorString(e != hipSuccess ? e : e2));
  return MAG_OK;
}

static int check_canonical(mag_job* j) {
  if (!j) return fail(MAG_ERR_ARG, "NULL job");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (!j->G64) return fail(MAG_ERR_STATE, "the job does not track its gauge (track_gauge = 0)");
  if (mixed32(j)) return fail(MAG_ERR_STATE, "the job is in its FP32 phase: run or finalize it first");
  return MAG_OK;
}

int mag_job_download_spins(mag_job* j, double* spins) {
  CHK(check_canonical(j));
  if (!spins) return fail(MAG_ERR_ARG, "NULL spins");
  CHK(set_device(j));
  const size_t n = (size_t)j->V * 3;
  double *S = nullptr, *stage = nullptr;
  HIPCHK(hipMalloc((void**)&S, n * sizeof(double)));
  hipError_t e = hipMalloc((void**)&stage, n * sizeof(double));
  if (e == hipSuccess) {
    mag::launch_gauge_spins(j->G64, S, j->g, j->st);
    mag::launch_spins_layout(stage, S, j->g, 0, j->st);
    e = hipMemcpyAsync(spins, stage, n * sizeof(double), hipMemcpyDeviceToHost, j->st);
  }
  hipError_t e2 = hipStreamSynchronize(j->st);
  hipFree(S);
  hipFree(stage);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(MAG_ERR_HIP, "download_spins failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  return MAG_OK;
}

int mag_job_download_links_canonical(mag_job* j, double* links) {
  CHK(check_canonical(j));
  if (!links) return fail(MAG_ERR_ARG, "NULL links");
  CHK(set_device(j));
  const size_t n = (size_t)j->V * 16;
  double *S = nullptr, *soa = nullptr, *stage = nullptr;
  HIPCHK(hipMalloc((void**)&S, (size_t)j->V * 3 * sizeof(double)));
  hipError_t e = hipMalloc((void**)&soa, j->nreal * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&stage, n * sizeof(double));
  if (e == hipSuccess) {
    mag::launch_gauge_spins(j->G64, S, j->g, j->st);
    mag::launch_canonical_links(j->U64, S, j->G64, soa, j->g, j->st);
    mag::launch_su2_soa_to_aos<double>(soa, stage, j->g, j->st);
    e = hipMemcpyAsync(links, stage, n * sizeof(double), hipMemcpyDeviceToHost, j->st);
  }
  hipError_t e2 = hipStreamSynchronize(j->st);
  hipFree(S);
  hipFree(soa);
  hipFree(stage);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(MAG_ERR_HIP, "download_links_canonical failed: %s",
                hipGetErrorString(e != hipSuccess ? e : e2));
  if (!all_finite(links, n))
    return fail(MAG_ERR_NONFINITE, "a site's spin is (0,0,1), the pole of GetGaugeMatrix");
  return MAG_OK;
}

int mag_job_functional(mag_job* j, double* F) {
  if (!j || !F) return fail(MAG_ERR_ARG, "NULL argument");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  CHK(set_device(j));
  double* tmp = nullptr;
  CHK(alloc_tmp(j, &tmp));
  if (tmp) {
    su3_copy_links(j, 0, tmp);
    mag::launch_su3_functional<double>(tmp, j->g, j->part, j->st);
  } else {
    launch_functional(j);
  }
  mag::launch_reduce(j->part, 2 * j->nf, mag::RED_PLAIN, j->invF, j->invV, j->ctl, j->hF, j->hR,
                     j->hA, j->out3, j->st);
  if (tmp) {
    (void)hipStreamSynchronize(j->st);
    (void)hipFree(tmp);
  }
  HIPCHK(hipMemcpyAsync(j->out3_h, j->out3, 3 * sizeof(double), hipMemcpyDeviceToHost, j->st));
  HIPCHK(hipStreamSynchronize(j->st));
  *F = j->out3_h[0];
  return MAG_OK;
}

int mag_job_plaquette(mag_job* j, double* P) {
  if (!j || !P) return fail(MAG_ERR_ARG, "NULL argument");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  CHK(set_device(j));
  // U64 is either the current links or (mixed, FP32 phase) the master,
  // which is gauge-equivalent: the plaquette is the same.
  if (j->prm.group == MAG_SU2) mag::launch_su2_plaquette(j->U64, j->g, j->part, j->st);
  else mag::launch_su3_plaquette(j->U64, j->g, j->part, j->st);
  mag::launch_reduce(j->part, 2 * j->nf, mag::RED_PLAIN, j->invV, j->invV, j->ctl, j->hF, j->hR,
                     j->hA, j->out3, j->st);
  HIPCHK(hipMemcpyAsync(j->out3_h, j->out3, 3 * sizeof(double), hipMemcpyDeviceToHost, j->st));
  HIPCHK(hipStreamSynchronize(j->st));
  *P = j->out3_h[0];
  return MAG_OK;
}

int mag_job_sweeps_timed(mag_job* j, int n, double* total_ms) {
  if (!j || n < 1 || !total_ms) return fail(MAG_ERR_ARG, "bad argument");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  CHK(set_device(j));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, j->st));
  for (int i = 0; i < n; i++) enqueue_sweep(j);
  flush_slots(j, false);
  HIPCHK(hipEventRecord(e1, j->st));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *total_ms = ms;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return MAG_OK;
}

int mag_job_halfsweep_timed(mag_job* j, int n, double* avg_ms) {
  if (!j || n < 1 || !avg_ms) return fail(MAG_ERR_ARG, "bad argument");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  CHK(set_device(j));
  CHK(read_ctl(j));
  if (all_done(j)) return fail(MAG_ERR_STATE, "job is done: its kernels would return at once");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, j->st));
  for (int i = 0; i < n; i++) launch_half(j, i & 1);
  HIPCHK(hipEventRecord(e1, j->st));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = ms / n;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return MAG_OK;
}

int mag_job_shape(mag_job* j, int fp32, int* nt, int* spt) {
  if (!j || !nt || !spt) return fail(MAG_ERR_ARG, "NULL argument");
  const int k = fp32 ? 1 : 0;
  *nt = j->shape_ok[k] ? j->shape[k].nt : 0;
  *spt = j->shape_ok[k] ? j->shape[k].spt : 0;
  return MAG_OK;
}

// ------------------------------------------------------------ gauge copies
static int check_copy(mag_job* j, int c) {
  if (!j) return fail(MAG_ERR_ARG, "NULL job");
  if (c < 0 || c >= j->ncopies) return fail(MAG_ERR_ARG, "copy %d out of range 0..%d", c, j->ncopies - 1);
  return MAG_OK;
}

int mag_job_set_copy_spins(mag_job* j, int c, const double* spins) {
  CHK(check_copy(j, c));
  if (!spins) return fail(MAG_ERR_ARG, "NULL spins");
  if (!j->spin || j->prm.group != MAG_SU2) return fail(MAG_ERR_STATE, "not an SU(2) spin-form job");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (j->batch) return fail(MAG_ERR_STATE, "job is in a batch: sweep it through the batch");
  const size_t n = (size_t)j->V * 3;
  if (!all_finite(spins, n)) return fail(MAG_ERR_NONFINITE, "spins hold NaN or Inf");
  CHK(set_device(j));
  double* stage = nullptr;
  HIPCHK(hipMalloc((void**)&stage, n * sizeof(double)));
  hipError_t e = hipMemcpyAsync(stage, spins, n * sizeof(double), hipMemcpyHostToDevice, j->st);
  if (e == hipSuccess) {
    mag::launch_spins_layout(stage, j->S + (size_t)c * j->sstride, j->g, 1, j->st);
    if (mixed32(j))
      mag::launch_spins_to32(j->S + (size_t)c * j->sstride, j->S32 + (size_t)c * j->sstride, j->g,
                             j->st);
    init_copy(j, c, nullptr);
    e = hipStreamSynchronize(j->st);
  }
  hipFree(stage);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "set_copy_spins failed: %s", hipGetErrorString(e));
  return MAG_OK;
}

int mag_job_get_copy_spins(mag_job* j, int c, double* spins) {
  CHK(check_copy(j, c));
  if (!spins) return fail(MAG_ERR_ARG, "NULL spins");
  if (!j->spin || j->prm.group != MAG_SU2) return fail(MAG_ERR_STATE, "not an SU(2) spin-form job");
  CHK(set_device(j));
  const size_t n = (size_t)j->V * 3;
  double* stage = nullptr;
  HIPCHK(hipMalloc((void**)&stage, n * sizeof(double)));
  mag::launch_spins_layout(stage, spins64(j, c), j->g, 0, j->st);
  hipError_t e = hipMemcpyAsync(spins, stage, n * sizeof(double), hipMemcpyDeviceToHost, j->st);
  hipError_t e2 = hipStreamSynchronize(j->st);
  hipFree(stage);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(MAG_ERR_HIP, "get_copy_spins failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  return MAG_OK;
}

int mag_job_copy_status(mag_job* j, int c, int* sweeps, int* converged, double* F) {
  CHK(check_copy(j, c));
  CHK(set_device(j));
  CHK(read_ctl(j));
  if (sweeps) *sweeps = j->ctl_h[c].sweep;
  if (converged) *converged = j->ctl_h[c].converged;
  if (F) *F = j->ctl_h[c].F_prev;
  return MAG_OK;
}

int mag_job_copy_history(mag_job* j, int c, double* F, double* rmax, double* ravg, int cap,
                         int* n) {
  CHK(check_copy(j, c));
  if (cap < 0) return fail(MAG_ERR_ARG, "bad argument");
  CHK(set_device(j));
  CHK(read_ctl(j));
  const int m = j->ctl_h[c].sweep < cap ? j->ctl_h[c].sweep : cap;
  const size_t hs = (size_t)c * j->prm.max_sweeps;
  if (m > 0) {
    if (F) HIPCHK(hipMemcpy(F, j->hF + hs, m * sizeof(double), hipMemcpyDeviceToHost));
    if (rmax) HIPCHK(hipMemcpy(rmax, j->hR + hs, m * sizeof(double), hipMemcpyDeviceToHost));
    if (ravg) HIPCHK(hipMemcpy(ravg, j->hA + hs, m * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (n) *n = m;
  return MAG_OK;
}

int mag_job_best_copy(mag_job* j, int* copy, double* F) {
  if (!j || !copy) return fail(MAG_ERR_ARG, "NULL argument");
  CHK(set_device(j));
  CHK(read_ctl(j));
  int best = 0;
  for (int c = 1; c < j->ncopies; c++)
    if (j->ctl_h[c].F_prev > j->ctl_h[best].F_prev) best = c;
  *copy = best;
  if (F) *F = j->ctl_h[best].F_prev;
  return MAG_OK;
}

int mag_job_download_copy_links(mag_job* j, int c, double* links) {
  CHK(check_copy(j, c));
  if (!links) return fail(MAG_ERR_ARG, "NULL links");
  if (!j->loaded) return fail(MAG_ERR_STATE, "no configuration uploaded");
  if (!j->spin) {
    if (c != 0) return fail(MAG_ERR_ARG, "a link-form job holds copy 0 only");
    return mag_job_download_links(j, links);
  }
  CHK(set_device(j));
  const size_t n = (size_t)j->V * 4 * j->nr_in;
  double *soa = nullptr, *stage = nullptr;
  HIPCHK(hipMalloc((void**)&soa, j->nreal * sizeof(double)));
  hipError_t e = hipMalloc((void**)&stage, n * sizeof(double));
  if (e != hipSuccess) {
    hipFree(soa);
    return fail(MAG_ERR_NOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
  }
  if (su3_site(j)) {
    su3_copy_links(j, c, soa);
    mag::launch_su3_soa_to_aos<double>(soa, stage, j->g, j->st);
  } else {
    mag::launch_spin_links(j->U64, spins64(j, c), soa, j->g, j->st);
    mag::launch_su2_soa_to_aos<double>(soa, stage, j->g, j->st);
  }
  e = hipMemcpyAsync(links, stage, n * sizeof(double), hipMemcpyDeviceToHost, j->st);
  hipError_t e2 = hipStreamSynchronize(j->st);
  hipFree(soa);
  hipFree(stage);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(MAG_ERR_HIP, "download failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  return MAG_OK;
}

// ------------------------------------------------------------ one-shot
static int oneshot(int group, const int* dims, const double* links, int what, double* out) {
  if (!links || !out) return fail(MAG_ERR_ARG, "NULL argument");
  mag_params p;
  mag_params_default(&p);
  p.group = group;
  for (int i = 0; i < 4; i++) p.dims[i] = dims ? dims[i] : 0;
  p.max_sweeps = 1;
  CHK(check_params(&p));
  mag_job* j = nullptr;
  CHK(mag_job_create(&p, &j));
  int r = mag_job_upload_links(j, links);
  if (r == MAG_OK) r = what == 0 ? mag_job_functional(j, out) : mag_job_plaquette(j, out);
  mag_job_destroy(j);
  return r;
}

int mag_su2_functional(const int* dims, const double* links, double* F) {
  return oneshot(MAG_SU2, dims, links, 0, F);
}
int mag_su3_functional(const int* dims, const double* links18, double* F) {
  return oneshot(MAG_SU3, dims, links18, 0, F);
}
int mag_plaquette(int group, const int* dims, const double* links, double* P) {
  if (group != MAG_SU2 && group != MAG_SU3) return fail(MAG_ERR_ARG, "group must be 2 or 3");
  return oneshot(group, dims, links, 1, P);
}

}  // extern "C"

extern "C" int mag_hbm_copy_bandwidth(int device, uint64_t bytes, int iters, double* gbs) {
  if (!gbs || iters < 1 || bytes < 16) return fail(MAG_ERR_ARG, "bad argument");
  bytes &= ~(uint64_t)15;
  HIPCHK(hipSetDevice(device));
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, bytes));
  hipError_t e = hipMalloc(&b, bytes);
  if (e != hipSuccess) {
    hipFree(a);
    return fail(MAG_ERR_NOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
  }
  hipMemset(a, 0, bytes);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  mag::launch_copy(a, b, bytes, nullptr);  // warm-up
  hipEventRecord(e0, nullptr);
  for (int i = 0; i < iters; i++) mag::launch_copy(a, b, bytes, nullptr);
  hipEventRecord(e1, nullptr);
  e = hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "copy failed: %s", hipGetErrorString(e));
  *gbs = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
  return MAG_OK;
}
