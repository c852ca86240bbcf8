// mag_job_impl.h -- host-side state of a MAG job (C ABI handle), shared by
// the job API (mag_capi.hip) and the batch / serving scheduler
// (mag_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "mag_gpu.h"
#include "mag_internal.h"

struct mag_job {
  mag_params prm;
  mag::Geom g;
  int64_t V = 0;
  int nr = 0;        // reals per link in the SoA layout (4 or 12)
  int nr_in = 0;     // reals per link in the reference layout (4 or 18)
  size_t nreal = 0;  // reals in one link array (8 * Vh * nr)
  int nb = 0;        // workgroups per half-sweep at the default shape
  int nw = 0;        // partial triples per half-sweep (one per wave)
  mag::Shape shape[2];           // link-form launch shape, FP64 / FP32 kernel
  bool shape_ok[2] = {false, false};
  int nf = 0;        // workgroups per parity of the functional / plaquette kernels
  hipStream_t st = nullptr;
  double* U64 = nullptr;  // FP64 links (plain) / FP64 master then output (mixed)
  float* W32 = nullptr;   // mixed: FP32 working links
  float* G32 = nullptr;   // mixed: accumulated site gauge (red sites)
  double* G64 = nullptr;  // track_gauge: total gauge of the fixing, FP64
  double* part = nullptr;
  mag::JobCtl* ctl = nullptr;
  double* hF = nullptr;
  double* hR = nullptr;
  double* hA = nullptr;
  double* out3 = nullptr;
  mag::JobCtl* ctl_h = nullptr;  // pinned host mirrors
  double* out3_h = nullptr;
  bool loaded = false;
  int phase = 0;  // mixed: -1 = FP64 warm-up, 0 = FP32 sweeps, 1 = FP64 after rebuild
  int fp64_start = -1;
  int since_rebuild = 0;  // sweeps enqueued since the last FP64 rebuild
  int pending = 0;        // mixed FP32 phase: swept slots not yet reduced
  double invF = 0, invV = 0;
  mag_batch* batch = nullptr;  // batch the job belongs to, if any
  // site form (gauge copies): U64 stays read-only; copy c has its site
  // variable (SU(2) spins, SU(3) site gauges) at S + c*sstride, partials at
  // part + c*pstride, control ctl[c] / ctl_h[c] and history at
  // hF/hR/hA + c*max_sweeps
  // SU(3) in the site form: S holds the FP64 site gauges of the copies
  // (2 rows, chunked site layout).  S32: the FP32 working copies of the
  // spins / site gauges that the FP32 phase of the mixed form sweeps (S is
  // refreshed from them when read during that phase and at its end)
  bool spin = false;
  int ncopies = 1;
  double* S = nullptr;
  float* S32 = nullptr;
  size_t sstride = 0;
  size_t pstride = 0;
};

namespace mag {
// job operations shared with the batch scheduler (mag_capi.hip)
int job_read_ctl(mag_job* j);
int job_finalize_mixed(mag_job* j);
bool job_mixed32(const mag_job* j);
// every copy done / converged (read_ctl first)
bool job_done(const mag_job* j);
// mixed with an FP64 warm-up: enter the FP32 phase once the warm-up sweeps
// are done (read_ctl first); *switched = whether it did
int job_maybe_start_fp32(mag_job* j, bool* switched);
// site form: the FP64 links of copy c (G U G^+ of its site variable) into
// soa (parity-split SoA, nreal doubles), enqueued on the job's stream
void job_copy_links64(mag_job* j, int c, double* soa);
// launch shape of the job's link-form half-sweep (mag_tune.hip): cached per
// (group, precision, device, lattice), tuned on first use
int job_shape(const mag_job* j, bool fp32, Shape* out);
// launch shape of a batched link-form half-sweep over `jobs` (mag_tune.hip):
// tuned on scratch copies of the jobs through the batched kernel, cached
// per (group, precision, device, log2 of the total lanes)
int batch_shape(int group, bool fp32, const std::vector<mag_job*>& jobs, Shape* out);
}  // namespace mag
