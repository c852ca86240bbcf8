// mag_batch.hip -- batched, graph-captured sweeps over many MAG jobs.
//
// The active members are split into classes: SU(2) / SU(3) x FP32 working
// links of the mixed phase / FP64 links, spin-form gauge-copy sets by copy
// count and link precision, SU(3) site-form sets by precision (FP32: and
// copy count).  One sweep of
// a class is
//   for parity 0, 1: one half-sweep launch whose grid concatenates the
//                    member jobs' grids
//   one reduction launch (a workgroup per job, per copy of a copy set);
//   link-form FP32 classes instead reduce once per reunit_every sweeps,
//   after the black half-sweep that rebuilds the links from the FP64
//   master, and FP32 SU(3) site-form classes after the F evaluation in FP64
//   that ends their period
// and `replay` sweeps (default 20) plus a gather of every member's control
// block into pinned host memory are captured in one hipGraph per class.
//
// Classes share no data, so each class is an independent chain on a stream
// of its own: the scheduler (mag_batch_step) relaunches a class's graph as
// soon as that class's previous replay has finished and its members have
// been examined, without waiting for the other classes.  A class's tables
// and graph are rebuilt only while it is idle; a job that changes class
// (FP64 warm-up over, FP32 phase over) joins its new class at that class's
// next launch.  The device tables (BatchJob) map a workgroup to its job; a
// job's workgroups run the code and partition of its own single launch, so
// per-job results do not depend on the batch.
#include "mag_capi_util.h"
#include "mag_gpu.h"
#include "mag_internal.h"
#include "mag_job_impl.h"

#include <algorithm>
#include <cstdlib>
#include <new>
#include <thread>
#include <utility>
#include <vector>

using mag::BatchJob;
using mag::fail;
using mag::JobCtl;

namespace {

// 0..3: link form (SU(2)/SU(3) x FP32/FP64); kSpinClass + K - 1 (+
// MAG_MAX_COPIES with FP32 links): spin-form copy sets of K copies (one
// kernel instance per K and link type)
// kSiteClass3: SU(3) site-form copy sets with FP64 links (any copy count);
// kSiteClass3 + K: FP32 sets of K copies (one kernel instance per K)
constexpr int kSpinClass = 4;
constexpr int kSiteClass3 = kSpinClass + 2 * MAG_MAX_COPIES;
constexpr int kClasses = kSiteClass3 + 1 + MAG_MAX_COPIES;

struct ClassRun {
  int group = 2;
  bool fp32 = false;
  std::vector<mag_job*> jobs;  // current members (host view)
  bool dirty = true;           // jobs changed since the tables were built
  // what the device tables and the captured graph reference
  std::vector<BatchJob> h;                      // half-sweep table
  std::vector<BatchJob> red;                    // reduction entries
  std::vector<std::pair<mag_job*, int>> owner;  // red[i] -> (job, copy)
  std::vector<mag_job*> cap_jobs;
  BatchJob* d = nullptr;
  size_t cap = 0;
  BatchJob* d_red = nullptr;
  size_t cap_red = 0;
  JobCtl* ctl_out = nullptr;  // pinned host memory, one entry per red[i]
  size_t cap_out = 0;
  int nblocks = 0;     // half-sweep grid
  int red_blocks = 0;  // per-sweep reduction grid (red_blocks per entry)
  mag::Shape shape;    // link-form classes: launch shape of the half-sweep
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool inflight = false;  // replays launched and not yet examined
};

int upload_table(std::vector<BatchJob>& h, BatchJob** d, size_t* cap) {
  if (h.size() > *cap) {
    (void)hipFree(*d);
    *d = nullptr;
    *cap = 0;
    const size_t c = std::max<size_t>(h.size(), 16);
    HIPCHK(hipMalloc((void**)d, c * sizeof(BatchJob)));
    *cap = c;
  }
  if (!h.empty()) HIPCHK(hipMemcpy(*d, h.data(), h.size() * sizeof(BatchJob), hipMemcpyHostToDevice));
  return MAG_OK;
}

}  // namespace

struct mag_batch {
  int device = 0;
  hipStream_t st = nullptr;  // the batch stream: joins every class after replays()
  std::vector<mag_job*> members;
  std::vector<mag_job*> active;  // members not finished
  int replay = 20;
  int reunit = 0;  // shared reunit_every of mixed link-form members (0: none yet)
  ClassRun cls[kClasses];
  // MAG_BATCH_DIRECT=1 in the environment: issue each replay's launches on
  // the stream instead of replaying a graph (for tools that do not follow
  // graph launches); the launches and results are the same
  bool direct = false;
  // MAG_BATCH_SERIAL=1: every class on the batch stream, one after the
  // other, instead of a stream per class; per-job results are the same
  bool fork = true;
};

namespace {

int class_of(const mag_job* j) {
  if (j->spin && j->prm.group == MAG_SU3) return kSiteClass3 + (mag::job_mixed32(j) ? j->ncopies : 0);
  if (j->spin) return kSpinClass + j->ncopies - 1 + (mag::job_mixed32(j) ? MAG_MAX_COPIES : 0);
  return (j->prm.group == MAG_SU2 ? 0 : 2) + (mag::job_mixed32(j) ? 0 : 1);
}

bool finished(const mag_job* j) { return mag::job_done(j) && !mag::job_mixed32(j); }

BatchJob describe(mag_job* j) {
  BatchJob b;
  const bool fp32 = mag::job_mixed32(j);
  b.U = fp32 ? (void*)j->W32 : (void*)j->U64;
  b.G = fp32 ? (void*)j->G32 : nullptr;
  b.U0 = j->U64;
  b.g = j->g;
  b.part = j->part;
  b.ctl = j->ctl;
  b.hF = j->hF;
  b.hR = j->hR;
  b.hA = j->hA;
  b.omega = j->prm.omega;
  b.invF = j->invF;
  b.invV = j->invV;
  b.nb = j->nb;
  b.nw = j->nw;
  b.nparts = j->spin ? 2 * j->nw : j->nw;  // link form: the black triples (red folded in)
  b.roff = j->spin ? 0 : 3 * j->nw;
  b.block0 = 0;
  b.red0 = 0;
  b.S = (j->spin && fp32) ? (void*)j->S32 : (void*)j->S;  // FP32 phase: FP32 spins / gauges
  b.sstride = j->sstride;
  b.pstride = j->pstride;
  b.ncopies = j->spin ? j->ncopies : 0;
  return b;
}

void drop_graph(ClassRun& t) {
  if (t.exec) (void)hipGraphExecDestroy(t.exec);
  if (t.graph) (void)hipGraphDestroy(t.graph);
  t.exec = nullptr;
  t.graph = nullptr;
}

// half-sweep of one class
void launch_class_half(const mag_batch* b, int c, int p, hipStream_t st, int slot = 0) {
  const ClassRun& t = b->cls[c];
  if (c < kSpinClass) {
    mag::launch_half_batch(t.group, t.fp32, t.shape, t.d, (int)t.h.size(), t.nblocks, p, st);
  } else if (c >= kSiteClass3) {
    mag::launch_su3_site_copies_batch(c - kSiteClass3, c > kSiteClass3, t.d, (int)t.h.size(),
                                      t.nblocks, p, st, slot);
  } else {
    const int k = (c - kSpinClass) % MAG_MAX_COPIES + 1;
    const bool fp32 = c >= kSpinClass + MAG_MAX_COPIES;
    mag::launch_copies_half_batch(k, fp32, t.d, (int)t.h.size(), t.nblocks, p, st);
  }
}

// one replay of class c: `replay` sweeps, then the control-block gather
void enqueue_class_replay(const mag_batch* b, int c, hipStream_t s) {
  const ClassRun& t = b->cls[c];
  const int nred = (int)t.red.size();
  for (int i = 0; i < b->replay; i++) {
    if (c < kSpinClass && t.fp32) {
      // mixed FP32 phase, as a job alone (mag_capi.hip enqueue_sweep):
      // periods of reunit sweeps, the replay a whole number of them
      const int R = b->reunit, k = i % R;
      mag::launch_half_batch(t.group, true, t.shape, t.d, (int)t.h.size(), t.nblocks, 0, s, k,
                             k == 0 ? mag::HALF_RENORM : mag::HALF_PLAIN);
      mag::launch_half_batch(t.group, true, t.shape, t.d, (int)t.h.size(), t.nblocks, 1, s, k,
                             k == R - 1 ? mag::HALF_REBUILD : mag::HALF_PLAIN);
      if (k == R - 1) {
        mag::launch_reduce_slots_batch(t.d_red, nred, R, s);
        mag::launch_reduce_period_batch(t.d_red, nred, R, 1, s);
      }
      continue;
    }
    if (c > kSiteClass3) {
      // FP32 SU(3) site-form sets, as a job alone (mag_capi.hip
      // enqueue_sweep): sweep partials in slots, and at the end of each
      // period of reunit sweeps F in FP64, the slots and the stopping test
      const int R = b->reunit, k = i % R;
      launch_class_half(b, c, 0, s, k);
      launch_class_half(b, c, 1, s, k);
      if (k == R - 1) {
        mag::launch_su3_site_F_batch(t.d, (int)t.h.size(), t.nblocks, s);
        mag::launch_reduce_slots_batch(t.d_red, nred, R, s);
        mag::launch_reduce_period_batch(t.d_red, nred, R, 1, s);
      }
      continue;
    }
    launch_class_half(b, c, 0, s);
    launch_class_half(b, c, 1, s);
    mag::launch_reduce_batch(t.d_red, nred, t.red_blocks, s);
  }
  mag::launch_gather_ctl(t.d_red, (int)t.red.size(), t.ctl_out, s);
}

// the class's stream and completion event (created on first use, outside
// any capture)
int ensure_stream(mag_batch* b, int c) {
  ClassRun& t = b->cls[c];
  if (!t.s) {
    if (b->fork) HIPCHK(hipStreamCreateWithFlags(&t.s, hipStreamNonBlocking));
    else t.s = b->st;
  }
  if (!t.ev) HIPCHK(hipEventCreateWithFlags(&t.ev, hipEventDisableTiming));
  return MAG_OK;
}

// rebuild class c's device tables and graph from its members; c is idle
int refresh_class(mag_batch* b, int c) {
  ClassRun& t = b->cls[c];
  if (!t.dirty) return MAG_OK;
  drop_graph(t);
  t.group = (c < 2 || (c >= kSpinClass && c < kSiteClass3)) ? 2 : 3;
  t.fp32 = c < kSpinClass && (c % 2) == 0;
  t.h.clear();
  t.red.clear();
  t.red_blocks = 0;
  t.owner.clear();
  t.cap_jobs = t.jobs;
  t.nblocks = 0;
  t.shape = mag::Shape{};
  if (c < kSpinClass && !t.jobs.empty())
    CHK(mag::batch_shape(t.group, t.fp32, t.jobs, &t.shape));  // autotuned; same results any shape
  for (mag_job* j : t.jobs) {
    BatchJob d = describe(j);
    if (c < kSpinClass) {
      const size_t tile = (size_t)t.shape.nt * t.shape.spt;
      d.nb = (int)((mag::sweep_lanes(j->g, j->prm.group) + tile - 1) / tile);
    }
    d.block0 = t.nblocks;
    t.nblocks += d.nb;
    t.h.push_back(d);
    if (!j->spin) {
      d.red0 = t.red_blocks;
      t.red_blocks += mag::red_blocks(d.nparts);
      t.red.push_back(d);
      t.owner.emplace_back(j, 0);
    } else {
      // one reduction entry per copy
      for (int k = 0; k < j->ncopies; k++) {
        BatchJob e = d;
        const size_t hs = (size_t)k * j->prm.max_sweeps;
        e.part = j->part + (size_t)k * j->pstride;
        e.ctl = j->ctl + k;
        e.hF = j->hF + hs;
        e.hR = j->hR + hs;
        e.hA = j->hA + hs;
        e.red0 = t.red_blocks;
        t.red_blocks += mag::red_blocks(e.nparts);
        t.red.push_back(e);
        t.owner.emplace_back(j, k);
      }
    }
  }
  CHK(upload_table(t.h, &t.d, &t.cap));
  CHK(upload_table(t.red, &t.d_red, &t.cap_red));
  if (t.red.size() > t.cap_out) {
    if (t.ctl_out) (void)hipHostFree(t.ctl_out);
    t.ctl_out = nullptr;
    t.cap_out = 0;
    const size_t n = std::max<size_t>(t.red.size(), 16);
    HIPCHK(hipHostMalloc((void**)&t.ctl_out, n * sizeof(JobCtl), hipHostMallocMapped | hipHostMallocCoherent));
    t.cap_out = n;
  }
  t.dirty = false;
  CHK(ensure_stream(b, c));
  if (t.jobs.empty() || b->direct) return MAG_OK;
  HIPCHK(hipStreamBeginCapture(t.s, hipStreamCaptureModeThreadLocal));
  enqueue_class_replay(b, c, t.s);
  const hipError_t el = hipGetLastError();
  const hipError_t e = hipStreamEndCapture(t.s, &t.graph);
  if (el != hipSuccess) return fail(MAG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(el));
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  HIPCHK(hipGraphInstantiate(&t.exec, t.graph, nullptr, nullptr, 0));
  return MAG_OK;
}

// launch n replays of idle class c (nothing if it has no members)
int launch_class(mag_batch* b, int c, int n) {
  ClassRun& t = b->cls[c];
  CHK(refresh_class(b, c));
  if (t.jobs.empty() || n == 0) return MAG_OK;
  for (int i = 0; i < n; i++) {
    if (b->direct) enqueue_class_replay(b, c, t.s);
    else HIPCHK(hipGraphLaunch(t.exec, t.s));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(t.ev, t.s));
  t.inflight = true;
  for (mag_job* j : t.cap_jobs)
    if (mag::job_mixed32(j) && (!j->spin || j->prm.group == MAG_SU3))
      j->since_rebuild = (j->since_rebuild + n * b->replay) % j->prm.reunit_every;
  return MAG_OK;
}

void erase_job(std::vector<mag_job*>& v, mag_job* j) {
  auto it = std::find(v.begin(), v.end(), j);
  if (it != v.end()) v.erase(it);
}

// class c's launched replays are complete: take every member's control block
// from the gather, move mixed jobs between phases (and classes), drop
// finished jobs
int complete_class(mag_batch* b, int c) {
  ClassRun& t = b->cls[c];
  t.inflight = false;
  for (size_t i = 0; i < t.owner.size(); i++) t.owner[i].first->ctl_h[t.owner[i].second] = t.ctl_out[i];
  for (mag_job* j : t.cap_jobs) {
    if (std::find(t.jobs.begin(), t.jobs.end(), j) == t.jobs.end()) continue;  // removed meanwhile
    bool moved = false;
    if (mag::job_done(j) && mag::job_mixed32(j)) {
      CHK(mag::job_finalize_mixed(j));  // FP32 phase over: FP64 rebuild, FP64 class
      moved = true;
    }
    bool switched = false;
    CHK(mag::job_maybe_start_fp32(j, &switched));  // FP64 warm-up over: FP32 class
    moved = moved || switched;
    if (moved) HIPCHK(hipStreamSynchronize(j->st));  // before another stream uses the job
    if (finished(j)) {
      erase_job(t.jobs, j);
      erase_job(b->active, j);
      t.dirty = true;
    } else if (class_of(j) != c) {
      erase_job(t.jobs, j);
      t.dirty = true;
      ClassRun& u = b->cls[class_of(j)];
      u.jobs.push_back(j);
      u.dirty = true;
    }
  }
  return MAG_OK;
}

// wait for class c's replays and examine them
int wait_class(mag_batch* b, int c) {
  ClassRun& t = b->cls[c];
  if (!t.inflight) return MAG_OK;
  HIPCHK(hipEventSynchronize(t.ev));
  return complete_class(b, c);
}

int wait_all(mag_batch* b) {
  for (int c = 0; c < kClasses; c++) CHK(wait_class(b, c));
  return MAG_OK;
}

void counts(const mag_batch* b, int* n_running, int* n_finished) {
  if (n_running) *n_running = (int)b->active.size();
  if (n_finished) *n_finished = (int)(b->members.size() - b->active.size());
}

}  // namespace

extern "C" {

int mag_batch_create(int device, mag_batch** out) {
  if (!out) return fail(MAG_ERR_ARG, "out is NULL");
  *out = nullptr;
  mag_batch* b = new (std::nothrow) mag_batch();
  if (!b) return fail(MAG_ERR_NOMEM, "host allocation failed");
  b->device = device;
  const char* dv = std::getenv("MAG_BATCH_DIRECT");
  b->direct = dv && dv[0] == '1';
  const char* sv = std::getenv("MAG_BATCH_SERIAL");
  b->fork = !(sv && sv[0] == '1');
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return fail(MAG_ERR_HIP, "batch setup failed: %s", hipGetErrorString(e));
  }
  *out = b;
  return MAG_OK;
}

int mag_batch_destroy(mag_batch* b) {
  if (!b) return MAG_OK;
  // best-effort teardown: errors here have nowhere to go
  (void)hipSetDevice(b->device);
  for (int c = 0; c < kClasses; c++) {
    ClassRun& t = b->cls[c];
    if (t.inflight) (void)hipEventSynchronize(t.ev);
  }
  (void)hipStreamSynchronize(b->st);
  for (mag_job* j : b->members) j->batch = nullptr;
  for (int c = 0; c < kClasses; c++) {
    ClassRun& t = b->cls[c];
    drop_graph(t);
    (void)hipFree(t.d);
    (void)hipFree(t.d_red);
    if (t.ctl_out) (void)hipHostFree(t.ctl_out);
    if (t.ev) (void)hipEventDestroy(t.ev);
    if (t.s && t.s != b->st) (void)hipStreamDestroy(t.s);
  }
  (void)hipStreamDestroy(b->st);
  delete b;
  return MAG_OK;
}

int mag_batch_add(mag_batch* b, mag_job* j) {
  if (!b || !j) return fail(MAG_ERR_ARG, "NULL argument");
  if (j->batch) return fail(MAG_ERR_STATE, "job is already in a batch");
  if (!j->loaded) return fail(MAG_ERR_STATE, "job has no configuration");
  if (j->prm.device != b->device) return fail(MAG_ERR_ARG, "job and batch on different devices");
  if (j->prm.track_gauge) return fail(MAG_ERR_ARG, "a job that tracks its gauge runs alone");
  if (j->prm.precision == MAG_PREC_MIXED && (!j->spin || j->prm.group == MAG_SU3)) {
    if (b->reunit == 0) {
      if (b->replay % j->prm.reunit_every != 0)
        return fail(MAG_ERR_ARG, "reunit_every %d does not divide the replay length %d",
                    j->prm.reunit_every, b->replay);
      b->reunit = j->prm.reunit_every;
    } else if (j->prm.reunit_every != b->reunit) {
      return fail(MAG_ERR_ARG, "batch members must share reunit_every (%d)", b->reunit);
    }
    if (j->since_rebuild != 0)
      return fail(MAG_ERR_STATE, "job is between two rebuilds (sweeps not a multiple of %d)",
                  j->prm.reunit_every);
  }
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(j->st));
  CHK(mag::job_read_ctl(j));
  j->batch = b;
  b->members.push_back(j);
  if (!finished(j)) {
    b->active.push_back(j);
    ClassRun& t = b->cls[class_of(j)];
    t.jobs.push_back(j);
    t.dirty = true;
  }
  return MAG_OK;
}

int mag_batch_remove(mag_batch* b, mag_job* j) {
  if (!b || !j) return fail(MAG_ERR_ARG, "NULL argument");
  auto it = std::find(b->members.begin(), b->members.end(), j);
  if (it == b->members.end()) return fail(MAG_ERR_ARG, "job is not in this batch");
  HIPCHK(hipSetDevice(b->device));
  // a class whose launched replays reference the job must finish them first
  for (int c = 0; c < kClasses; c++) {
    ClassRun& t = b->cls[c];
    if (t.inflight && std::find(t.cap_jobs.begin(), t.cap_jobs.end(), j) != t.cap_jobs.end())
      CHK(wait_class(b, c));
  }
  for (int c = 0; c < kClasses; c++) {
    ClassRun& t = b->cls[c];
    if (std::find(t.jobs.begin(), t.jobs.end(), j) != t.jobs.end()) {
      erase_job(t.jobs, j);
      t.dirty = true;
    }
  }
  b->members.erase(it);
  erase_job(b->active, j);
  j->batch = nullptr;
  return MAG_OK;
}

int mag_batch_size(mag_batch* b, int* n) {
  if (!b || !n) return fail(MAG_ERR_ARG, "NULL argument");
  *n = (int)b->members.size();
  return MAG_OK;
}

int mag_batch_replay_sweeps(mag_batch* b, int* n) {
  if (!b || !n) return fail(MAG_ERR_ARG, "NULL argument");
  *n = b->replay;
  return MAG_OK;
}

int mag_batch_stream(mag_batch* b, void** stream) {
  if (!b || !stream) return fail(MAG_ERR_ARG, "NULL argument");
  *stream = (void*)b->st;
  return MAG_OK;
}

int mag_batch_replays(mag_batch* b, int n) {
  if (!b || n < 0) return fail(MAG_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(b->device));
  for (int c = 0; c < kClasses; c++) {
    ClassRun& t = b->cls[c];
    if (t.inflight) CHK(wait_class(b, c));
  }
  for (int c = 0; c < kClasses; c++) {
    ClassRun& t = b->cls[c];
    CHK(launch_class(b, c, n));
    if (t.inflight && t.s != b->st) HIPCHK(hipStreamWaitEvent(b->st, t.ev, 0));
  }
  return MAG_OK;
}

int mag_batch_poll(mag_batch* b, int* n_running, int* n_finished) {
  if (!b) return fail(MAG_ERR_ARG, "NULL batch");
  HIPCHK(hipSetDevice(b->device));
  CHK(wait_all(b));
  HIPCHK(hipStreamSynchronize(b->st));
  counts(b, n_running, n_finished);
  return MAG_OK;
}

int mag_batch_step(mag_batch* b, int* n_running, int* n_finished) {
  if (!b) return fail(MAG_ERR_ARG, "NULL batch");
  HIPCHK(hipSetDevice(b->device));
  for (int c = 0; c < kClasses; c++)
    if (!b->cls[c].inflight) CHK(launch_class(b, c, 1));
  bool any = false;
  for (int c = 0; c < kClasses; c++) any = any || b->cls[c].inflight;
  if (any) {
    // wait until at least one class has finished its replay
    for (bool got = false; !got;) {
      for (int c = 0; c < kClasses; c++) {
        ClassRun& t = b->cls[c];
        if (!t.inflight) continue;
        const hipError_t e = hipEventQuery(t.ev);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess) return fail(MAG_ERR_HIP, "batch: %s", hipGetErrorString(e));
        CHK(complete_class(b, c));
        got = true;
      }
      if (!got) std::this_thread::yield();
    }
    // keep the device busy while the caller looks at the results
    for (int c = 0; c < kClasses; c++)
      if (!b->cls[c].inflight) CHK(launch_class(b, c, 1));
  }
  counts(b, n_running, n_finished);
  return MAG_OK;
}

int mag_batch_run(mag_batch* b) {
  if (!b) return fail(MAG_ERR_ARG, "NULL batch");
  int running = 0;
  CHK(mag_batch_poll(b, &running, nullptr));
  while (running > 0) CHK(mag_batch_step(b, &running, nullptr));
  return mag_batch_poll(b, nullptr, nullptr);
}

int mag_batch_replays_timed(mag_batch* b, int n, double* ms) {
  if (!b || n < 1 || !ms) return fail(MAG_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(b->device));
  CHK(wait_all(b));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, b->st));
  for (int c = 0; c < kClasses; c++) {
    ClassRun& t = b->cls[c];
    CHK(refresh_class(b, c));
    if (t.jobs.empty()) continue;
    if (t.s != b->st) HIPCHK(hipStreamWaitEvent(t.s, e0, 0));
    CHK(launch_class(b, c, n));
    if (t.s != b->st) HIPCHK(hipStreamWaitEvent(b->st, t.ev, 0));
  }
  HIPCHK(hipEventRecord(e1, b->st));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return MAG_OK;
}

}  // extern "C"
