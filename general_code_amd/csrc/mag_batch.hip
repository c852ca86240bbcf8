// mag_batch.hip -- batched, graph-captured sweeps over many MAG jobs.
//
// The active members are split into classes: SU(2) / SU(3) x FP32 working
// links of the mixed phase / FP64 links, and spin-form gauge-copy sets by
// copy count.  One sweep of the batch is
//   for parity 0, 1: one half-sweep launch per non-empty class
//   one reduction launch (a workgroup per active job, per copy of a copy set)
//   every reunit_every sweeps: one FP64 rebuild + one normalisation launch
//   per FP32 class
// and `replay` sweeps (default 20) are captured in one hipGraph that is
// replayed until the membership or a job's phase changes.  The device
// tables (BatchJob) map a workgroup to its job; a job's workgroups run the
// code and partition of its own single launch, so per-job results do not
// depend on the batch.
#include "mag_capi_util.h"
#include "mag_gpu.h"
#include "mag_internal.h"
#include "mag_job_impl.h"

#include <algorithm>
#include <cstdlib>
#include <new>
#include <vector>

using mag::BatchJob;
using mag::fail;

namespace {

// 0..3: link form (SU(2)/SU(3) x FP32/FP64); kSpinClass + K - 1 (+
// MAG_MAX_COPIES with FP32 links): spin-form copy sets of K copies (one
// kernel instance per K and link type)
// kSiteClass3 (+1 with FP32 links): SU(3) site-form copy sets
constexpr int kSpinClass = 4;
constexpr int kSiteClass3 = kSpinClass + 2 * MAG_MAX_COPIES;
constexpr int kClasses = kSiteClass3 + 2;

struct ClassTable {
  int group = 2;
  bool fp32 = false;
  std::vector<BatchJob> h;
  BatchJob* d = nullptr;
  size_t cap = 0;
  std::vector<BatchJob> red;  // reduction entries (one per job, per copy)
  BatchJob* d_red = nullptr;
  size_t cap_red = 0;
  int nblocks = 0;   // half-sweep grid
  int nrblocks = 0;  // rebuild grid (FP32 classes)
};

int upload_table(std::vector<BatchJob>& h, BatchJob** d, size_t* cap) {
  if (h.size() > *cap) {
    hipFree(*d);
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
  hipStream_t st = nullptr;
  std::vector<mag_job*> members;
  std::vector<mag_job*> active;  // members not finished
  int replay = 20;
  int reunit = 0;  // shared reunit_every of mixed members (0: none yet)
  ClassTable cls[kClasses];
  std::vector<BatchJob> all;
  BatchJob* d_all = nullptr;
  size_t cap_all = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool dirty = true;
  // MAG_BATCH_DIRECT=1 in the environment: issue each replay's launches on
  // the stream instead of replaying a graph (for tools that do not follow
  // graph launches); the launches and results are the same
  bool direct = false;
  // Jobs of different classes never share data, so each class's sweeps form
  // an independent chain: with `fork` the chains run on streams of their own
  // (graph branches) and overlap on the device; without it
  // (MAG_BATCH_SERIAL=1) every sweep launches all classes in turn on one
  // stream.  Per-job results are the same either way.
  bool fork = true;
  hipStream_t fs[kClasses] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[kClasses] = {};
};

namespace {

int class_of(const mag_job* j) {
  if (j->spin && j->prm.group == MAG_SU3) return kSiteClass3 + (mag::job_mixed32(j) ? 1 : 0);
  if (j->spin) return kSpinClass + j->ncopies - 1 + (mag::job_mixed32(j) ? MAG_MAX_COPIES : 0);
  return (j->prm.group == MAG_SU2 ? 0 : 2) + (mag::job_mixed32(j) ? 0 : 1);
}

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
  b.nparts = 2 * j->nw;
  b.nrparts = mag::rebuild_blocks_of(j->g);
  b.block0 = 0;
  b.rb0 = 0;
  b.S = (j->spin && j->prm.group == MAG_SU3 && fp32) ? (void*)j->S32 : (void*)j->S;
  b.sstride = j->sstride;
  b.pstride = j->pstride;
  b.ncopies = j->spin ? j->ncopies : 0;
  return b;
}

void drop_graph(mag_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

hipError_t enqueue_replay(mag_batch* b);
int ensure_streams(mag_batch* b);

// rebuild the device tables and re-capture the replay graph
int refresh(mag_batch* b) {
  if (!b->dirty) return MAG_OK;
  drop_graph(b);
  b->all.clear();
  for (int c = 0; c < kClasses; c++) {
    b->cls[c].group = (c < 2 || (c >= kSpinClass && c < kSiteClass3)) ? 2 : 3;
    b->cls[c].fp32 = c < kSpinClass && (c % 2) == 0;
    b->cls[c].h.clear();
    b->cls[c].red.clear();
    b->cls[c].nblocks = 0;
    b->cls[c].nrblocks = 0;
  }
  for (mag_job* j : b->active) {
    ClassTable& t = b->cls[class_of(j)];
    BatchJob d = describe(j);
    d.block0 = t.nblocks;
    d.rb0 = t.nrblocks;
    t.nblocks += d.nb;
    t.nrblocks += mag::rebuild_blocks_of(j->g);
    t.h.push_back(d);
    if (!j->spin) {
      b->all.push_back(d);
      t.red.push_back(d);
    } else {
      // one reduction entry per copy
      for (int c = 0; c < j->ncopies; c++) {
        BatchJob e = d;
        const size_t hs = (size_t)c * j->prm.max_sweeps;
        e.part = j->part + (size_t)c * j->pstride;
        e.ctl = j->ctl + c;
        e.hF = j->hF + hs;
        e.hR = j->hR + hs;
        e.hA = j->hA + hs;
        b->all.push_back(e);
        t.red.push_back(e);
      }
    }
  }
  for (int c = 0; c < kClasses; c++) {
    ClassTable& t = b->cls[c];
    CHK(upload_table(t.h, &t.d, &t.cap));
    CHK(upload_table(t.red, &t.d_red, &t.cap_red));
  }
  CHK(upload_table(b->all, &b->d_all, &b->cap_all));
  b->dirty = false;
  CHK(ensure_streams(b));
  if (b->active.empty() || b->direct) return MAG_OK;
  HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  const hipError_t er = enqueue_replay(b);
  hipError_t e = hipStreamEndCapture(b->st, &b->graph);
  if (er != hipSuccess) return fail(MAG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(er));
  if (e != hipSuccess) return fail(MAG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return MAG_OK;
}

// the launches of one replay (captured into the graph, or issued directly)
// half-sweep of one class
void launch_class_half(mag_batch* b, int c, int p, hipStream_t st) {
  const ClassTable& t = b->cls[c];
  if (c < kSpinClass) {
    mag::launch_half_batch(t.group, t.fp32, t.d, (int)t.h.size(), t.nblocks, p, st);
  } else if (c >= kSiteClass3) {
    mag::launch_su3_site_copies_batch(c == kSiteClass3 + 1, t.d, (int)t.h.size(), t.nblocks, p,
                                      st);
  } else {
    const int k = (c - kSpinClass) % MAG_MAX_COPIES + 1;
    const bool fp32 = c >= kSpinClass + MAG_MAX_COPIES;
    mag::launch_copies_half_batch(k, fp32, t.d, (int)t.h.size(), t.nblocks, p, st);
  }
}

// sweep i's rebuild of a link-form FP32 class
void launch_class_rebuild(mag_batch* b, int c, int i, hipStream_t st) {
  const ClassTable& t = b->cls[c];
  if (c < kSpinClass && t.fp32 && b->reunit > 0 && (i + 1) % b->reunit == 0) {
    mag::launch_rebuild_batch(t.group, t.d, (int)t.h.size(), t.nrblocks, st);
    mag::launch_reduce_batch(t.d_red, (int)t.red.size(), mag::RED_REBUILD, st);
  }
}

int live_classes(const mag_batch* b, std::vector<int>& live) {
  live.clear();
  for (int c = 0; c < kClasses; c++)
    if (!b->cls[c].h.empty()) live.push_back(c);
  return (int)live.size();
}

// the side streams and join events of the forked replay (created on first
// use, outside any capture)
int ensure_streams(mag_batch* b) {
  std::vector<int> live;
  const int n = live_classes(b, live);
  if (!b->fork || n < 2) return MAG_OK;
  for (int k = 0; k < n; k++) {
    if (!b->fs[k]) HIPCHK(hipStreamCreateWithFlags(&b->fs[k], hipStreamNonBlocking));
    if (!b->ev_join[k]) HIPCHK(hipEventCreateWithFlags(&b->ev_join[k], hipEventDisableTiming));
  }
  return MAG_OK;
}

// the launches of one replay (captured into the graph, or issued directly);
// returns the first error of the stream-ordering calls
hipError_t enqueue_replay(mag_batch* b) {
  std::vector<int> live;
  live_classes(b, live);
  if (!b->fork || live.size() < 2) {
    for (int i = 0; i < b->replay; i++) {
      for (int p = 0; p < 2; p++)
        for (int c : live) launch_class_half(b, c, p, b->st);
      mag::launch_reduce_batch(b->d_all, (int)b->all.size(), mag::RED_SWEEP, b->st);
      for (int c : live) launch_class_rebuild(b, c, i, b->st);
    }
    return hipSuccess;
  }
  hipError_t e = hipEventRecord(b->ev_fork, b->st);
  for (size_t k = 0; k < live.size() && e == hipSuccess; k++) {
    const int c = live[k];
    hipStream_t s = b->fs[k];
    e = hipStreamWaitEvent(s, b->ev_fork, 0);
    if (e != hipSuccess) break;
    const ClassTable& t = b->cls[c];
    for (int i = 0; i < b->replay; i++) {
      launch_class_half(b, c, 0, s);
      launch_class_half(b, c, 1, s);
      mag::launch_reduce_batch(t.d_red, (int)t.red.size(), mag::RED_SWEEP, s);
      launch_class_rebuild(b, c, i, s);
    }
    e = hipEventRecord(b->ev_join[k], s);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->st, b->ev_join[k], 0);
  }
  return e;
}

bool finished(const mag_job* j) { return mag::job_done(j) && !mag::job_mixed32(j); }

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
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
  if (e != hipSuccess) {
    if (b->st) (void)hipStreamDestroy(b->st);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
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
  (void)hipStreamSynchronize(b->st);
  drop_graph(b);
  for (mag_job* j : b->members) j->batch = nullptr;
  for (int c = 0; c < kClasses; c++) {
    (void)hipFree(b->cls[c].d);
    (void)hipFree(b->cls[c].d_red);
    if (b->fs[c]) (void)hipStreamDestroy(b->fs[c]);
    if (b->ev_join[c]) (void)hipEventDestroy(b->ev_join[c]);
  }
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  (void)hipFree(b->d_all);
  (void)hipStreamDestroy(b->st);
  delete b;
  return MAG_OK;
}

int mag_batch_add(mag_batch* b, mag_job* j) {
  if (!b || !j) return fail(MAG_ERR_ARG, "NULL argument");
  if (j->batch) return fail(MAG_ERR_STATE, "job is already in a batch");
  if (!j->loaded) return fail(MAG_ERR_STATE, "job has no configuration");
  if (j->prm.device != b->device) return fail(MAG_ERR_ARG, "job and batch on different devices");
  if (j->prm.precision == MAG_PREC_MIXED && !j->spin) {
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
  if (!finished(j)) b->active.push_back(j);
  b->dirty = true;
  return MAG_OK;
}

int mag_batch_remove(mag_batch* b, mag_job* j) {
  if (!b || !j) return fail(MAG_ERR_ARG, "NULL argument");
  auto it = std::find(b->members.begin(), b->members.end(), j);
  if (it == b->members.end()) return fail(MAG_ERR_ARG, "job is not in this batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->st));
  b->members.erase(it);
  auto a = std::find(b->active.begin(), b->active.end(), j);
  if (a != b->active.end()) {
    b->active.erase(a);
    b->dirty = true;
  }
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
  CHK(refresh(b));
  if (b->direct) {
    if (b->active.empty()) return MAG_OK;
    for (int i = 0; i < n; i++) HIPCHK(enqueue_replay(b));
    HIPCHK(hipGetLastError());
  } else {
    if (!b->exec) return MAG_OK;
    for (int i = 0; i < n; i++) HIPCHK(hipGraphLaunch(b->exec, b->st));
  }
  for (mag_job* j : b->active)
    if (mag::job_mixed32(j)) j->since_rebuild = (j->since_rebuild + n * b->replay) % j->prm.reunit_every;
  return MAG_OK;
}

int mag_batch_poll(mag_batch* b, int* n_running, int* n_finished) {
  if (!b) return fail(MAG_ERR_ARG, "NULL batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->st));
  for (mag_job* j : b->active)
    HIPCHK(hipMemcpyAsync(j->ctl_h, j->ctl, j->ncopies * sizeof(mag::JobCtl),
                          hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  std::vector<mag_job*> still;
  for (mag_job* j : b->active) {
    if (mag::job_done(j) && mag::job_mixed32(j)) {
      CHK(mag::job_finalize_mixed(j));  // FP32 phase over: FP64 rebuild, FP64 class
      b->dirty = true;
    }
    bool switched = false;
    CHK(mag::job_maybe_start_fp32(j, &switched));  // FP64 warm-up over: FP32 class
    if (switched) b->dirty = true;
    if (finished(j)) b->dirty = true;
    else still.push_back(j);
  }
  b->active.swap(still);
  if (n_running) *n_running = (int)b->active.size();
  if (n_finished) *n_finished = (int)(b->members.size() - b->active.size());
  return MAG_OK;
}

int mag_batch_run(mag_batch* b) {
  if (!b) return fail(MAG_ERR_ARG, "NULL batch");
  int running = 0;
  CHK(mag_batch_poll(b, &running, nullptr));
  while (running > 0) {
    CHK(mag_batch_replays(b, 1));
    CHK(mag_batch_poll(b, &running, nullptr));
  }
  return MAG_OK;
}

int mag_batch_replays_timed(mag_batch* b, int n, double* ms) {
  if (!b || n < 1 || !ms) return fail(MAG_ERR_ARG, "bad argument");
  HIPCHK(hipSetDevice(b->device));
  CHK(refresh(b));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, b->st));
  CHK(mag_batch_replays(b, n));
  HIPCHK(hipEventRecord(e1, b->st));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = t;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return MAG_OK;
}

}  // extern "C"
