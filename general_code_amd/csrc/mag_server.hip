// mag_server.hip -- continuous-batching MAG server (no counterpart in the
// reference; the reference runs one configuration per process,
// apps/mag_su2/mag_fixation.cpp).
//
// One scheduler thread per server owns a mag_batch and every admitted job.
// Its loop:
//   1. one batch step (mag_batch_step): every idle class of the batch
//      launches its next replay (mag_batch_replay_sweeps sweeps of its
//      jobs), the step returns when some class's replay is done and its
//      jobs have been examined (mixed jobs moved between phases), with
//      the classes relaunched;
//   2. while they run: admit queued requests in arrival order as long as
//      the HBM footprint of the admitted jobs stays within the budget (and
//      the active count within max_active): create the job, upload or
//      generate its links on the job's own stream, add it to the batch (it
//      joins at its class's next launch);
//   3. finished jobs leave the batch, their result is published and their
//      HBM freed at once.
// Requests that arrive while others are on the GPU are therefore admitted
// as memory frees up and their sweeps run in the same launches as the jobs
// already there.  A job's sweeps are exactly those of the job run alone (a
// batch runs each job's own partition and reductions), so a result does not
// depend on what else was being served.
#include "mag_capi_util.h"
#include "mag_gpu.h"
#include "mag_internal.h"
#include "mag_job_impl.h"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <set>
#include <thread>
#include <vector>

using mag::fail;

namespace {

struct Pending {
  uint64_t id;
  mag_request req;
  std::vector<double> links;  // copy of the caller's links (empty: synthetic)
  uint64_t footprint;
  double t_submit;
};

struct Active {
  uint64_t id;
  mag_job* job;
  uint64_t footprint;
  double t_submit, t_admit;
};

}  // namespace

struct mag_server {
  int device = 0;
  uint64_t budget = 0;
  int max_active = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work;  // scheduler: new request or stop
  std::condition_variable cv_done;  // waiters: a result was published
  std::deque<Pending> queue;
  std::map<uint64_t, mag_result> results;
  std::set<uint64_t> cancel;  // admitted requests to cancel (scheduler acts)
  uint64_t next_id = 1;
  uint64_t submitted = 0, finished = 0;
  bool stop = false;
  // scheduler-thread state
  mag_batch* batch = nullptr;
  std::vector<Active> active;
  uint64_t used = 0;
  mag_server_stats stats{};
  std::thread th;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void publish(mag_server* s, const mag_result& r) {
  std::lock_guard<std::mutex> g(s->mu);
  s->results[r.id] = r;
  s->finished++;
  s->cv_done.notify_all();
}

mag_result result_of(mag_server* s, const Active& a, int status) {
  mag_result r;
  std::memset(&r, 0, sizeof(r));
  r.id = a.id;
  r.t_submit = a.t_submit;
  r.t_admit = a.t_admit;
  r.status = status;
  if (status == MAG_OK || status == MAG_ERR_CANCELLED) {
    mag_job* j = a.job;
    int sw = 0, conv = 0;
    double F = 0;
    mag_job_status(j, &sw, &conv, &F);
    int best = 0;
    mag_job_best_copy(j, &best, &F);
    int swb = 0;
    mag_job_copy_status(j, best, &swb, nullptr, nullptr);
    r.sweeps = swb;
    r.converged = conv;
    r.F = F;
    r.best_copy = best;
    r.fp64_start = j->fp64_start;
    if (!conv && status == MAG_OK) r.status = MAG_ERR_NOT_CONVERGED;
  }
  r.t_done = s->now();
  return r;
}

// admit the head of the queue if it fits; returns false when it must wait
bool admit_one(mag_server* s) {
  Pending p;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->queue.empty()) return false;
    const Pending& h = s->queue.front();
    if (!s->active.empty()) {
      if (s->used + h.footprint > s->budget) return false;
      if (s->max_active > 0 && (int)s->active.size() >= s->max_active) return false;
    }
    p = std::move(s->queue.front());
    s->queue.pop_front();
  }
  Active a{p.id, nullptr, p.footprint, p.t_submit, s->now()};
  int rc = mag_job_create(&p.req.params, &a.job);
  if (rc == MAG_OK) {
    if (!p.links.empty()) rc = mag_job_upload_links(a.job, p.links.data());
    else if (p.req.params.form == MAG_FORM_SPIN) rc = mag_job_generate(a.job, p.req.eps, p.req.seed);
    else rc = mag_job_generate_copy(a.job, p.req.eps, p.req.seed, p.req.copy);
  }
  if (rc == MAG_OK) rc = mag_batch_add(s->batch, a.job);
  if (rc == MAG_ERR_NOMEM && !s->active.empty()) {
    // the device is fuller than the budget says: retry once jobs leave
    if (a.job) mag_job_destroy(a.job);
    std::lock_guard<std::mutex> g(s->mu);
    s->queue.push_front(std::move(p));
    return false;
  }
  if (rc != MAG_OK) {
    mag_result r;
    std::memset(&r, 0, sizeof(r));
    r.id = p.id;
    r.status = rc;
    r.t_submit = p.t_submit;
    r.t_admit = a.t_admit;
    r.t_done = s->now();
    if (a.job) mag_job_destroy(a.job);
    publish(s, r);
    return true;
  }
  s->used += p.footprint;
  s->active.push_back(a);
  std::lock_guard<std::mutex> g(s->mu);
  s->stats.admitted++;
  if (s->used > s->stats.peak_hbm) s->stats.peak_hbm = s->used;
  if ((int)s->active.size() > s->stats.peak_active) s->stats.peak_active = (int)s->active.size();
  return true;
}

void retire_finished(mag_server* s) {
  std::vector<Active> still;
  for (const Active& a : s->active) {
    if (mag::job_done(a.job) && !mag::job_mixed32(a.job)) {
      const mag_result r = result_of(s, a, MAG_OK);
      mag_batch_remove(s->batch, a.job);
      mag_job_destroy(a.job);
      s->used -= a.footprint;
      publish(s, r);
    } else {
      still.push_back(a);
    }
  }
  s->active.swap(still);
}

// admitted requests the caller cancelled: out of the batch (which first
// lets a replay that holds the job finish), result published, memory freed
void cancel_requested(mag_server* s) {
  std::set<uint64_t> ids;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->cancel.empty()) return;
    ids.swap(s->cancel);
  }
  std::vector<Active> still;
  for (const Active& a : s->active) {
    if (ids.count(a.id)) {
      mag_batch_remove(s->batch, a.job);
      (void)mag::job_read_ctl(a.job);
      const mag_result r = result_of(s, a, MAG_ERR_CANCELLED);
      mag_job_destroy(a.job);
      s->used -= a.footprint;
      publish(s, r);
    } else {
      still.push_back(a);
    }
  }
  s->active.swap(still);
}

void fail_all_active(mag_server* s, int rc) {
  for (const Active& a : s->active) {
    const mag_result r = result_of(s, a, rc);
    mag_batch_remove(s->batch, a.job);
    mag_job_destroy(a.job);
    publish(s, r);
  }
  s->active.clear();
  s->used = 0;
}

void scheduler(mag_server* s) {
  hipSetDevice(s->device);
  for (;;) {
    {
      std::unique_lock<std::mutex> g(s->mu);
      s->cv_work.wait(g, [&] { return s->stop || !s->queue.empty() || !s->active.empty(); });
      if (s->stop && s->queue.empty() && s->active.empty()) return;
    }
    int rc = MAG_OK;
    if (!s->active.empty()) {
      // returns once some class's replay is done, with the others (and
      // that class's next replay) still running
      rc = mag_batch_step(s->batch, nullptr, nullptr);
      std::lock_guard<std::mutex> g(s->mu);
      s->stats.replays++;
    }
    while (admit_one(s)) {
    }
    if (rc != MAG_OK) {
      fail_all_active(s, rc);
      continue;
    }
    retire_finished(s);
    cancel_requested(s);
  }
}

}  // namespace

extern "C" {

int mag_server_create(int device, uint64_t hbm_budget, int max_active, mag_server** out) {
  if (!out) return fail(MAG_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (max_active < 0) return fail(MAG_ERR_ARG, "max_active must be >= 0");
  HIPCHK(hipSetDevice(device));
  if (hbm_budget == 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    hbm_budget = (uint64_t)(0.8 * (double)fr);
  }
  mag_server* s = new (std::nothrow) mag_server();
  if (!s) return fail(MAG_ERR_NOMEM, "host allocation failed");
  s->device = device;
  s->budget = hbm_budget;
  s->max_active = max_active;
  s->t0 = std::chrono::steady_clock::now();
  int rc = mag_batch_create(device, &s->batch);
  if (rc != MAG_OK) {
    delete s;
    return rc;
  }
  s->th = std::thread(scheduler, s);
  *out = s;
  return MAG_OK;
}

int mag_server_destroy(mag_server* s) {
  if (!s) return MAG_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  s->th.join();
  mag_batch_destroy(s->batch);
  delete s;
  return MAG_OK;
}

int mag_server_submit(mag_server* s, const mag_request* r, uint64_t* id) {
  if (!s || !r) return fail(MAG_ERR_ARG, "NULL argument");
  uint64_t fp = 0;
  CHK(mag_job_footprint(&r->params, &fp));
  if (fp > s->budget) return fail(MAG_ERR_NOMEM, "job footprint %llu exceeds the budget %llu",
                                  (unsigned long long)fp, (unsigned long long)s->budget);
  if (r->params.device != s->device) return fail(MAG_ERR_ARG, "request for another device");
  if (r->params.track_gauge) return fail(MAG_ERR_ARG, "served jobs do not track their gauge");
  Pending p;
  p.req = *r;
  p.req.links = nullptr;
  p.footprint = fp;
  if (r->links) {
    size_t n = 4;
    for (int i = 0; i < 4; i++) n *= (size_t)r->params.dims[i];
    n *= r->params.group == MAG_SU2 ? 4 : 18;
    p.links.assign(r->links, r->links + n);
  }
  std::lock_guard<std::mutex> g(s->mu);
  p.id = s->next_id++;
  p.t_submit = s->now();
  s->submitted++;
  if (id) *id = p.id;
  s->queue.push_back(std::move(p));
  s->cv_work.notify_all();
  return MAG_OK;
}

int mag_server_wait(mag_server* s, uint64_t id, mag_result* out) {
  if (!s || !out) return fail(MAG_ERR_ARG, "NULL argument");
  std::unique_lock<std::mutex> g(s->mu);
  if (id == 0 || id >= s->next_id) return fail(MAG_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  s->cv_done.wait(g, [&] { return s->results.count(id) > 0; });
  *out = s->results[id];
  return MAG_OK;
}

int mag_server_drain(mag_server* s) {
  if (!s) return fail(MAG_ERR_ARG, "NULL server");
  std::unique_lock<std::mutex> g(s->mu);
  s->cv_done.wait(g, [&] { return s->finished == s->submitted; });
  return MAG_OK;
}

int mag_server_cancel(mag_server* s, uint64_t id) {
  if (!s) return fail(MAG_ERR_ARG, "NULL server");
  std::lock_guard<std::mutex> g(s->mu);
  if (id == 0 || id >= s->next_id) return fail(MAG_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  if (s->results.count(id))
    return fail(MAG_ERR_STATE, "request %llu has already finished", (unsigned long long)id);
  for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
    if (it->id != id) continue;
    mag_result r;
    std::memset(&r, 0, sizeof(r));
    r.id = id;
    r.status = MAG_ERR_CANCELLED;
    r.t_submit = it->t_submit;
    r.t_admit = -1.0;  // never admitted
    r.t_done = s->now();
    s->queue.erase(it);
    s->results[id] = r;
    s->finished++;
    s->cv_done.notify_all();
    return MAG_OK;
  }
  s->cancel.insert(id);  // admitted (or being admitted): the scheduler removes it
  s->cv_work.notify_all();
  return MAG_OK;
}

int mag_server_clock(mag_server* s, double* t) {
  if (!s || !t) return fail(MAG_ERR_ARG, "NULL argument");
  *t = s->now();
  return MAG_OK;
}

int mag_server_get_stats(mag_server* s, mag_server_stats* st) {
  if (!s || !st) return fail(MAG_ERR_ARG, "NULL argument");
  std::lock_guard<std::mutex> g(s->mu);
  *st = s->stats;
  st->budget = s->budget;
  st->submitted = s->submitted;
  st->finished = s->finished;
  return MAG_OK;
}

}  // extern "C"
