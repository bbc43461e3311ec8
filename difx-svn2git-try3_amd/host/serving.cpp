// serving.cpp — the request loop of fxcorr_serve.h: per-configuration slots
// (fxg_job instances), each with a worker thread that runs one
// subintegration at a time on the slot's own stream; slots are created on
// demand under an HBM budget.
#include "fxcorr_serve.h"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  int cfg = 0;
  fxg_subint_desc sub{};
  std::vector<fxg_station_input> st;
  std::vector<std::vector<int32_t>> flags;   // copies of the valid flags
  float* results = nullptr;
  Clock::time_point arrival, done;
  int rc = 0;
  bool finished = false;
  bool waited = false;                       // a thread is in fxs_wait for it
};

struct Config {
  fxg_job_desc desc{};
  std::vector<fxg_freq> freqs;
  std::vector<fxg_datastream> dss;
  std::vector<fxg_baseline> bls;
  int max_slots = 1;
  int64_t slot_bytes = 0;
  int nslots = 0, busy = 0;
  int creating = 0;                          // slots reserved, job being created
  std::deque<Request*> queue;
};

}  // namespace

struct fxs_server {
  fxg_ctx* ctx = nullptr;
  int device = 0;
  int64_t budget = 0, used = 0;
  std::mutex m;
  std::condition_variable cv_work, cv_done, cv_gpu;
  // requests a slot has taken, waiting to start on the device: they start
  // in arrival (id) order, at most max_inflight at a time (0: no limit)
  std::set<int64_t> ready;
  int max_inflight = 0, inflight = 0;
  bool stop = false;
  std::vector<std::unique_ptr<Config>> cfgs;
  std::vector<std::thread> workers;
  std::vector<fxg_job*> jobs;
  std::map<int64_t, Request*> reqs;
  int64_t next_id = 1;
  fxs_stats_t st{};
};

namespace {

// worker of one slot: takes its configuration's requests in arrival order,
// then starts each when it is the oldest request taken by any slot and the
// device has room for one more (max_inflight)
void slot_loop(fxs_server* s, Config* c, fxg_job* job) {
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->m);
      s->cv_work.wait(lk, [&] { return s->stop || !c->queue.empty(); });
      if (c->queue.empty()) return;                 // stopping and nothing left
      r = c->queue.front();
      c->queue.pop_front();
      c->busy++;
      s->st.queued--;
      s->ready.insert(r->id);
      s->cv_gpu.wait(lk, [&] {
        return (s->max_inflight <= 0 || s->inflight < s->max_inflight) && *s->ready.begin() == r->id;
      });
      s->ready.erase(r->id);
      s->inflight++;
    }
    s->cv_gpu.notify_all();                         // the next oldest may start too
    int rc = fxg_subint_process(job, &r->sub, r->st.data(), 1, r->results, 1);
    {
      std::lock_guard<std::mutex> lk(s->m);
      r->done = Clock::now();
      r->rc = rc;
      r->finished = true;
      c->busy--;
      s->inflight--;
      s->st.completed++;
    }
    s->cv_gpu.notify_all();
    s->cv_done.notify_all();
  }
}

// Whether c should get another slot: its requests outnumber its idle and
// coming slots and the budget allows.  A yes reserves the slot's footprint;
// the caller creates the job without the lock and then publishes it (or
// gives the reservation back).  Caller holds s->m.
bool reserve_slot(fxs_server* s, Config* c) {
  const int idle = c->nslots + c->creating - c->busy;
  if ((int)c->queue.size() <= idle || c->nslots + c->creating >= c->max_slots) return false;
  if (s->used + c->slot_bytes > s->budget) {
    s->st.refused++;
    return false;
  }
  s->used += c->slot_bytes;
  s->st.used_bytes = s->used;
  c->creating++;
  return true;
}

// the reserved slot as a running worker (caller holds s->m)
void publish_slot(fxs_server* s, Config* c, fxg_job* job) {
  c->creating--;
  s->jobs.push_back(job);
  c->nslots++;
  s->st.nslots++;
  s->workers.emplace_back(slot_loop, s, c, job);
}

}  // namespace

extern "C" {

int fxs_create(int device, int64_t budget_bytes, fxs_server** out) {
  if (!out) return FXG_EINVAL;
  fxs_server* s = new fxs_server();
  s->device = device;
  int rc = fxg_ctx_create(device, &s->ctx);
  if (rc) {
    delete s;
    return rc;
  }
  if (budget_bytes <= 0) {
    int64_t fr = 0, tot = 0;
    if ((rc = fxg_ctx_memory(s->ctx, &fr, &tot))) {
      fxg_ctx_destroy(s->ctx);
      delete s;
      return rc;
    }
    budget_bytes = fr / 10 * 9;
  }
  s->budget = budget_bytes;
  s->st.budget_bytes = budget_bytes;
  *out = s;
  return 0;
}

int fxs_destroy(fxs_server* s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (std::thread& t : s->workers) t.join();   // (workers drain their queues before they stop)
  for (fxg_job* j : s->jobs) fxg_job_destroy(j);
  for (auto& kv : s->reqs) delete kv.second;
  fxg_ctx_destroy(s->ctx);
  delete s;
  return 0;
}

int fxs_add_config(fxs_server* s, const fxg_job_desc* desc, int max_slots, int* config_id) {
  if (!s || !desc || !config_id || max_slots < 1) return FXG_EINVAL;
  std::unique_ptr<Config> c(new Config());
  c->desc = *desc;
  c->freqs.assign(desc->freqs, desc->freqs + desc->nfreqs);
  c->dss.assign(desc->datastreams, desc->datastreams + desc->ndatastreams);
  if (desc->nbaselines) c->bls.assign(desc->baselines, desc->baselines + desc->nbaselines);
  c->desc.freqs = c->freqs.data();
  c->desc.datastreams = c->dss.data();
  c->desc.baselines = c->bls.empty() ? nullptr : c->bls.data();
  fxg_plan_info info{};
  int rc = fxg_plan(&c->desc, &info);
  if (rc) return rc;
  c->slot_bytes = info.device_bytes;
  c->max_slots = max_slots;
  std::lock_guard<std::mutex> lk(s->m);
  if (s->used + c->slot_bytes > s->budget) return FXG_ENOMEM;   // admission by HBM footprint
  *config_id = (int)s->cfgs.size();
  s->cfgs.push_back(std::move(c));
  s->st.nconfigs = (int)s->cfgs.size();
  return 0;
}

int fxs_submit(fxs_server* s, int config_id, const fxg_subint_desc* sub, const fxg_station_input* stations,
               float* results_device, int64_t* request) {
  if (!s || !sub || !stations || !results_device || !request) return FXG_EINVAL;
  std::unique_lock<std::mutex> lk(s->m);
  if (config_id < 0 || config_id >= (int)s->cfgs.size()) return FXG_EINVAL;
  Config* c = s->cfgs[config_id].get();
  Request* r = new Request();
  r->arrival = Clock::now();
  r->id = s->next_id++;
  r->cfg = config_id;
  r->sub = *sub;
  r->st.assign(stations, stations + c->desc.ndatastreams);
  r->flags.resize(r->st.size());
  for (size_t d = 0; d < r->st.size(); d++) {
    if (r->st[d].validflags && r->st[d].nvalidflags > 0)
      r->flags[d].assign(r->st[d].validflags, r->st[d].validflags + r->st[d].nvalidflags);
    r->st[d].validflags = r->flags[d].empty() ? nullptr : r->flags[d].data();
  }
  r->results = results_device;
  s->reqs[r->id] = r;
  c->queue.push_back(r);
  s->st.submitted++;
  s->st.queued++;
  const bool reserved = reserve_slot(s, c);
  if (!reserved && c->nslots + c->creating == 0) {   // nothing can ever run it
    c->queue.pop_back();
    s->reqs.erase(r->id);
    s->st.submitted--;
    s->st.queued--;
    delete r;
    return FXG_ENOMEM;
  }
  if (!reserved) {
    *request = r->id;                    // queued; a refused extra slot only delays it
    lk.unlock();
    s->cv_work.notify_all();
    return 0;
  }
  // the new slot's job (device allocations, table uploads) is created
  // without the server lock, so that other slots keep completing
  lk.unlock();
  fxg_job* job = nullptr;
  const int rc = fxg_job_create(s->ctx, &c->desc, &job);
  lk.lock();
  if (!rc) {
    publish_slot(s, c, job);
    *request = r->id;
    lk.unlock();
    s->cv_work.notify_all();
    return 0;
  }
  c->creating--;
  s->used -= c->slot_bytes;
  s->st.used_bytes = s->used;
  if (c->nslots + c->creating > 0) {     // another slot will take it
    *request = r->id;
    lk.unlock();
    s->cv_work.notify_all();
    return 0;
  }
  // no slot of this configuration exists: every request queued for it
  // fails with the creation error (this one is withdrawn)
  for (Request* q : c->queue) {
    s->st.queued--;
    if (q == r) continue;
    q->rc = rc;
    q->done = Clock::now();
    q->finished = true;
  }
  c->queue.clear();
  s->reqs.erase(r->id);
  s->st.submitted--;
  delete r;
  lk.unlock();
  s->cv_done.notify_all();
  return rc;
}

int fxs_wait(fxs_server* s, int64_t request, int64_t* latency_ns) {
  if (!s) return FXG_EINVAL;
  std::unique_lock<std::mutex> lk(s->m);
  auto it = s->reqs.find(request);
  if (it == s->reqs.end()) return FXG_EINVAL;
  Request* r = it->second;
  if (r->waited) return FXG_ESTATE;     // one waiter per request: it frees the request
  r->waited = true;
  s->cv_done.wait(lk, [&] { return r->finished; });
  if (latency_ns)
    *latency_ns = std::chrono::duration_cast<std::chrono::nanoseconds>(r->done - r->arrival).count();
  int rc = r->rc;
  s->reqs.erase(it);
  delete r;
  return rc;
}

int fxs_set_max_inflight(fxs_server* s, int n) {
  if (!s || n < 0) return FXG_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->max_inflight = n;
  }
  s->cv_gpu.notify_all();
  return 0;
}

int fxs_stats(fxs_server* s, fxs_stats_t* out) {
  if (!s || !out) return FXG_EINVAL;
  std::lock_guard<std::mutex> lk(s->m);
  *out = s->st;
  return 0;
}

}  // extern "C"
