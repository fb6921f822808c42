// Serving PCG-OLE expansions with continuous batching on one GPU.
//
// Clients submit expansion requests (N, c, t, w, PRG, seed, pub_seed) at any
// time.  A worker thread keeps a pool of expansion plans, each bound to a
// context of its own (a stream, NTT tables and scratch), and runs at most
// `slots` requests at once.  The oldest waiting request is admitted as soon
// as fewer than `slots` run and either an idle plan of its shape exists or
// the device-memory budget holds a new one (idle plans of other shapes are
// evicted, least recently used first): its seed is dealt on the device, the
// expansion is enqueued on the plan's stream and the outputs are copied into
// the request's own device buffer, followed by an event.  Requests on
// different plans run concurrently on the device (the streams overlap); a
// request arriving while others run is admitted the moment a slot and memory
// free up, without waiting for the others to finish.  Completion is the
// event, polled by the worker; outputs stay on the device until the client
// copies or releases them.  Results do not depend on the plan or on what else
// runs: every request's work is a pure function of its parameters and seeds.
//
// No counterpart in the reference (its benchmarks run one expansion per
// process: CPU_baseline/test/bench_pcg_ole_2pc.cpp).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "context.hpp"
#include "pcgsq.h"

namespace pcgsq {
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

bool same_params(const pcgsq_params& a, const pcgsq_params& b) {
  return a.N == b.N && a.c == b.c && a.t == b.t && a.w == b.w && a.prg == b.prg;
}

struct Request {
  uint64_t id = 0;
  pcgsq_params p{};
  uint64_t seed = 0, pub_seed = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  int state = PCGSQ_REQ_QUEUED;
  std::string error;
  DevBuf out;            // [x0, z0, x1, z1][N]
  size_t out_bytes = 0;
  int slot = -1;
};

struct PlanEntry {
  pcgsq_ctx* ctx = nullptr;
  pcgsq_ole_plan* plan = nullptr;
  pcgsq_params pp{};
  uint64_t pub_seed = 0;
  size_t bytes = 0;
  Request* req = nullptr;
  hipEvent_t ev = nullptr;
  double last_used = 0;
};

void destroy_entry(PlanEntry& e) {
  if (e.plan) (void)pcgsq_ole_plan_destroy(e.plan);
  if (e.ev) (void)hipEventDestroy(e.ev);
  if (e.ctx) (void)pcgsq_ctx_destroy(e.ctx);
  e = PlanEntry{};
}

}  // namespace
}  // namespace pcgsq

using namespace pcgsq;

struct pcgsq_server {
  int device = 0, max_running = 1;
  size_t budget = 0, used = 0, peak = 0, plan_cap = 0;
  std::vector<std::unique_ptr<PlanEntry>> pool;
  std::deque<Request*> queue;
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::mutex mu;
  std::condition_variable cv;
  std::thread worker;
  bool stop = false;
  uint64_t next_id = 1;
  long completed = 0;

  size_t plan_budget() const { return plan_cap; }

  int running() const {
    int n = 0;
    for (const auto& e : pool) n += e->req != nullptr;
    return n;
  }

  static bool fits(const PlanEntry& e, const Request& r) {
    return same_params(e.pp, r.p) && e.pub_seed == r.pub_seed;
  }

  // Evict idle plans, least recently used first, until `need` more bytes fit.
  bool make_room(size_t need) {
    while (used + need > budget) {
      int victim = -1;
      for (int i = 0; i < (int)pool.size(); ++i)
        if (!pool[(size_t)i]->req &&
            (victim < 0 || pool[(size_t)i]->last_used < pool[(size_t)victim]->last_used))
          victim = i;
      if (victim < 0) return false;
      used -= pool[(size_t)victim]->bytes;
      destroy_entry(*pool[(size_t)victim]);
      pool.erase(pool.begin() + victim);
    }
    return true;
  }

  // A new plan of r's shape (lock held); nullptr if memory does not allow.
  PlanEntry* new_entry(const pcgsq_params& p, uint64_t pub_seed, size_t extra) {
    const size_t est = pcgsq_ole_plan_estimate(&p, plan_budget());
    if (est == 0) throw Error(kErrInvalid, "invalid expansion parameters");
    if (!make_room(est + extra)) return nullptr;
    std::unique_ptr<PlanEntry> e(new PlanEntry);
    try {
      int rc = pcgsq_ctx_create(device, &e->ctx);
      if (rc != PCGSQ_OK) throw Error(rc, pcgsq_last_error());
      rc = pcgsq_ole_plan_create(e->ctx, &p, pub_seed, plan_budget(), &e->plan);
      if (rc != PCGSQ_OK) throw Error(rc, pcgsq_last_error());
      hip_check(hipSetDevice(device), "server hipSetDevice");
      hip_check(hipEventCreateWithFlags(&e->ev, hipEventDisableTiming), "server event");
    } catch (...) {
      destroy_entry(*e);
      throw;
    }
    e->pp = p;
    e->pub_seed = pub_seed;
    e->bytes = est;
    used += est;
    peak = std::max(peak, used);
    pool.push_back(std::move(e));
    return pool.back().get();
  }

  // Admit the oldest request if a slot and the memory budget allow (lock held).
  bool admit_one() {
    if (queue.empty() || running() >= max_running) return false;
    Request* r = queue.front();
    PlanEntry* e = nullptr;
    for (auto& x : pool)
      if (!x->req && fits(*x, *r)) {
        e = x.get();
        break;
      }
    try {
      if (!e) {
        e = new_entry(r->p, r->pub_seed, r->out_bytes);
        if (!e) return false;  // wait for memory
      } else if (!make_room(r->out_bytes)) {
        return false;
      }
    } catch (const std::exception& ex) {
      queue.pop_front();
      r->state = PCGSQ_REQ_FAILED;
      r->error = ex.what();
      return true;
    }
    queue.pop_front();
    r->t_start = now_s();
    r->state = PCGSQ_REQ_RUNNING;
    used += r->out_bytes;
    peak = std::max(peak, used);
    e->last_used = r->t_start;
    try {
      CtxGuard g(e->ctx);
      r->out.ensure(r->out_bytes, "hipMalloc served outputs");
      if (pcgsq_ole_deal(e->plan, r->seed) != PCGSQ_OK || pcgsq_ole_run(e->plan) != PCGSQ_OK)
        throw Error(kErrDevice, pcgsq_last_error());
      hip_check(hipMemcpyAsync(r->out.ptr, pcgsq_ole_outputs_dev(e->plan), r->out_bytes,
                               hipMemcpyDeviceToDevice, e->ctx->stream),
                "served outputs");
      hip_check(hipEventRecord(e->ev, e->ctx->stream), "served event");
      e->req = r;
    } catch (const std::exception& ex) {
      r->state = PCGSQ_REQ_FAILED;
      r->error = ex.what();
      used -= r->out_bytes;
      r->out.release();
    }
    return true;
  }

  void run() {
    std::unique_lock<std::mutex> lk(mu);
    while (!stop) {
      bool progressed = false;
      for (auto& e : pool) {
        if (!e->req) continue;
        const hipError_t q = hipEventQuery(e->ev);
        if (q == hipErrorNotReady) continue;
        Request* r = e->req;
        r->t_done = now_s();
        e->last_used = r->t_done;
        if (q == hipSuccess) {
          r->state = PCGSQ_REQ_DONE;
        } else {
          r->state = PCGSQ_REQ_FAILED;
          r->error = hipGetErrorString(q);
          used -= r->out_bytes;
          r->out.release();
        }
        e->req = nullptr;
        ++completed;
        progressed = true;
      }
      while (admit_one()) progressed = true;
      if (progressed) cv.notify_all();
      // poll completions every 50 us while work runs; a waiting request with
      // nothing running waits for memory (a release); otherwise sleep until a
      // submit
      if (running() > 0)
        cv.wait_for(lk, std::chrono::microseconds(50));
      else if (!queue.empty())
        cv.wait_for(lk, std::chrono::milliseconds(1));
      else
        cv.wait(lk, [&] { return stop || !queue.empty(); });
    }
  }
};

extern "C" {

int pcgsq_server_create(int device, size_t hbm_budget_bytes, int slots, pcgsq_server** out) {
  return guarded([&] {
    require(out, "server_create: null out");
    require(slots >= 1 && slots <= 16, "server_create: slots must be 1..16");
    std::unique_ptr<pcgsq_server> s(new pcgsq_server);
    s->device = device;
    size_t free_b = 0, total_b = 0;
    hip_check(hipSetDevice(device), "server hipSetDevice");
    hip_check(hipMemGetInfo(&free_b, &total_b), "server hipMemGetInfo");
    s->budget = hbm_budget_bytes ? hbm_budget_bytes : free_b / 10 * 7;
    require(s->budget <= free_b, "server_create: budget exceeds free device memory", kErrNoMemory);
    // each plan keeps its DPF frontier within a share of the budget
    s->plan_cap = std::max<size_t>((size_t)1 << 30, s->budget / (size_t)slots / 2);
    s->max_running = slots;
    s->worker = std::thread([p = s.get()] {
      (void)hipSetDevice(p->device);
      p->run();
    });
    *out = s.release();
  });
}

int pcgsq_server_destroy(pcgsq_server* s) {
  if (!s) return PCGSQ_OK;
  return guarded([&] {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto& e : s->pool) {
      if (e->ctx) (void)pcgsq_ctx_synchronize(e->ctx);
      destroy_entry(*e);
    }
    s->pool.clear();
    s->reqs.clear();
    delete s;
  });
}

int pcgsq_server_submit(pcgsq_server* s, const pcgsq_params* p, uint64_t seed, uint64_t pub_seed,
                        uint64_t* id) {
  return guarded([&] {
    require(s && p && id, "server_submit: null argument");
    std::unique_ptr<Request> r(new Request);
    r->p = *p;
    r->seed = seed;
    r->pub_seed = pub_seed;
    r->t_submit = now_s();
    r->out_bytes = (size_t)4 * (size_t)p->N * sizeof(uint64_t);
    const size_t est = pcgsq_ole_plan_estimate(p, s->plan_budget());
    require(est != 0, "server_submit: bad parameters");
    require(est + r->out_bytes <= s->budget,
            "server_submit: the request does not fit the server's memory budget", kErrNoMemory);
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    *id = r->id;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
    s->cv.notify_all();
  });
}

int pcgsq_server_poll(pcgsq_server* s, uint64_t id, int* state) {
  return guarded([&] {
    require(s && state, "server_poll: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    require(it != s->reqs.end(), "server_poll: unknown request");
    *state = it->second->state;
  });
}

int pcgsq_server_wait(pcgsq_server* s, uint64_t id, double timeout_s) {
  return guarded([&] {
    require(s, "server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    require(it != s->reqs.end(), "server_wait: unknown request");
    Request* r = it->second.get();
    const auto pred = [&] { return r->state == PCGSQ_REQ_DONE || r->state == PCGSQ_REQ_FAILED; };
    if (timeout_s > 0) {
      require(s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred),
              "server_wait: timed out", kErrDevice);
    } else {
      s->cv.wait(lk, pred);
    }
    require(r->state == PCGSQ_REQ_DONE, "request failed: " + r->error, kErrDevice);
  });
}

int pcgsq_server_outputs(pcgsq_server* s, uint64_t id, uint64_t* x0, uint64_t* z0, uint64_t* x1,
                         uint64_t* z1) {
  return guarded([&] {
    require(s, "server_outputs: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    require(it != s->reqs.end(), "server_outputs: unknown request");
    Request* r = it->second.get();
    require(r->state == PCGSQ_REQ_DONE, "server_outputs: request not done");
    hip_check(hipSetDevice(s->device), "server hipSetDevice");
    const size_t N = (size_t)r->p.N;
    uint64_t* dst[4] = {x0, z0, x1, z1};
    for (int i = 0; i < 4; ++i)
      if (dst[i])
        hip_check(hipMemcpy(dst[i], r->out.as<uint64_t>() + i * N, N * 8, hipMemcpyDeviceToHost),
                  "server outputs");
  });
}

int pcgsq_server_outputs_dev(pcgsq_server* s, uint64_t id, const uint64_t** out) {
  return guarded([&] {
    require(s && out, "server_outputs_dev: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    require(it != s->reqs.end(), "server_outputs_dev: unknown request");
    require(it->second->state == PCGSQ_REQ_DONE, "server_outputs_dev: request not done");
    *out = it->second->out.as<uint64_t>();
  });
}

int pcgsq_server_release(pcgsq_server* s, uint64_t id) {
  return guarded([&] {
    require(s, "server_release: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    require(it != s->reqs.end(), "server_release: unknown request");
    Request* r = it->second.get();
    require(r->state == PCGSQ_REQ_DONE || r->state == PCGSQ_REQ_FAILED,
            "server_release: request still queued or running");
    if (r->state == PCGSQ_REQ_DONE) s->used -= r->out_bytes;
    s->reqs.erase(it);
    s->cv.notify_all();
  });
}

int pcgsq_server_timing(pcgsq_server* s, uint64_t id, double* t_submit, double* t_start,
                        double* t_done) {
  return guarded([&] {
    require(s, "server_timing: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    require(it != s->reqs.end(), "server_timing: unknown request");
    if (t_submit) *t_submit = it->second->t_submit;
    if (t_start) *t_start = it->second->t_start;
    if (t_done) *t_done = it->second->t_done;
  });
}

int pcgsq_server_stats(pcgsq_server* s, size_t* budget, size_t* used, size_t* peak,
                       long* completed, int* queued) {
  return guarded([&] {
    require(s, "server_stats: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (budget) *budget = s->budget;
    if (used) *used = s->used;
    if (peak) *peak = s->peak;
    if (completed) *completed = s->completed;
    if (queued) *queued = (int)s->queue.size();
  });
}

int pcgsq_server_reserve(pcgsq_server* s, const pcgsq_params* p, uint64_t pub_seed, int count) {
  return guarded([&] {
    require(s && p, "server_reserve: null argument");
    require(count >= 0 && count <= 16, "server_reserve: count must be 0..16");
    std::lock_guard<std::mutex> lk(s->mu);
    int have = 0;
    for (auto& e : s->pool)
      have += same_params(e->pp, *p) && e->pub_seed == pub_seed;
    for (; have < count; ++have)
      require(s->new_entry(*p, pub_seed, 0) != nullptr,
              "server_reserve: the plans do not fit the memory budget", kErrNoMemory);
  });
}

int pcgsq_server_charge(pcgsq_server* s, const pcgsq_params* p, size_t* bytes) {
  return guarded([&] {
    require(s && p && bytes, "server_charge: null argument");
    const size_t est = pcgsq_ole_plan_estimate(p, s->plan_budget());
    require(est != 0, "server_charge: bad parameters");
    *bytes = est + (size_t)4 * (size_t)p->N * sizeof(uint64_t);
  });
}

double pcgsq_server_clock(void) { return now_s(); }

}  // extern "C"
