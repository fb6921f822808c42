// Serving PCG-OLE expansions with continuous batching on one GPU.
//
// Clients submit expansion requests (N, c, t, w, PRG, seed, pub_seed) at any
// time.  A worker thread keeps a pool of expansion plans, each with a context
// of its own (NTT tables and scratch, all allocated when the plan is built),
// and runs at most `slots` requests at once, one per slot stream: a running
// plan's context borrows the stream of its slot.  The server runs requests on
// exactly `slots` streams, so that with slots <= the device's hardware queues
// (GPU_MAX_HW_QUEUES, 4 by default) every running request has a hardware
// queue of its own and never waits in order behind another one's kernels; one
// more stream builds the plans that pcgsq_server_reserve asks for.  Waiting
// requests are admitted oldest first, except that a request whose shape has
// no idle plan is passed over (requests of one shape keep their order): a
// slot going free is taken by the oldest request that can start now.  A
// shape with no plan at all gets one built (idle plans of other shapes are
// evicted, least recently used first, if the device-memory budget needs it);
// more plans of a shape, for requests of one shape to overlap, come from
// pcgsq_server_reserve.  On admission the seed is dealt on the device, the
// expansion is enqueued on the plan's stream and the outputs are copied into
// a device buffer of the request's own, followed by an event that the worker
// polls.  Output buffers are pooled by size, so that in the steady state
// nothing is allocated or freed: a device allocation or free can wait for
// every stream's work and would stall the other requests.  Requests on
// different plans overlap on the device (the streams run concurrently).
// Results do not depend on the plan or on what else runs: every request's
// work is a pure function of its parameters and seeds.
//
// No counterpart in the reference (its benchmarks run one expansion per
// process: CPU_baseline/test/bench_pcg_ole_2pc.cpp).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <list>
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
  void* out = nullptr;   // pooled device buffer [x0, z0, x1, z1][N]
  size_t out_bytes = 0;
  int readers = 0;       // outputs() copies in progress
};

bool same_shape(const Request& a, const Request& b) {
  return same_params(a.p, b.p) && a.pub_seed == b.pub_seed;
}

struct PlanEntry {
  pcgsq_ctx* ctx = nullptr;
  pcgsq_ole_plan* plan = nullptr;
  pcgsq_params pp{};
  uint64_t pub_seed = 0;
  size_t bytes = 0;
  Request* req = nullptr;
  int slot = -1;          // slot stream while running
  hipEvent_t ev = nullptr;
  double last_used = 0;
};

void destroy_entry(PlanEntry& e) {
  if (e.plan) (void)pcgsq_ole_plan_destroy(e.plan);
  if (e.ev) (void)hipEventDestroy(e.ev);
  if (e.ctx) (void)pcgsq_ctx_destroy(e.ctx);
  e = PlanEntry{};
}

// A plan of shape p on a context of its own that borrows stream `st` (no
// server state touched).
std::unique_ptr<PlanEntry> build_entry(int device, const pcgsq_params& p, uint64_t pub_seed,
                                       size_t plan_budget, hipStream_t st) {
  std::unique_ptr<PlanEntry> e(new PlanEntry);
  try {
    int rc = pcgsq_ctx_create(device, &e->ctx);
    if (rc != PCGSQ_OK) throw Error(rc, pcgsq_last_error());
    hip_check(hipStreamDestroy(e->ctx->stream), "server stream");
    e->ctx->stream = st;
    e->ctx->owns_stream = false;
    rc = pcgsq_ole_plan_create(e->ctx, &p, pub_seed, plan_budget, &e->plan);
    if (rc != PCGSQ_OK) throw Error(rc, pcgsq_last_error());
    hip_check(hipSetDevice(device), "server hipSetDevice");
    hip_check(hipEventCreateWithFlags(&e->ev, hipEventDisableTiming), "server event");
  } catch (...) {
    destroy_entry(*e);
    throw;
  }
  e->pp = p;
  e->pub_seed = pub_seed;
  return e;
}

}  // namespace
}  // namespace pcgsq

using namespace pcgsq;

struct pcgsq_server {
  int device = 0, max_running = 1;
  size_t budget = 0, used = 0, peak = 0, plan_cap = 0;
  std::vector<std::unique_ptr<PlanEntry>> pool;
  std::vector<hipStream_t> streams;        // one per slot
  hipStream_t build_stream = nullptr;      // plans built by reserve (never runs requests)
  std::vector<char> slot_busy;
  std::multimap<size_t, void*> free_out;   // idle output buffers by size (counted in used)
  std::list<Request*> queue;               // list: iterators survive submits while unlocked
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::mutex mu;
  std::condition_variable cv;
  std::thread worker;
  bool stop = false;
  uint64_t next_id = 1;
  uint64_t window = 0;     // most ring dimension (sum of N) running at once; 0 = no limit
  uint64_t exclusive_N = 0;  // requests with N >= this run one at a time, in arrival order
  double age_bound = 0;      // seconds: a request waiting longer is passed by no later one
  long completed = 0, plans_built = 0, plans_evicted = 0;

  size_t plan_budget() const { return plan_cap; }

  int running() const {
    int n = 0;
    for (char b : slot_busy) n += b != 0;
    return n;
  }

  int free_slot() const {
    for (int k = 0; k < (int)slot_busy.size(); ++k)
      if (!slot_busy[(size_t)k]) return k;
    return -1;
  }

  static bool fits(const PlanEntry& e, const Request& r) {
    return same_params(e.pp, r.p) && e.pub_seed == r.pub_seed;
  }

  void charge(size_t b) {
    used += b;
    peak = std::max(peak, used);
  }

  // Free pooled output buffers, then idle plans (least recently used first),
  // until `need` more bytes fit the budget (lock held).  Freeing waits for the
  // device, so it happens only under memory pressure.
  bool make_room(size_t need) {
    while (used + need > budget && !free_out.empty()) {
      auto it = std::prev(free_out.end());
      (void)hipSetDevice(device);
      (void)hipFree(it->second);
      used -= it->first;
      free_out.erase(it);
    }
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
      ++plans_evicted;
    }
    return true;
  }

  // An output buffer of `bytes`: a pooled one, else a new one if the budget
  // holds it (lock held); nullptr if memory does not allow.
  void* take_out(size_t bytes) {
    auto it = free_out.find(bytes);
    if (it != free_out.end()) {
      void* p = it->second;
      free_out.erase(it);
      return p;
    }
    if (!make_room(bytes)) return nullptr;
    void* p = nullptr;
    hip_check(hipSetDevice(device), "server hipSetDevice");
    hip_check(hipMalloc(&p, bytes), "hipMalloc served outputs");
    charge(bytes);
    return p;
  }

  void put_out(Request* r) {
    if (r->out) free_out.emplace(r->out_bytes, r->out);
    r->out = nullptr;
  }

  // A new plan of this shape for reserve: room is made and charged under the
  // lock, the plan is built with the lock released, on the build stream (no
  // request runs there, so the build's synchronization waits only for
  // itself, and the worker keeps polling and admitting meanwhile).
  void reserve_one(std::unique_lock<std::mutex>& lk, const pcgsq_params& p, uint64_t pub_seed) {
    const size_t est = pcgsq_ole_plan_estimate(&p, plan_budget());
    if (est == 0) throw Error(kErrInvalid, "server_reserve: invalid expansion parameters");
    if (!make_room(est))
      throw Error(kErrNoMemory, "server_reserve: the plans do not fit the memory budget");
    charge(est);
    std::unique_ptr<PlanEntry> e;
    std::string err;
    lk.unlock();
    try {
      e = build_entry(device, p, pub_seed, plan_budget(), build_stream);
    } catch (const std::exception& ex) {
      err = ex.what();
    }
    lk.lock();
    if (!e) {
      used -= est;
      throw Error(kErrDevice, "server_reserve: " + err);
    }
    e->bytes = est;
    e->last_used = now_s();
    pool.push_back(std::move(e));
    ++plans_built;
    cv.notify_all();
  }

  void fail(Request* r, const std::string& why) {
    r->state = PCGSQ_REQ_FAILED;
    r->error = why;
    put_out(r);
  }

  // Start r on the idle plan e in free slot k (lock held).
  void start(Request* r, PlanEntry* e, int k) {
    r->t_start = now_s();
    r->state = PCGSQ_REQ_RUNNING;
    e->last_used = r->t_start;
    e->req = r;
    e->slot = k;
    slot_busy[(size_t)k] = 1;
    e->ctx->stream = streams[(size_t)k];
    try {
      CtxGuard g(e->ctx);
      if (pcgsq_ole_deal(e->plan, r->seed) != PCGSQ_OK || pcgsq_ole_run(e->plan) != PCGSQ_OK)
        throw Error(kErrDevice, pcgsq_last_error());
      hip_check(hipMemcpyAsync(r->out, pcgsq_ole_outputs_dev(e->plan), r->out_bytes,
                               hipMemcpyDeviceToDevice, e->ctx->stream),
                "served outputs");
      hip_check(hipEventRecord(e->ev, e->ctx->stream), "served event");
    } catch (const std::exception& ex) {
      e->req = nullptr;
      e->slot = -1;
      slot_busy[(size_t)k] = 0;
      fail(r, ex.what());
      ++completed;
    }
  }

  // One admission pass over the queue (lock held on entry and exit; released
  // while a plan is built).  Returns whether anything changed.
  bool admit(std::unique_lock<std::mutex>& lk) {
    bool changed = false;
    std::vector<const Request*> passed;  // shapes passed over: later ones of a shape wait too
    uint64_t run_n = 0;                  // ring dimension of the running requests
    bool excl_running = false;           // an exclusive-class request runs
    for (const auto& x : pool)
      if (x->req) {
        run_n += (uint64_t)x->req->p.N;
        excl_running = excl_running || (exclusive_N && (uint64_t)x->req->p.N >= exclusive_N);
      }
    uint64_t held_n = ~0ULL;  // smallest N held back by the window: no larger one passes it
    bool excl_waiting = false;  // an exclusive-class request waits: no later one of the class starts
    const double now = now_s();
    auto it = queue.begin();
    while (it != queue.end() && running() < max_running && !stop) {
      Request* r = *it;
      const uint64_t n = (uint64_t)r->p.N;
      const bool excl = exclusive_N && n >= exclusive_N;
      bool wait = n >= held_n;
      for (const Request* q : passed) wait = wait || same_shape(*q, *r);
      if (!wait && window && run_n > 0 && run_n + n > window) {
        held_n = n;
        wait = true;
      }
      if (!wait && excl && (excl_running || excl_waiting)) wait = true;
      if (excl && wait) excl_waiting = true;
      // an aged request that cannot start stops every later one from passing it
      const bool aged = age_bound > 0 && now - r->t_submit > age_bound;
      if (wait) {
        if (aged) break;
        ++it;
        continue;
      }
      PlanEntry* e = nullptr;
      bool have_shape = false;
      for (auto& x : pool)
        if (fits(*x, *r)) {
          have_shape = true;
          if (!x->req) {
            e = x.get();
            break;
          }
        }
      const int k = free_slot();
      if (!e && !have_shape) {
        // no plan of this shape: build one, outside the lock, on the slot
        // stream it will run on (idle, so its synchronization waits for nothing)
        const size_t est = pcgsq_ole_plan_estimate(&r->p, plan_budget());
        if (!make_room(est)) {
          passed.push_back(r);
          ++it;
          continue;
        }
        charge(est);
        std::unique_ptr<PlanEntry> ne;
        std::string err;
        const pcgsq_params pp = r->p;
        const uint64_t ps = r->pub_seed;
        slot_busy[(size_t)k] = 1;   // held through the build
        lk.unlock();
        try {
          ne = build_entry(device, pp, ps, plan_budget(), streams[(size_t)k]);
        } catch (const std::exception& ex) {
          err = ex.what();
        }
        lk.lock();
        slot_busy[(size_t)k] = 0;
        changed = true;
        if (!ne) {
          used -= est;
          it = queue.erase(it);
          fail(r, err);
          ++completed;
          continue;
        }
        ne->bytes = est;
        pool.push_back(std::move(ne));
        ++plans_built;
        e = pool.back().get();
      }
      if (!e) {  // its shape's plans are busy: wait for one
        if (aged) break;
        if (excl) excl_waiting = true;
        passed.push_back(r);
        ++it;
        continue;
      }
      e->req = r;  // keeps e from eviction while the outputs find room
      void* ob = nullptr;
      try {
        ob = take_out(r->out_bytes);
      } catch (const std::exception& ex) {
        e->req = nullptr;
        it = queue.erase(it);
        fail(r, ex.what());
        ++completed;
        changed = true;
        continue;
      }
      e->req = nullptr;
      if (!ob) {
        passed.push_back(r);
        ++it;
        continue;
      }
      r->out = ob;
      it = queue.erase(it);
      start(r, e, k);
      if (e->req) {
        run_n += n;
        excl_running = excl_running || excl;
      }
      changed = true;
    }
    return changed;
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
        if (q == hipSuccess)
          r->state = PCGSQ_REQ_DONE;
        else
          fail(r, hipGetErrorString(q));
        e->req = nullptr;
        slot_busy[(size_t)e->slot] = 0;
        e->slot = -1;
        ++completed;
        progressed = true;
      }
      progressed = admit(lk) || progressed;
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
    s->streams.assign((size_t)slots, nullptr);
    s->slot_busy.assign((size_t)slots, 0);
    for (auto& st : s->streams)
      hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "server stream");
    hip_check(hipStreamCreateWithFlags(&s->build_stream, hipStreamNonBlocking), "server stream");
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
    (void)hipSetDevice(s->device);
    for (hipStream_t st : s->streams)
      if (st) (void)hipStreamSynchronize(st);
    for (auto& e : s->pool) destroy_entry(*e);
    s->pool.clear();
    for (hipStream_t st : s->streams)
      if (st) (void)hipStreamDestroy(st);
    if (s->build_stream) (void)hipStreamDestroy(s->build_stream);
    for (auto& kv : s->reqs)
      if (kv.second->out) (void)hipFree(kv.second->out);
    for (auto& kv : s->free_out) (void)hipFree(kv.second);
    s->reqs.clear();
    s->free_out.clear();
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
    Request* r = nullptr;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->reqs.find(id);
      require(it != s->reqs.end(), "server_outputs: unknown request");
      r = it->second.get();
      require(r->state == PCGSQ_REQ_DONE, "server_outputs: request not done");
      ++r->readers;  // release() refuses while the copy runs
    }
    // copy without the server lock: the worker keeps admitting meanwhile
    struct Done {
      pcgsq_server* s;
      Request* r;
      ~Done() {
        std::lock_guard<std::mutex> lk(s->mu);
        --r->readers;
      }
    } done{s, r};
    hip_check(hipSetDevice(s->device), "server hipSetDevice");
    const size_t N = (size_t)r->p.N;
    const uint64_t* src = static_cast<const uint64_t*>(r->out);
    uint64_t* dst[4] = {x0, z0, x1, z1};
    for (int i = 0; i < 4; ++i)
      if (dst[i])
        hip_check(hipMemcpy(dst[i], src + i * N, N * 8, hipMemcpyDeviceToHost), "server outputs");
  });
}

int pcgsq_server_outputs_dev(pcgsq_server* s, uint64_t id, const uint64_t** out) {
  return guarded([&] {
    require(s && out, "server_outputs_dev: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    require(it != s->reqs.end(), "server_outputs_dev: unknown request");
    require(it->second->state == PCGSQ_REQ_DONE, "server_outputs_dev: request not done");
    *out = static_cast<const uint64_t*>(it->second->out);
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
    require(r->readers == 0, "server_release: outputs are being copied");
    s->put_out(r);  // the buffer goes back to the pool, still charged
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
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
      int have = 0;
      for (auto& e : s->pool) have += same_params(e->pp, *p) && e->pub_seed == pub_seed;
      if (have >= count) break;
      s->reserve_one(lk, *p, pub_seed);
    }
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

int pcgsq_server_set_window(pcgsq_server* s, uint64_t max_running_N) {
  return guarded([&] {
    require(s, "server_set_window: null server");
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->window = max_running_N;
    }
    s->cv.notify_all();
  });
}

int pcgsq_server_set_policy(pcgsq_server* s, uint64_t exclusive_N, double age_bound_s) {
  return guarded([&] {
    require(s, "server_set_policy: null server");
    require(age_bound_s >= 0, "server_set_policy: age bound must be >= 0");
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->exclusive_N = exclusive_N;
      s->age_bound = age_bound_s;
    }
    s->cv.notify_all();
  });
}

int pcgsq_server_plans(pcgsq_server* s, int* plans, long* built, long* evicted) {
  return guarded([&] {
    require(s, "server_plans: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (plans) *plans = (int)s->pool.size();
    if (built) *built = s->plans_built;
    if (evicted) *evicted = s->plans_evicted;
  });
}

double pcgsq_server_clock(void) { return now_s(); }

}  // extern "C"
