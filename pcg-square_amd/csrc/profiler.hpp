// Per-launch device timing on the context stream (the build's counterpart of
// the reference's phase-wall timers and PCG_DPF_LEVEL_MS / PCG_SQUARE_STAGE_MS
// splits, dpf_gpu.cu:427-493, poly_mul_gpuntt_square.cu:105-154).
// When enabled, every instrumented launch records a hipEvent pair on the
// stream; totals per zone name are resolved lazily at read time.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace pcgsq {

struct Profiler {
  bool enabled = false;
  struct Pending {
    std::string zone;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  struct Acc {
    long count = 0;
    double ms = 0;
  };
  std::map<std::string, Acc> totals;

  // Returns an open record (events recorded) or -1 when disabled.
  int begin(hipStream_t st, const char* zone) {
    if (!enabled) return -1;
    Pending p{zone, nullptr, nullptr};
    if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) return -1;
    (void)hipEventRecord(p.a, st);
    pending.push_back(p);
    return (int)pending.size() - 1;
  }
  void end(hipStream_t st, int id) {
    if (id < 0) return;
    (void)hipEventRecord(pending[(size_t)id].b, st);
  }
  void resolve() {
    for (auto& p : pending) {
      (void)hipEventSynchronize(p.b);
      float ms = 0;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        auto& t = totals[p.zone];
        t.count += 1;
        t.ms += ms;
      }
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending.clear();
  }
  void reset() {
    resolve();
    totals.clear();
  }
};

// RAII zone: records around whatever is enqueued in its lifetime.
struct ProfZone {
  Profiler* p;
  hipStream_t st;
  int id;
  ProfZone(Profiler* prof, hipStream_t s, const char* zone)
      : p(prof), st(s), id(prof ? prof->begin(s, zone) : -1) {}
  ~ProfZone() {
    if (p) p->end(st, id);
  }
};

}  // namespace pcgsq
