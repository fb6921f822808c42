// Context object behind the C ABI: one device, one stream, owned scratch.
#pragma once

#include <mutex>

#include "convert.hpp"
#include "dpf.hpp"
#include "internal.hpp"
#include "ntt.hpp"
#include "profiler.hpp"

struct pcgsq_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = true;   // false: the stream is lent (a server's slot stream)
  std::recursive_mutex mu;
  pcgsq::DpfScratch dpf;
  pcgsq::NttEngine ntt;
  pcgsq::DevBuf in_a, in_b, in_c, in_d, leaves, part, sums, cwbuf, gbuf, t0, rks;
  pcgsq::Profiler prof;
  // device address of this device's copy of the library's no-allocation
  // scratch array (capi.hip), taken when the context is created
  uint64_t* noalloc_scratch = nullptr;

  void activate() { pcgsq::hip_check(hipSetDevice(device), "hipSetDevice"); }
};

namespace pcgsq {

// Scoped device activation + context lock for an API call.
struct CtxGuard {
  std::lock_guard<std::recursive_mutex> lk;
  explicit CtxGuard(pcgsq_ctx* c) : lk(c->mu) { c->activate(); }
};

// Host dealer (seeded stand-in for the reference's noise sampling and 2PC
// preprocessing; the spec is written out in ole.cpp).
uint64_t dealer_u64(uint64_t seed, uint32_t domain, uint64_t a, uint64_t b);

}  // namespace pcgsq
