// Negacyclic NTT over F_p, P = 2^62 - 3*2^25 + 1, on gfx950 (host side).
//
// Semantics are those of CPU_baseline/ntt.h:
//   forward(a)[k] = a(psi^(2k+1)) in natural order (psi = 3^((P-1)/2N)),
//   inverse = forward^-1, poly_mul(a, b) = a*b mod (X^N + 1).
// Two device backends compute exactly that:
//   * kFourStep: for N <= 2^13 one pass per transform (the whole transform in
//     one workgroup tile, one HBM round trip).  Larger N split as n1*n2 with
//     n1 = 2^ceil(L/2), n2 = 2^floor(L/2) by default: pass 1 runs length-n1
//     negacyclic NTTs down the columns (root psi^n2, twist merged into the
//     butterflies), multiplies by psi^(i2(2k1+1)) and writes back; pass 2
//     runs length-n2 cyclic NTTs along the rows and writes the result
//     transposed, i.e. in natural order (two HBM round trips).  Passes hold
//     their data in registers and exchange through LDS between register
//     groups (ntt_pass.hpp).  The inverse mirrors it; N^-1 and the pointwise
//     product of a poly-mul are folded into the first inverse pass.
//   * kNaive: the textbook per-stage radix-2 form (one launch per butterfly
//     stage: psi-twist, DIF forward, DIT inverse), the structurally
//     independent cross-check of the reference's naive tier
//     (GPU_baseline/naive_ntt_bench.cu:46-88,147-159).
#pragma once

#include <stdint.h>

#include <memory>

#include "internal.hpp"
#include "profiler.hpp"

namespace pcgsq {

enum NttBackend : int { kNttFourStep = 0, kNttNaive = 1 };

// Longest transform run as a single pass (2^13 words = 64 KiB of LDS).
constexpr int kNttSingleMax = 13;

struct NttPlan;

// Cached per-(N) plans (tables live on the device of the calling context).
class NttEngine {
 public:
  NttEngine();
  ~NttEngine();
  static bool supported(int N, int backend, std::string* reason);
  // In-place batched transforms on device buffers of batch*N words.
  void forward(hipStream_t st, uint64_t* d, int batch, int N, int backend);
  void inverse(hipStream_t st, uint64_t* d, int batch, int N, int backend);
  // out = a*b mod X^N+1 (device buffers; out may alias a or b).
  void poly_mul(hipStream_t st, uint64_t* out, const uint64_t* a, const uint64_t* b, int batch,
                int N, int backend);
  // out = inverse(x .* y) with x, y in the forward domain (device buffers).
  void inverse_of_product(hipStream_t st, uint64_t* out, const uint64_t* x, const uint64_t* y,
                          int batch, int N);
  // Scratch for a batch of transforms (grows on demand); trim() frees it.
  uint64_t* scratch(size_t words, int which);
  void trim();
  NttPlan& plan(int N);
  // Shape of the passes for length N (chosen by the autotuner): n1 = 2^n1_log
  // column length (n1_log = log2 N: single pass; 0: default_split), variant
  // bit 0/1 = alternative register grouping of the n1/n2 passes, cols1/cols2 =
  // sub-transforms per workgroup tile of those passes (0 = automatic).
  void tune(int N, int n1_log, int variant, int cols1, int cols2);
  static int default_split(int L);
  Profiler* prof = nullptr;  // optional per-launch timing

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

// Pointwise helpers (device buffers of n words, values < P).
void launch_pointwise_mul(hipStream_t st, uint64_t* out, const uint64_t* a, const uint64_t* b,
                          size_t n);

}  // namespace pcgsq
