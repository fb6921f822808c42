// PCG-OLE output layer on the device (host-side launchers).
//
// Semantics follow pcg_ole_impl.h:330-449 and leaf_convert_cuda.cu:133-234:
//   C = lo64(H'(leaf)) mod P, tau = lsb(leaf);
//   A_s = (-1)^s sum tau, B_s = (-1)^(1-s) sum C + beta_s;
//   Beaver open d = A - r, e = B - s;  CW_s = d s_s + e r_s + rs_s (+ d e, s=0);
//   y_s = (-1)^s (C + tau CW);  g[(kk+ll) bin + d] += y, folded X^N = -1.
// The fold into g is a gather: every output coefficient sums the (at most
// 4t) leaves that land on it, so no atomics are needed and the result does
// not depend on scheduling.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "device_common.hpp"
#include "profiler.hpp"

namespace pcgsq {

// partials [ntrees][nparts][2] -> sumC[ntrees] (mod P), sumT[ntrees].
void launch_fold_sums(hipStream_t st, const uint64_t* partials, int ntrees, int nparts,
                      uint64_t* sumC, uint64_t* sumT);

// Per-instance output correction word for both parties in one process.
// partials as above for 2*ninst trees (party 0 first); beta/r/s/rs are
// [2][ninst] (party-major).  CW[ninst].
// Element (party p, instance b) of beta/r/s/rs is at [p*pstride + b].
void launch_beaver_cw(hipStream_t st, const uint64_t* partials, int ninst, int nparts,
                      const uint64_t* beta, const uint64_t* r, const uint64_t* s,
                      const uint64_t* rs, size_t pstride, uint64_t* CW);

// Folded g for `nsets` (party, block) sets at once.  Set q belongs to party
// p = q / sets_per_party and block bb = q % sets_per_party; it reads
// src + q*set_src_stride (C|tau<<63 words or leaves, [B][D]) and
// CW + bb*cw_stride [B], and writes g + p*g_party_stride + bb*N.  Party 1
// negates.
struct GatherArgs {
  const uint64_t* ct = nullptr;   // C|tau words (convert path) ...
  const Blk* leaves = nullptr;    // ... or raw leaves (out-hash recomputed)
  size_t set_src_stride = 0;      // elements between sets
  const uint64_t* CW = nullptr;
  size_t cw_stride = 0;
  uint64_t* g = nullptr;
  size_t g_party_stride = 0;
  int nsets = 1;
  int sets_per_party = 1;
  int first_party = 0;            // party of set 0 (single-party calls)
  int B = 0, t = 0, N = 0;
  int D_log = 0;                  // D = 2*bin = 2^D_log
  Profiler* prof = nullptr;
};
void launch_gather_g(hipStream_t st, const GatherArgs& a);

// Sums straight from stored leaves (drop-in dpf_out_sums): partials then fold.
void launch_leaf_sums(hipStream_t st, const Blk* leaves, int B, int D_log, uint64_t* partials,
                      int nparts, uint64_t* sumC, uint64_t* sumT);
int leaf_sums_nparts(int D_log);

}  // namespace pcgsq
