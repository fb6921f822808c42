// PCG-OLE output layer on the device (host-side launchers).
//
// Semantics follow pcg_ole_impl.h:330-449 and leaf_convert_cuda.cu:133-234:
//   C = lo64(H'(leaf)) mod P, tau = lsb(leaf);
//   A_s = (-1)^s sum tau, B_s = (-1)^(1-s) sum C + beta_s;
//   Beaver open d = A - r, e = B - s;  CW_s = d s_s + e r_s + rs_s (+ d e, s=0);
//   y_s = (-1)^s (C + tau CW);  g[(kk+ll) bin + d] += y, folded X^N = -1.
// The fold into g is a gather: every output coefficient sums the (at most
// 4t) leaves that land on it, so no atomics are needed and the result does
// not depend on scheduling.  The PCG-OLE expansion does the gather inside the
// DPF leaf level (dpf.hip) and adds the tau*CW terms afterwards
// (launch_gather_fix); launch_gather_g serves the drop-in scatter and bins of
// fewer than kGatherTile coefficients.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "device_common.hpp"
#include "profiler.hpp"

namespace pcgsq {

// The leaves that land on output coefficient q (0 <= q < N) of one
// (party, block) set, in the order the output layer sums them: for h = 0, 1
// the raw position pos = q + h*N (h = 1 folds with a minus sign), for ds =
// 0, 1 the diagonal sd = (pos >> bin_log) - ds and, along it, the instances
// inst = kk*t + (sd - kk) < B, each at leaf d = (pos & (bin - 1)) + ds*bin
// (pcg_ole_impl.h:432-446).  Calls f(slot, h, ds, inst) with slot = 0, 1, ...
// and returns the number of calls (at most 4t).  Every q of an aligned run of
// bin coefficients sees the same sequence, so a workgroup whose coefficients
// share one bin enumerates it with uniform control flow.
template <class F>
__device__ __forceinline__ int for_each_landing(uint32_t q, int N_log, int bin_log, int t, int B,
                                                F&& f) {
  int slot = 0;
  for (int h = 0; h < 2; ++h) {
    const int sb = (int)(((uint64_t)q + ((uint64_t)h << N_log)) >> bin_log);
    for (int ds = 0; ds < 2; ++ds) {
      const int sd = sb - ds;
      if (sd < 0 || sd > 2 * t - 2) continue;
      const int kk_lo = sd - t + 1 > 0 ? sd - t + 1 : 0;
      const int kk_hi = sd < t - 1 ? sd : t - 1;
      for (int kk = kk_lo; kk <= kk_hi; ++kk) {
        const int inst = kk * t + (sd - kk);
        if (inst >= B) break;
        f(slot++, h, ds, inst);
      }
    }
  }
  return slot;
}

// Slot `slot` of that sequence in closed form: sets h and inst and returns
// true, or returns false past the last slot.
__device__ __forceinline__ bool landing_slot(uint32_t q, int N_log, int bin_log, int t, int B,
                                             int slot, int& h, int& inst) {
  for (int hh = 0; hh < 2; ++hh) {
    const int sb = (int)(((uint64_t)q + ((uint64_t)hh << N_log)) >> bin_log);
    for (int ds = 0; ds < 2; ++ds) {
      const int sd = sb - ds;
      if (sd < 0 || sd > 2 * t - 2) continue;
      const int kk_lo = sd - t + 1 > 0 ? sd - t + 1 : 0;
      int kk_hi = sd < t - 1 ? sd : t - 1;
      // inst = sd + kk*(t-1) grows with kk; the sequence stops at inst >= B
      if (t > 1) {
        const int kmax = B - 1 - sd >= 0 ? (B - 1 - sd) / (t - 1) : -1;
        kk_hi = kk_hi < kmax ? kk_hi : kmax;
      } else if (sd >= B) {
        kk_hi = kk_lo - 1;
      }
      const int n = kk_hi - kk_lo + 1 > 0 ? kk_hi - kk_lo + 1 : 0;
      if (slot < n) {
        h = hh;
        inst = (kk_lo + slot) * t + (sd - kk_lo - slot);
        return true;
      }
      slot -= n;
    }
  }
  return false;
}

// The leaf level with the gather fused in gives each workgroup at least
// kGatherTile output coefficients of one bin (two per thread; more when the
// bins are larger), and handles at most kGatherMaxSlots landing slots
// (t <= 64).
constexpr int kGatherTileLog = 9;
constexpr int kGatherTile = 1 << kGatherTileLog;
constexpr int kGatherMaxSlots = 256;

// Second half of the fused output layer: g = (-1)^party (A + sum over the
// landing slots whose tau bit is set of +-CW[inst]), in place over A.
// For the launch's party p (party first_party + p): A/g at g +
// p*g_party_stride + bb*N; masks [mask_words][parties][sets][N]; CW
// [sets][B] for the set's instances.  Needs bin >= 512.
struct GatherFixArgs {
  uint64_t* g = nullptr;
  size_t g_party_stride = 0;
  const uint64_t* mask = nullptr;
  int mask_words = 0;
  const uint64_t* CW = nullptr;
  int sets = 0, B = 0, t = 0, N_log = 0, bin_log = 0;
  int parties = 2, first_party = 0;
  Profiler* prof = nullptr;
};
void launch_gather_fix(hipStream_t st, const GatherFixArgs& a);

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
