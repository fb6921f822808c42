// Host-side tables of one NTT length (shared by the per-length launches in
// ntt.hip and the batched persistent launch in ntt_batch.hip).
#pragma once

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "internal.hpp"
#include "ntt_pass.hpp"

namespace pcgsq {

inline int brv_bits(uint64_t x, int bits) {
  uint64_t r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1ULL) << (bits - 1 - i);
  return (int)r;
}

inline std::vector<uint64_t> with_companions(const std::vector<uint64_t>& w) {
  std::vector<uint64_t> t(2 * w.size());
  for (size_t i = 0; i < w.size(); ++i) {
    t[2 * i] = w[i];
    t[2 * i + 1] = shoup_companion(w[i]);
  }
  return t;
}

inline void upload(DevBuf& buf, const std::vector<uint64_t>& v, const char* what) {
  buf.ensure(v.size() * sizeof(uint64_t), what);
  hip_check(hipMemcpy(buf.ptr, v.data(), v.size() * sizeof(uint64_t), hipMemcpyHostToDevice),
            what);
}

// Fill the four-step twiddle tables of a split N = 2^n1_log * 2^n2_log, one
// word w < P per element, indexed by the word offset at which the first pass
// stores that element (ntt.hip):
//   fwd[p 2^n2_log + c] = psi^(c (2p + 1)),       p < 2^n1_log, c < 2^n2_log
//   inv[k 2^n2_log + p] = psi^-(p (2k + 1)) N^-1, k < 2^n1_log, p < 2^n2_log
// from the two-level tables lo[e mod 2^lo_bits], hi[e >> lo_bits] of psi^(+-e)
// (the inverse high table carries N^-1).  Runs on a stream of its own and
// returns when the tables are written.
void ntt_build_twiddles(uint64_t* fwd, uint64_t* inv, const uint64_t* lo, const uint64_t* hi,
                        const uint64_t* loi, const uint64_t* hii, int n1_log, int n2_log,
                        int lo_bits);

struct NttPlan {
  int N = 0, L = 0, n1_log = 0, n2_log = 0;
  int variant = 0;           // bit 0: alternative stage grouping of the n1 passes, bit 1: n2
  int cols1 = 0, cols2 = 0;  // columns per tile of the n1 / n2 passes (0 = automatic)
  uint64_t psi = 0, ninv = 0, ninv_q = 0;
  DevBuf n1_neg_fwd, n1_neg_inv, n2_cyc_fwd, n2_cyc_inv;
  DevBuf tw_fwd, tw_inv;  // four-step twiddles, [N] each (ntt_build_twiddles)
  bool naive_ready = false;
  DevBuf nv_w, nv_winv, nv_twf, nv_twi;

  bool single() const { return n2_log == 0; }

  // n1_log == L: one pass (whole transform in one tile); otherwise the
  // four-step split N = n1 * n2.
  NttPlan(int N_, int n1_log_) : N(N_) {
    L = ilog2_exact((uint64_t)N);
    n1_log = n1_log_;
    n2_log = L - n1_log;
    psi = hpow(3, (kPrime - 1) / (2 * (uint64_t)N));
    const uint64_t psi_inv = hinv(psi);
    ninv = hinv((uint64_t)N);
    ninv_q = shoup_companion(ninv);
    const int n1 = 1 << n1_log, n2 = 1 << n2_log;
    // length-n1 negacyclic sub-transform, root psi1 = psi^n2 (order 2*n1)
    const uint64_t psi1 = hpow(psi, (uint64_t)n2), psi1_inv = hinv(psi1);
    std::vector<uint64_t> f(n1, 0), fi(n1, 0);
    for (int k = 1; k < n1; ++k) {
      f[k] = hpow(psi1, (uint64_t)brv_bits(k, n1_log));
      fi[k] = hpow(psi1_inv, (uint64_t)brv_bits(k, n1_log));
    }
    upload(n1_neg_fwd, with_companions(f), "ntt tables");
    upload(n1_neg_inv, with_companions(fi), "ntt tables");
    if (single()) return;
    // length-n2 cyclic sub-transform, root w2 = psi^(2 n1) (order n2)
    const uint64_t w2 = hpow(psi, 2 * (uint64_t)n1), w2_inv = hinv(w2);
    std::vector<uint64_t> g(n2, 0), gi(n2, 0);
    for (int k = 1; k < n2; ++k) {
      int s = 0;
      while ((2 << s) <= k) ++s;
      const int m = 1 << s, i = k - m;
      const uint64_t e = (uint64_t)brv_bits(i, s) * (uint64_t)(n2 / (2 * m));
      g[k] = hpow(w2, e);
      gi[k] = hpow(w2_inv, e);
    }
    upload(n2_cyc_fwd, with_companions(g), "ntt tables");
    upload(n2_cyc_inv, with_companions(gi), "ntt tables");
    // two-level psi^(+-e), e < 2N (the inverse high table carries N^-1), from
    // which the device fills the per-element tables
    const int lo_bits = (L + 2) / 2;
    const int hi_bits = L + 1 - lo_bits;
    std::vector<uint64_t> lo(1 << lo_bits), hi(1 << hi_bits), loi(1 << lo_bits),
        hii(1 << hi_bits);
    for (size_t x = 0; x < lo.size(); ++x) {
      lo[x] = hpow(psi, x);
      loi[x] = hpow(psi_inv, x);
    }
    for (size_t y = 0; y < hi.size(); ++y) {
      hi[y] = hpow(psi, y << lo_bits);
      hii[y] = hmul(hpow(psi_inv, y << lo_bits), ninv);
    }
    DevBuf dlo, dhi, dloi, dhii;
    upload(dlo, with_companions(lo), "ntt twiddles");
    upload(dhi, with_companions(hi), "ntt twiddles");
    upload(dloi, with_companions(loi), "ntt twiddles");
    upload(dhii, with_companions(hii), "ntt twiddles");
    tw_fwd.ensure((size_t)N * 8, "ntt twiddles");
    tw_inv.ensure((size_t)N * 8, "ntt twiddles");
    ntt_build_twiddles(tw_fwd.as<uint64_t>(), tw_inv.as<uint64_t>(), dlo.as<uint64_t>(),
                       dhi.as<uint64_t>(), dloi.as<uint64_t>(), dhii.as<uint64_t>(), n1_log,
                       n2_log, lo_bits);
  }

  void ensure_naive() {
    if (naive_ready) return;
    const uint64_t omega = hmul(psi, psi), omega_inv = hinv(omega);
    const uint64_t psi_inv = hinv(psi);
    std::vector<uint64_t> w(std::max(N / 2, 1)), wi(std::max(N / 2, 1)), tf(N), ti(N);
    w[0] = wi[0] = 1;
    for (int k = 1; k < N / 2; ++k) {
      w[k] = hmul(w[k - 1], omega);
      wi[k] = hmul(wi[k - 1], omega_inv);
    }
    tf[0] = 1;
    ti[0] = ninv;
    for (int k = 1; k < N; ++k) {
      tf[k] = hmul(tf[k - 1], psi);
      ti[k] = hmul(ti[k - 1], psi_inv);
    }
    upload(nv_w, with_companions(w), "naive tables");
    upload(nv_winv, with_companions(wi), "naive tables");
    upload(nv_twf, with_companions(tf), "naive tables");
    upload(nv_twi, with_companions(ti), "naive tables");
    naive_ready = true;
  }
};

// The four passes of a transform: 0 forward pass 1 (or the whole forward
// transform), 1 forward pass 2, 2 inverse pass 1 (or the whole inverse), 3
// inverse pass 2.  A one-pass plan has no phases 1 and 3.
enum NttPhase : int { kPhFwd1 = 0, kPhFwd2 = 1, kPhInv1 = 2, kPhInv2 = 3 };

struct PassSpec {
  int kind = -1;     // PassKind; -1: the phase does not exist
  int nlog = 0;      // sub-transform length of the pass
  int variant = 0;   // register grouping (0 / 1)
  int cols = 0;      // requested columns per tile (0 = automatic)
  int ncols = 0;     // columns of one transform (one-pass: transforms in the batch)
  NttPassDesc d{};
};

struct Profiler;

// Launch one pass over `batch` transforms (ntt.hip): the per-length pass
// kernels, tiles sized by the spec's cols.  With ps.d.boff set, transform b of
// the launch starts at word boff[b] of in, in2 and out.
void ntt_launch_pass(hipStream_t st, const PassSpec& ps, int batch, Profiler* prof,
                     const char* zone);

// Descriptor of pass `phase` of plan `pl` reading `in` (and `in2`, the second
// factor of a product, for phase 2) and writing `out`.  `batch` transforms
// lie back to back (stride N).
inline PassSpec plan_pass(const NttPlan& pl, int phase, const uint64_t* in, const uint64_t* in2,
                          uint64_t* out, int batch) {
  PassSpec p;
  NttPassDesc& d = p.d;
  d.in = in;
  d.in2 = in2;
  d.out = out;
  d.N_log = pl.L;
  d.scale_w = pl.ninv;
  d.scale_wq = pl.ninv_q;
  const bool v1 = pl.variant & 1, v2 = (pl.variant >> 1) & 1;
  if (pl.single()) {
    if (phase == kPhFwd2 || phase == kPhInv2) return p;
    const bool fwd = phase == kPhFwd1;
    p.kind = fwd ? kFwdWhole : (in2 ? kInvWholeProd : kInvWhole);
    p.nlog = pl.L; p.variant = v1; p.cols = pl.cols1; p.ncols = batch;
    d.tab = (fwd ? pl.n1_neg_fwd : pl.n1_neg_inv).as<uint64_t>();
    d.in_ps_log = 0; d.in_cs_log = pl.L; d.out_ps_log = 0; d.out_cs_log = pl.L;
    return p;
  }
  switch (phase) {
    case kPhFwd1:  // columns i2, length-n1 negacyclic over i1, twiddle psi^(i2(2k1+1))
      p.kind = kFwdCols; p.nlog = pl.n1_log; p.variant = v1; p.cols = pl.cols1;
      p.ncols = 1 << pl.n2_log;
      d.tab = pl.n1_neg_fwd.as<uint64_t>();
      d.tw = pl.tw_fwd.as<uint64_t>();
      d.in_ps_log = pl.n2_log; d.in_cs_log = 0; d.out_ps_log = pl.n2_log; d.out_cs_log = 0;
      break;
    case kPhFwd2:  // rows k1, length-n2 cyclic over i2, written transposed (natural k)
      p.kind = kFwdRows; p.nlog = pl.n2_log; p.variant = v2; p.cols = pl.cols2;
      p.ncols = 1 << pl.n1_log;
      d.tab = pl.n2_cyc_fwd.as<uint64_t>();
      d.in_ps_log = 0; d.in_cs_log = pl.n2_log; d.out_ps_log = pl.n1_log; d.out_cs_log = 0;
      break;
    case kPhInv1:  // columns k1 of [k2][k1], inverse cyclic over k2, twiddle psi^-(..) N^-1
      p.kind = in2 ? kInvRowsProd : kInvRows; p.nlog = pl.n2_log; p.variant = v2;
      p.cols = pl.cols2; p.ncols = 1 << pl.n1_log;
      d.tab = pl.n2_cyc_inv.as<uint64_t>();
      d.tw = pl.tw_inv.as<uint64_t>();
      d.in_ps_log = pl.n1_log; d.in_cs_log = 0; d.out_ps_log = 0; d.out_cs_log = pl.n2_log;
      break;
    default:  // columns i2, inverse negacyclic over k1
      p.kind = kInvCols; p.nlog = pl.n1_log; p.variant = v1; p.cols = pl.cols1;
      p.ncols = 1 << pl.n2_log;
      d.tab = pl.n1_neg_inv.as<uint64_t>();
      d.in_ps_log = pl.n2_log; d.in_cs_log = 0; d.out_ps_log = pl.n2_log; d.out_cs_log = 0;
      break;
  }
  return p;
}

}  // namespace pcgsq
