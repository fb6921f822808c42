// PCG-OLE expansion plan: both parties on one MI355X.
//
// Follows pcg_ole::PCG_OLE<P>::gen_and_expand
// (CPU_baseline/pcg_ole_2pc/pcg_ole_impl.h:46-490):
//   1. noise and preprocessing (here: the seeded dealer below),
//   2. for every DPF block (ii, jj): B = t^2 half-tree DPFs of domain
//      D = 2N/t for both parties (fused protocol, in-device exchange), the
//      output layer (ChaCha8 out-hash fused into the last tree level,
//      Beaver opens, folded scatter into g_ij),
//   3. step 4: x = sum_i a_i e_i, z = sum_ij (a_i a_j) fold(g_ij).
// Step 4 is evaluated in the NTT domain: with A_i = NTT(a_i),
// X = sum_i A_i .* NTT(e_i) and Z = sum_ij A_i .* A_j .* NTT(g_ij), then one
// inverse each.  Products in R_p are unique, so this is bit-identical to the
// reference's c + 2c^2 separate poly-muls while computing 2c + 2c^2 forward
// and 4 inverse transforms per expansion (both parties).
//
// Dealer spec (shared with the oracle, oracle/orc_pcg.c, which must agree):
//   U(seed, dom, a, b) = sm(sm(sm(seed ^ dom*K) ^ a) ^ b), sm = splitmix64
//   finalizer, K = 0xD6E8FEB86659FD93.
//   pos[s][i][k] = k*bin + U(seed,1,s,i*t+k) % bin;
//   pay[s][i][k] = U(seed,2,s,i*t+k) % (P-1) + 1;
//   instance g = blk*B + idx (blk = ii*c + jj, idx = kk*t + ll);
//   delta_s = (U(seed,3,s,2g), U(seed,3,s,2g+1)), lsb forced 1 (s=0) / 0 (s=1);
//   root_s = delta_s ^ W, W = (0x0123456789abcdef, 0xdeadbeefcafe0001);
//   alpha = pos0[ii][kk] % bin + pos1[jj][ll] % bin;
//   dltsft0[i][k] = (U(seed,4,g,2e), U(seed,4,g,2e+1)), e = i*nk + k;
//   dltsft1 = dltsft0 ^ ([digit_i(alpha) = k] ? Delta : 0);
//   beta0 = U(seed,5,g,0) % P, beta1 = pay0*pay1 - beta0;
//   r_s = U(seed,6,g,s) % P, s_s = U(seed,7,g,s) % P,
//   rs0 = U(seed,8,g,0) % P, rs1 = (r0+r1)(s0+s1) - rs0;
//   public_a[i][j] = U(pub_seed,9,i,j) % P.
// Preprocessing as the parties run it (preproc.hip): each party draws its own
// positions, payloads, Delta share and r, s with the formulas above and its
// own seed, and the parties compute alpha, DltSft, beta and rs together by
// OT extension, GMW and Gilboa products, exchanging every message:
//   * one party per process: pcgsq_ole_party_deal_own, then
//     pcgsq_ole_party_run runs the protocol through the caller's exchange;
//   * both parties in one plan: pcgsq_ole_deal_gmw(seed) runs the two
//     parties' protocol in-process (each with `seed` as its own seed), the
//     messages moved by device copies.
// The shares differ from the dealer's but reconstruct to the same values, so
// the expansion's outputs are identical.
#include <string.h>

#include <algorithm>
#include <vector>

#include <memory>

#include "context.hpp"
#include "ot.hpp"
#include "pcgsq.h"
#include "preproc.hpp"

namespace pcgsq {

namespace {

// The dealer of the spec above, on the device: the seed is all a request
// carries to the GPU (t and bin are powers of two since t | N).  Noise: one
// thread per (i, k), one grid row per party s.
__global__ void deal_noise_kernel(uint64_t* __restrict__ noise, uint64_t seed, int c, int t_log,
                                  uint64_t bin) {
  const int s = blockIdx.y;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;  // i*t + k
  const int ct = c << t_log;
  if (r >= ct) return;
  const int k = r & ((1 << t_log) - 1);
  const int o = s * ct + r;
  noise[o] = (uint64_t)k * bin + (dealer_hd(seed, 1, (uint64_t)s, (uint64_t)r) & (bin - 1));
  noise[2 * ct + o] = dealer_hd(seed, 2, (uint64_t)s, (uint64_t)r) % (kP - 1) + 1;
}

// generate_triples (CPU_baseline/pcg_ole_2pc/triple.h:45-110) from the
// outputs [x0, z0, x1, z1] of two OLEs: a_0 = x_{0,1}, b_0 = x_{0,2},
// a_1 = x_{1,2}, b_1 = x_{1,1} and z_s = z_{s,1} + z_{s,2}, into
// T = [a0, b0, a1, b1, z0, z1][N].
__global__ void triple_compose_kernel(const uint64_t* __restrict__ o1,
                                      const uint64_t* __restrict__ o2,
                                      uint64_t* __restrict__ T, int N) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N) return;
  const size_t n = (size_t)N;
  T[i] = o1[i];                                  // a0 = x_{0,1}
  T[n + i] = o2[i];                              // b0 = x_{0,2}
  T[2 * n + i] = o2[2 * n + i];                  // a1 = x_{1,2}
  T[3 * n + i] = o1[2 * n + i];                  // b1 = x_{1,1}
  T[4 * n + i] = fadd(o1[n + i], o2[n + i]);     // z_{0,1} + z_{0,2}
  T[5 * n + i] = fadd(o1[3 * n + i], o2[3 * n + i]);
}

// With every row of T forward-transformed, slot i of the ring triple's c_s =
// a_s b_s + z_s is A_s[i] B_s[i] + Z_s[i] (the ring product is slot-wise in
// the NTT domain), written over Z_s.
__global__ void triple_finish_kernel(uint64_t* __restrict__ T, int N) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N) return;
  const size_t n = (size_t)N;
  T[4 * n + i] = fadd(fmul(T[i], T[n + i]), T[4 * n + i]);
  T[5 * n + i] = fadd(fmul(T[2 * n + i], T[3 * n + i]), T[5 * n + i]);
}

// a_0 = 1 and a_i = public_a[i-1] (pcg_ole_impl.h:455-458); grid row i.
__global__ void public_a_kernel(uint64_t* __restrict__ a, uint64_t pub_seed, int N) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  a[(size_t)i * N + j] =
      i == 0 ? (j == 0 ? 1u : 0u) : dealer_hd(pub_seed, 9, (uint64_t)(i - 1), (uint64_t)j) % kP;
}

// Per instance g = blk*B + idx: both roots, both DltSft shares, beta and the
// Beaver shares.  One thread per instance idx = kk*t + ll; grid rows ii, jj.
__global__ void deal_inst_kernel(const uint64_t* __restrict__ noise, Blk* __restrict__ roots,
                                 Blk* __restrict__ dlt, uint64_t* __restrict__ bt,
                                 uint64_t seed, int c, int t_log, int levels, int nk, int w,
                                 int D_log, uint64_t bin, int ninst) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = 1 << t_log, B = t << t_log;
  if (idx >= B) return;
  const int ii = blockIdx.y, jj = blockIdx.z;
  const int blk = ii * c + jj;
  const int g = blk * B + idx;
  const int kk = idx >> t_log, ll = idx & (t - 1);
  const uint64_t* pos = noise;
  const uint64_t* pay = noise + (size_t)2 * c * t;
  const uint64_t p0 = pos[(size_t)ii * t + kk], p1 = pos[((size_t)c + jj) * t + ll];
  const uint64_t alpha = (p0 & (bin - 1)) + (p1 & (bin - 1));
  const uint64_t G = (uint64_t)g;
  Blk d0{dealer_hd(seed, 3, 0, 2 * G), dealer_hd(seed, 3, 0, 2 * G + 1)};
  Blk d1{dealer_hd(seed, 3, 1, 2 * G), dealer_hd(seed, 3, 1, 2 * G + 1)};
  d0.lo |= 1ULL;   // dpf.cpp:78-82
  d1.lo &= ~1ULL;
  const Blk Delta{d0.lo ^ d1.lo, d0.hi ^ d1.hi};
  const size_t ni = (size_t)ninst;
  roots[g] = Blk{d0.lo ^ kWlo, d0.hi ^ kWhi};
  roots[ni + g] = Blk{d1.lo ^ kWlo, d1.hi ^ kWhi};
  for (int i = 0; i < levels; ++i) {
    const int digit = (int)((alpha >> (D_log - w * (i + 1))) & (uint64_t)((1 << w) - 1));
    for (int k = 0; k < nk; ++k) {
      const uint64_t e = (uint64_t)(i * nk + k);
      Blk s0{dealer_hd(seed, 4, G, 2 * e), dealer_hd(seed, 4, G, 2 * e + 1)};
      dlt[((size_t)g * levels + i) * nk + k] = s0;
      if (digit == k) { s0.lo ^= Delta.lo; s0.hi ^= Delta.hi; }
      dlt[((ni + g) * levels + i) * nk + k] = s0;
    }
  }
  const uint64_t pay0 = pay[(size_t)ii * t + kk], pay1 = pay[((size_t)c + jj) * t + ll];
  uint64_t* BETA = bt;
  uint64_t* R = bt + 2 * ni;
  uint64_t* S = bt + 4 * ni;
  uint64_t* RS = bt + 6 * ni;
  const uint64_t r0 = dealer_hd(seed, 6, G, 0) % kP, r1 = dealer_hd(seed, 6, G, 1) % kP;
  const uint64_t s0 = dealer_hd(seed, 7, G, 0) % kP, s1 = dealer_hd(seed, 7, G, 1) % kP;
  R[g] = r0;
  R[ni + g] = r1;
  S[g] = s0;
  S[ni + g] = s1;
  BETA[g] = dealer_hd(seed, 5, G, 0) % kP;
  BETA[ni + g] = fsub(fmul(pay0, pay1), BETA[g]);
  RS[g] = dealer_hd(seed, 8, G, 0) % kP;
  RS[ni + g] = fsub(fmul(fadd(r0, r1), fadd(s0, s1)), RS[g]);
}

__global__ void build_e_kernel(uint64_t* e, const uint64_t* pos, const uint64_t* pay, int c, int t,
                               int N) {
  // e: [2][c][N] zero-filled beforehand; pos/pay: [2][c][t]; one y-block row
  // per polynomial
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= t) return;
  const int poly = blockIdx.y;
  const int i = poly * t + k;
  e[(size_t)poly * N + pos[i]] = pay[i];
}

// X_s = sum_i A_i E_{s,i};  Z_s = sum_i A_i sum_j A_j G_{s,ij}  (NTT domain).
// C = c known at compile time: every load of a coefficient is issued before
// the products (independent loads in flight), and A is loaded once for both
// parties.
template <int C>
__global__ void step4_dot_c_kernel(const uint64_t* __restrict__ A, const uint64_t* __restrict__ G,
                                   const uint64_t* __restrict__ E, uint64_t* __restrict__ out,
                                   int N, int s0, int ns) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)N) return;
  uint64_t a[C];
#pragma unroll
  for (int i = 0; i < C; ++i) a[i] = A[(size_t)i * N + q];
  for (int s = s0; s < s0 + ns; ++s) {
    uint64_t e[C], g[C * C];
#pragma unroll
    for (int i = 0; i < C; ++i) e[i] = E[((size_t)s * C + i) * N + q];
#pragma unroll
    for (int k = 0; k < C * C; ++k) g[k] = G[((size_t)s * C * C + k) * N + q];
    uint64_t x = 0, z = 0;
#pragma unroll
    for (int i = 0; i < C; ++i) {
      x = fadd(x, fmul(a[i], e[i]));
      uint64_t inner = 0;
#pragma unroll
      for (int j = 0; j < C; ++j) inner = fadd(inner, fmul(a[j], g[i * C + j]));
      z = fadd(z, fmul(a[i], inner));
    }
    out[((size_t)2 * s) * N + q] = x;
    out[((size_t)2 * s + 1) * N + q] = z;
  }
}

// The same for any c (one party per grid row).
__global__ void step4_dot_kernel(const uint64_t* __restrict__ A, const uint64_t* __restrict__ G,
                                 const uint64_t* __restrict__ E, uint64_t* __restrict__ out,
                                 int c, int N, int s0) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = s0 + blockIdx.y;  // party
  if (q >= (size_t)N) return;
  uint64_t x = 0, z = 0;
  for (int i = 0; i < c; ++i) {
    const uint64_t ai = A[(size_t)i * N + q];
    x = fadd(x, fmul(ai, E[((size_t)s * c + i) * N + q]));
    uint64_t inner = 0;
    for (int j = 0; j < c; ++j) {
      const uint64_t aj = A[(size_t)j * N + q];
      inner = fadd(inner, fmul(aj, G[(((size_t)s * c + i) * c + j) * N + q]));
    }
    z = fadd(z, fmul(ai, inner));
  }
  out[((size_t)2 * s) * N + q] = x;
  out[((size_t)2 * s + 1) * N + q] = z;
}

void launch_step4_dot(hipStream_t st, const uint64_t* A, const uint64_t* G, const uint64_t* E,
                      uint64_t* out, int c, int N, int s0, int ns) {
  const dim3 grid((unsigned)((N + 255) / 256));
  switch (c) {
    case 2:
      hipLaunchKernelGGL(step4_dot_c_kernel<2>, grid, dim3(256), 0, st, A, G, E, out, N, s0, ns);
      break;
    case 3:
      hipLaunchKernelGGL(step4_dot_c_kernel<3>, grid, dim3(256), 0, st, A, G, E, out, N, s0, ns);
      break;
    case 4:
      hipLaunchKernelGGL(step4_dot_c_kernel<4>, grid, dim3(256), 0, st, A, G, E, out, N, s0, ns);
      break;
    default:
      hipLaunchKernelGGL(step4_dot_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)ns),
                         dim3(256), 0, st, A, G, E, out, c, N, s0);
  }
  hip_check(hipGetLastError(), "launch step4_dot");
}

__global__ void zero_kernel(uint64_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int ilog(uint64_t x) { return ilog2_exact(x); }

}  // namespace

uint64_t dealer_u64(uint64_t seed, uint32_t domain, uint64_t a, uint64_t b) {
  return dealer_hd(seed, domain, a, b);
}

struct OleDims {
  int N, c, t, w, prg;
  int B, bin, D, D_log, levels, nk, nblk, ninst, L;
};

static OleDims make_dims(const pcgsq_params& p) {
  OleDims d{};
  require(p.N > 0 && (p.N & (p.N - 1)) == 0, "ole: N must be a power of two");
  require(p.t > 0 && p.N % p.t == 0, "ole: t must divide N");
  require(p.c >= 1 && p.c <= 16, "ole: c out of range");
  require(p.prg == PCGSQ_PRG_AES || p.prg == PCGSQ_PRG_CHACHA8, "ole: unknown PRG");
  d.N = p.N; d.c = p.c; d.t = p.t; d.w = p.w; d.prg = p.prg;
  d.L = ilog((uint64_t)p.N);
  require(d.L >= 2 && d.L <= 24, "ole: N must be in [4, 2^24] (2N | P-1)", kErrUnsupported);
  // instance counts and every device index derived from them are int: bound
  // c^2 t^2 (both parties' trees, 2 c^2 t^2, too) in 64-bit arithmetic first
  const uint64_t inst64 = (uint64_t)p.c * (uint64_t)p.c * (uint64_t)p.t * (uint64_t)p.t;
  require(inst64 < ((uint64_t)1 << 30), "ole: c^2 t^2 DPF instances must be below 2^30",
          kErrUnsupported);
  d.B = p.t * p.t;
  d.bin = p.N / p.t;
  d.D = 2 * d.bin;
  d.D_log = ilog((uint64_t)d.D);
  require(d.D_log > 0, "ole: 2N/t must be a power of two");
  require(p.w >= 1 && p.w <= 4 && d.D_log % p.w == 0, "ole: w must be in [1,4] and divide log2(2N/t)");
  d.levels = d.D_log / p.w;
  d.nk = (1 << p.w) - 1;
  d.nblk = p.c * p.c;
  d.ninst = d.nblk * d.B;
  return d;
}

}  // namespace pcgsq

using namespace pcgsq;

struct pcgsq_ole_plan {
  pcgsq_ctx* ctx = nullptr;
  OleDims d{};
  uint64_t pub_seed = 0;
  int party = -1;        // -1: both parties in one process; 0 / 1: that party alone
  int group_blocks = 1;
  bool dealt = false;
  DpfScratch dpf;
  // device-resident inputs
  DevBuf roots, dlt, bt, noise;
  DevBuf aes_t0, aes_rks;
  // public NTT(a_i)
  DevBuf ahat;
  // working set
  DevBuf ct, part, cw, fw, out;   // ct: C|tau per leaf, or the fused gather's tau masks
  DevBuf tri;   // triple generation: first OLE's outputs [4][N] + shares [6][N]
  // the 2PC preprocessing (preproc.hpp): the run of each party this plan holds
  PreRun pre[2];
  bool have_base = false;   // base OTs done for this plan's party pair
  uint64_t sid = 0;         // next IKNP session (both parties count alike)
  bool own_deal = false;    // one-party plan: the next run starts with the preprocessing
  uint64_t own_seed = 0;
  int exchange_mode = PCGSQ_EXCHANGE_PER_BLOCK;
  PinnedBuf hsend, hrecv;   // host staging of one party's messages
};

namespace {

// The output layer's gather runs fused into the DPF leaf level when every
// workgroup tile of output coefficients lies in one bin and the landing
// slots fit its LDS (dpf.hip, leaf_gather_kernel); otherwise the leaf level
// writes C|tau per leaf and a separate gather reads them.
bool fused_gather(const OleDims& d) {
  return d.bin >= kGatherTile && d.t <= 64 && d.levels >= 2;
}
int gather_mask_words(const OleDims& d) { return (4 * d.t + 63) / 64; }
int leaf_nparts(const OleDims& d) {
  return fused_gather(d) ? dpf_gather_nparts(d.D_log) : dpf_leaf_nparts(d.D_log, d.w, 8);
}

// Whether the plan's DPF runs expand the last three tree levels inside the
// leaf level (dpf.hpp, dpf_deep_leaf): only smaller frontiers are stored.
bool plan_deep(const OleDims& d, int parties) {
  return dpf_deep_leaf(d.D_log, d.w, d.prg, fused_gather(d) ? kLeafGather : kLeafConvert,
                       parties == 1, parties);
}

// parties: 2 when one plan runs both parties, 1 for one party.
size_t plan_group_bytes(const OleDims& d, int gb, int parties) {
  const size_t trees = (size_t)parties * gb * d.B;
  const size_t maxS = dpf_frontier_nodes(d.D_log, d.w, plan_deep(d, parties));
  const size_t fh = trees * maxS * (1 + d.nk) * sizeof(Blk) * 2;
  const size_t leaf = fused_gather(d)
                          ? (size_t)gather_mask_words(d) * parties * gb * d.N * sizeof(uint64_t)
                          : trees * (size_t)d.D * sizeof(uint64_t);
  return fh + leaf;
}

// DPF blocks per group: as many as the budget holds, dividing c^2 (equal
// groups keep every launch the same shape).
int plan_group(const OleDims& d, size_t budget, int parties) {
  if (budget == 0) budget = (size_t)96 << 30;
  int gb = d.nblk;
  while (gb > 1 && plan_group_bytes(d, gb, parties) > budget) --gb;
  while (d.nblk % gb) --gb;
  return gb;
}

// Device bytes a plan (with its DPF scratch) and its context's NTT scratch
// take, without allocating anything.
size_t plan_estimate(const OleDims& d, size_t budget, int parties = 2) {
  const int gb = plan_group(d, budget, parties);
  const size_t gtrees = (size_t)parties * gb * d.B;
  const size_t trees_all = (size_t)2 * d.ninst;
  const int nparts = leaf_nparts(d);
  const size_t nfw = (size_t)2 * d.nblk + 2 * d.c;
  size_t b = plan_group_bytes(d, gb, parties);
  b += (size_t)d.levels * gtrees * d.nk * sizeof(Blk) * (parties == 1 ? 2 : 1);  // cw (+ peer)
  b += trees_all * sizeof(Blk) + trees_all * d.levels * d.nk * sizeof(Blk);
  b += (size_t)8 * d.ninst * sizeof(uint64_t) + (size_t)4 * d.c * d.t * sizeof(uint64_t);
  b += (size_t)d.c * d.N * sizeof(uint64_t);                       // ahat
  b += gtrees * nparts * 2 * sizeof(uint64_t) + (size_t)d.ninst * sizeof(uint64_t);
  b += 2 * nfw * d.N * sizeof(uint64_t);                           // fw + NTT scratch
  b += (size_t)4 * d.N * sizeof(uint64_t);                         // outputs
  return b + ((size_t)1 << 20);
}

void plan_alloc(pcgsq_ole_plan* pl, size_t budget) {
  const OleDims& d = pl->d;
  if (budget == 0) budget = (size_t)96 << 30;
  const int parties = pl->party < 0 ? 2 : 1;
  const int gb = plan_group(d, budget, parties);
  require(plan_group_bytes(d, gb, parties) <= budget || gb == 1, "ole: HBM budget too small",
          kErrNoMemory);
  pl->group_blocks = gb;
  const size_t trees_all = (size_t)2 * d.ninst;
  pl->roots.ensure(trees_all * sizeof(Blk), "ole roots");
  pl->dlt.ensure(trees_all * d.levels * d.nk * sizeof(Blk), "ole dltsft");
  pl->bt.ensure((size_t)4 * 2 * d.ninst * sizeof(uint64_t), "ole beaver");
  pl->noise.ensure((size_t)2 * 2 * d.c * d.t * sizeof(uint64_t), "ole noise");
  pl->ahat.ensure((size_t)d.c * d.N * sizeof(uint64_t), "ole ahat");
  const size_t gtrees = (size_t)parties * gb * d.B;
  if (fused_gather(d))
    pl->ct.ensure((size_t)gather_mask_words(d) * parties * gb * d.N * sizeof(uint64_t),
                  "ole tau masks");
  else
    pl->ct.ensure(gtrees * d.D * sizeof(uint64_t), "ole ct");
  const int nparts = leaf_nparts(d);
  pl->part.ensure(gtrees * nparts * 2 * sizeof(uint64_t), "ole partials");
  // CWs; a one-party plan also keeps its group's sums after them
  pl->cw.ensure(((size_t)d.ninst + (parties == 1 ? 2 * gtrees : 0)) * sizeof(uint64_t), "ole cw");
  pl->fw.ensure(((size_t)2 * d.nblk + 2 * d.c) * d.N * sizeof(uint64_t), "ole ntt buffers");
  pl->out.ensure((size_t)4 * d.N * sizeof(uint64_t), "ole outputs");
  // the DPF frontier and the NTT scratch too, so that a run allocates nothing
  // (an allocation inside a run would stall other streams' work)
  dpf_reserve(pl->dpf, d.D_log, d.w, gtrees, parties == 1, plan_deep(d, parties));
  (void)pl->ctx->ntt.scratch(((size_t)2 * d.nblk + 2 * d.c) * d.N, 0);
  if (d.prg == PCGSQ_PRG_AES) {
    uint32_t t0[256];
    aes_ttable_host(t0);
    std::vector<uint8_t> rks((size_t)d.nk * 176);
    static const uint8_t S[16] = {0x2b, 0x7e, 0x15, 0x16, 0x28, 0xae, 0xd2, 0xa6,
                                  0xab, 0xf7, 0x15, 0x88, 0x09, 0xcf, 0x4f, 0x3c};
    aes_tree_round_keys_host(S, d.nk, rks.data());  // pcg_ole_impl.h:89-92
    pl->aes_t0.ensure(sizeof(t0), "ole aes");
    pl->aes_rks.ensure(rks.size(), "ole aes");
    hip_check(hipMemcpy(pl->aes_t0.ptr, t0, sizeof(t0), hipMemcpyHostToDevice), "ole aes");
    hip_check(hipMemcpy(pl->aes_rks.ptr, rks.data(), rks.size(), hipMemcpyHostToDevice), "ole aes");
  }
  // ahat: a_0 = 1, a_i = public_a[i-1], generated on the device
  hipLaunchKernelGGL(public_a_kernel, dim3((unsigned)((d.N + 255) / 256), d.c), dim3(256), 0,
                     pl->ctx->stream, pl->ahat.as<uint64_t>(), pl->pub_seed, d.N);
  hip_check(hipGetLastError(), "launch public_a");
  pl->ctx->ntt.forward(pl->ctx->stream, pl->ahat.as<uint64_t>(), d.c, d.N, kNttFourStep);
  hip_check(hipStreamSynchronize(pl->ctx->stream), "ole ahat");
}

void plan_deal(pcgsq_ole_plan* pl, uint64_t seed) {
  const OleDims& d = pl->d;
  hipStream_t st = pl->ctx->stream;
  const int t_log = ilog((uint64_t)d.t);
  require(t_log >= 0, "ole: t must be a power of two");
  hipLaunchKernelGGL(deal_noise_kernel, dim3((unsigned)((d.c * d.t + 255) / 256), 2), dim3(256),
                     0, st, pl->noise.as<uint64_t>(), seed, d.c, t_log, (uint64_t)d.bin);
  hipLaunchKernelGGL(deal_inst_kernel, dim3((unsigned)((d.B + 255) / 256), d.c, d.c), dim3(256),
                     0, st, pl->noise.as<uint64_t>(), pl->roots.as<Blk>(), pl->dlt.as<Blk>(),
                     pl->bt.as<uint64_t>(), seed, d.c, t_log, d.levels, d.nk, d.w, d.D_log,
                     (uint64_t)d.bin, d.ninst);
  hip_check(hipGetLastError(), "launch dealer");
  pl->dealt = true;
  pl->own_deal = false;
}

PreOut plan_pre_out(pcgsq_ole_plan* pl) {
  PreOut o;
  o.noise = pl->noise.as<uint64_t>();
  o.roots = pl->roots.as<Blk>();
  o.dlt = pl->dlt.as<Blk>();
  o.bt = pl->bt.as<uint64_t>();
  return o;
}

// Both parties' preprocessing in one plan, in-process: the two parties' runs
// step in lockstep on the plan's stream and every message is a device copy
// from one party's send buffer into the other's receive buffer.  Base OTs
// (host, libcrypto) once per plan.
void plan_deal_gmw(pcgsq_ole_plan* pl, uint64_t seed) {
  require(pl->party < 0, "ole_deal_gmw: a one-party plan runs its preprocessing with the other "
                         "party: pcgsq_ole_party_deal_own, then pcgsq_ole_party_run");
  const OleDims& d = pl->d;
  hipStream_t st = pl->ctx->stream;
  if (!pl->have_base) {
    const PreShape sh = make_pre_shape(d.N, d.c, d.t, d.w);
    BaseOtParty a(0, seed), b(1, seed);
    const std::vector<uint8_t> a1 = a.round1(), b1 = b.round1();
    const std::vector<uint8_t> a2 = a.round2(b1.data(), b1.size());
    const std::vector<uint8_t> b2 = b.round2(a1.data(), a1.size());
    std::unique_ptr<BaseOts> ba(new BaseOts), bb(new BaseOts);
    a.finish(b2.data(), b2.size(), ba.get());
    b.finish(a2.data(), a2.size(), bb.get());
    pre_prepare(pl->pre[0], sh, 0, *ba);
    pre_prepare(pl->pre[1], sh, 1, *bb);
    pl->have_base = true;
  }
  const PreOut out = plan_pre_out(pl);
  const uint64_t sid = pl->sid++;
  pre_begin(pl->pre[0], seed, sid, out, st);
  pre_begin(pl->pre[1], seed, sid, out, st);
  for (;;) {
    PreMsg m0, m1;
    const bool more0 = pre_step(pl->pre[0], &m0);
    const bool more1 = pre_step(pl->pre[1], &m1);
    require(more0 == more1, "ole_deal_gmw: the parties' protocols diverged", kErrInternal);
    if (!more0) break;
    require(m0.send_bytes == m1.recv_bytes && m1.send_bytes == m0.recv_bytes,
            "ole_deal_gmw: message sizes disagree", kErrInternal);
    if (m0.send_bytes)
      hip_check(hipMemcpyAsync(m1.d_recv, m0.d_send, m0.send_bytes, hipMemcpyDeviceToDevice, st),
                "ole_deal_gmw exchange");
    if (m1.send_bytes)
      hip_check(hipMemcpyAsync(m0.d_recv, m1.d_send, m1.send_bytes, hipMemcpyDeviceToDevice, st),
                "ole_deal_gmw exchange");
  }
  pl->dealt = true;
  pl->own_deal = false;
}

// DPF blocks + output layer: leaves g_ij (folded) for both parties in fw.
void plan_run_blocks(pcgsq_ole_plan* pl) {
  require(pl->dealt, "ole: run before deal");
  require(pl->party < 0, "ole: a one-party plan runs with pcgsq_ole_party_run");
  const OleDims& d = pl->d;
  hipStream_t st = pl->ctx->stream;
  const int gb = pl->group_blocks;
  const size_t ni = (size_t)d.ninst;
  const int nparts = leaf_nparts(d);
  const bool fused = fused_gather(d);
  uint64_t* G = pl->fw.as<uint64_t>();                                  // [2][nblk][N]
  const uint64_t* BT = pl->bt.as<uint64_t>();
  for (int b0 = 0; b0 < d.nblk; b0 += gb) {
    const size_t i0 = (size_t)b0 * d.B;   // first instance of the group
    DpfRun r;
    r.ntrees = 2 * gb * d.B;
    r.n = d.D_log;
    r.w = d.w;
    r.prg = d.prg;
    r.roots = pl->roots.as<Blk>() + i0;
    r.roots1 = pl->roots.as<Blk>() + ni + i0;
    r.dltsft = pl->dlt.as<Blk>() + i0 * d.levels * d.nk;
    r.dltsft1 = pl->dlt.as<Blk>() + (ni + i0) * d.levels * d.nk;
    r.aes_t0 = pl->aes_t0.as<uint32_t>();
    r.aes_rks = pl->aes_rks.as<uint8_t>();
    r.partials = pl->part.as<uint64_t>();
    r.prof = &pl->ctx->prof;
    if (fused) {
      // sums of +-C straight into g; tau bits into the mask words
      r.mode = kLeafGather;
      r.g_acc = G + (size_t)b0 * d.N;
      r.g_party_stride = (size_t)d.nblk * d.N;
      r.g_mask = pl->ct.as<uint64_t>();
      r.mask_words = gather_mask_words(d);
      r.t = d.t;
      r.B = d.B;
      r.N_log = d.L;
      r.sets = gb;
    } else {
      r.mode = kLeafConvert;
      r.ct = pl->ct.as<uint64_t>();
    }
    dpf_expand(st, r, pl->dpf);
    {
      ProfZone pzb(&pl->ctx->prof, st, "out.beaver");
      launch_beaver_cw(st, pl->part.as<uint64_t>(), gb * d.B, nparts, BT + i0, BT + 2 * ni + i0,
                       BT + 4 * ni + i0, BT + 6 * ni + i0, ni, pl->cw.as<uint64_t>() + i0);
    }
    if (fused) {
      GatherFixArgs fa;
      fa.g = G + (size_t)b0 * d.N;
      fa.g_party_stride = (size_t)d.nblk * d.N;
      fa.mask = pl->ct.as<uint64_t>();
      fa.mask_words = gather_mask_words(d);
      fa.CW = pl->cw.as<uint64_t>() + i0;
      fa.sets = gb;
      fa.B = d.B;
      fa.t = d.t;
      fa.N_log = d.L;
      fa.bin_log = d.D_log - 1;
      fa.prof = &pl->ctx->prof;
      launch_gather_fix(st, fa);
      continue;
    }
    GatherArgs ga;
    ga.ct = pl->ct.as<uint64_t>();
    ga.set_src_stride = (size_t)d.B * d.D;
    ga.CW = pl->cw.as<uint64_t>() + i0;
    ga.cw_stride = (size_t)d.B;
    ga.g = G + (size_t)b0 * d.N;
    ga.g_party_stride = (size_t)d.nblk * d.N;
    ga.nsets = 2 * gb;
    ga.sets_per_party = gb;
    ga.B = d.B;
    ga.t = d.t;
    ga.N = d.N;
    ga.D_log = d.D_log;
    ga.prof = &pl->ctx->prof;
    launch_gather_g(st, ga);
  }
}

void plan_step4(pcgsq_ole_plan* pl) {
  require(pl->party < 0, "ole: a one-party plan runs with pcgsq_ole_party_run");
  const OleDims& d = pl->d;
  hipStream_t st = pl->ctx->stream;
  uint64_t* G = pl->fw.as<uint64_t>();
  uint64_t* E = G + (size_t)2 * d.nblk * d.N;
  // step 4 in the NTT domain (a fill kernel, not a memset node, so the
  // sequence can be captured into a graph)
  const size_t ne_words = (size_t)2 * d.c * d.N;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((ne_words + 255) / 256)), dim3(256), 0, st, E,
                     ne_words);
  const int ne = 2 * d.c * d.t;
  hipLaunchKernelGGL(build_e_kernel, dim3((d.t + 255) / 256, 2 * d.c), dim3(256), 0, st, E,
                     pl->noise.as<uint64_t>(), pl->noise.as<uint64_t>() + ne, d.c, d.t, d.N);
  hip_check(hipGetLastError(), "launch build_e");
  pl->ctx->ntt.prof = &pl->ctx->prof;
  pl->ctx->ntt.forward(st, G, 2 * d.nblk + 2 * d.c, d.N, kNttFourStep);
  ProfZone pzd(&pl->ctx->prof, st, "ole.step4_dot");
  launch_step4_dot(st, pl->ahat.as<uint64_t>(), G, E, pl->out.as<uint64_t>(), d.c, d.N, 0, 2);
  pl->ctx->ntt.inverse(st, pl->out.as<uint64_t>(), 4, d.N, kNttFourStep);
}

// ---- one party (pcg_ole_impl.h:46-490 as run by one process) --------------

struct PartyExchange {
  pcgsq_ole_exchange_fn fn;
  void* user;
  int B, nk, gb;
  int mode;   // PCGSQ_EXCHANGE_PER_BLOCK or PCGSQ_EXCHANGE_BATCHED
};

void party_call(const PartyExchange& x, const char* phase, int level, int batch, const void* send,
                size_t send_bytes, void* recv, size_t recv_bytes) {
  if (x.fn(x.user, phase, level, batch, send, send_bytes, recv, recv_bytes) != 0)
    throw Error(kErrCallback, std::string("ole: exchange callback failed at ") + phase);
}

// The DPF's per-level exchange of a group of blocks: one call per block, as
// the reference's serial arm exchanges one DPF batch per block
// (dpf.cpp:535-551), or one call for the whole group (its batched arm,
// GPU_baseline/e2e_full_bench.cu:128-153).
int party_dpf_exchange(void* user, int level, const Blk* mine, Blk* peer, size_t count) {
  const PartyExchange* x = static_cast<const PartyExchange*>(user);
  const size_t per = (size_t)x->B * x->nk;
  if (count != per * (size_t)x->gb) return -1;
  if (x->mode == PCGSQ_EXCHANGE_BATCHED)
    return x->fn(x->user, "dpf_gen", level, x->B * x->gb, mine, count * sizeof(Blk), peer,
                 count * sizeof(Blk));
  for (int bb = 0; bb < x->gb; ++bb) {
    const int rc = x->fn(x->user, "dpf_gen", level, x->B, mine + bb * per, per * sizeof(Blk),
                         peer + bb * per, per * sizeof(Blk));
    if (rc != 0) return rc;
  }
  return 0;
}

// One message of the preprocessing protocol through the host: copy out,
// call the exchange (per DPF block when the message has per-block pieces
// and the mode is per block), copy in.
void party_msg(pcgsq_ole_plan* pl, const PartyExchange& x, const PreMsg& m) {
  hipStream_t st = pl->ctx->stream;
  pl->hsend.ensure(std::max<size_t>(m.send_bytes, 1), "ole party staging");
  pl->hrecv.ensure(std::max<size_t>(m.recv_bytes, 1), "ole party staging");
  if (m.send_bytes)
    hip_check(hipMemcpyAsync(pl->hsend.ptr, m.d_send, m.send_bytes, hipMemcpyDeviceToHost, st),
              "ole party message");
  hip_check(hipStreamSynchronize(st), "ole party message");
  const int nb = x.mode == PCGSQ_EXCHANGE_BATCHED ? 1 : m.nblocks;
  require(m.send_bytes % (size_t)nb == 0 && m.recv_bytes % (size_t)nb == 0,
          "ole: message does not split into blocks", kErrInternal);
  const size_t sb = m.send_bytes / nb, rb = m.recv_bytes / nb;
  const int batch = m.batch * (m.nblocks / nb);
  for (int b = 0; b < nb; ++b)
    party_call(x, m.phase, m.level, batch, pl->hsend.as<uint8_t>() + (size_t)b * sb, sb,
               pl->hrecv.as<uint8_t>() + (size_t)b * rb, rb);
  if (m.recv_bytes)
    hip_check(hipMemcpyAsync(m.d_recv, pl->hrecv.ptr, m.recv_bytes, hipMemcpyHostToDevice, st),
              "ole party message");
}

}  // namespace

namespace pcgsq {

// Base OTs of one party through an exchange: two "base_ot" rows (the
// parties' A points, then their 128 B points).
void base_ots_exchange(int party, uint64_t seed, pcgsq_ole_exchange_fn fn, void* user,
                       BaseOts* out) {
  BaseOtParty bo(party, seed);
  const PartyExchange x{fn, user, 0, 0, 1, PCGSQ_EXCHANGE_PER_BLOCK};
  const std::vector<uint8_t> m1 = bo.round1();
  std::vector<uint8_t> p1(BaseOtParty::kMsg1Bytes), p2(BaseOtParty::kMsg2Bytes);
  party_call(x, "base_ot", 0, 1, m1.data(), m1.size(), p1.data(), p1.size());
  const std::vector<uint8_t> m2 = bo.round2(p1.data(), p1.size());
  party_call(x, "base_ot", 1, kKappa, m2.data(), m2.size(), p2.data(), p2.size());
  bo.finish(p2.data(), p2.size(), out);
}

}  // namespace pcgsq

namespace {

// One party's preprocessing (pcg_ole_impl.h:62-199) at the start of its run:
// base OTs the first time, then every message of preproc.hpp's protocol.
void party_preprocess(pcgsq_ole_plan* pl, const PartyExchange& x) {
  const OleDims& d = pl->d;
  const int p = pl->party;
  if (!pl->have_base) {
    std::unique_ptr<BaseOts> b(new BaseOts);
    base_ots_exchange(p, pl->own_seed, x.fn, x.user, b.get());
    pre_prepare(pl->pre[p], make_pre_shape(d.N, d.c, d.t, d.w), p, *b);
    pl->have_base = true;
  }
  pre_begin(pl->pre[p], pl->own_seed, pl->sid++, plan_pre_out(pl), pl->ctx->stream);
  PreMsg m;
  while (pre_step(pl->pre[p], &m)) party_msg(pl, x, m);
}

// This party's blocks: the DPF with the per-level exchange, the sums, the
// two Beaver opens per block on the host (pcg_ole_impl.h:213-295: d = A - r,
// e = B - s; CW share = d s + e r + rs (+ d e for party 0)), and this
// party's g.
void plan_run_party(pcgsq_ole_plan* pl, pcgsq_ole_exchange_fn fn, void* user) {
  require(pl->dealt, "ole: run before deal");
  require(pl->party == 0 || pl->party == 1, "ole: not a one-party plan");
  const OleDims& d = pl->d;
  const int p = pl->party;
  hipStream_t st = pl->ctx->stream;
  const int gb = pl->group_blocks;
  const size_t ni = (size_t)d.ninst, gi = (size_t)gb * d.B;
  const int nparts = leaf_nparts(d);
  const bool fused = fused_gather(d);
  uint64_t* G = pl->fw.as<uint64_t>() + (size_t)p * d.nblk * d.N;  // this party's g [nblk][N]
  const uint64_t* BT = pl->bt.as<uint64_t>();
  PartyExchange px{fn, user, d.B, d.nk, gb, pl->exchange_mode};
  if (pl->own_deal) party_preprocess(pl, px);
  std::vector<uint64_t> sums(2 * gi), bt(4 * gi), cw_full(gi);
  std::vector<uint64_t> de_my(2 * gi), de_peer(2 * gi), cw_my(gi), cw_peer(gi);
  uint64_t* dsums = pl->cw.as<uint64_t>() + ni;  // scratch after the CWs
  for (int b0 = 0; b0 < d.nblk; b0 += gb) {
    const size_t i0 = (size_t)b0 * d.B;
    DpfRun r;
    r.ntrees = (int)gi;
    r.n = d.D_log;
    r.w = d.w;
    r.prg = d.prg;
    r.roots = pl->roots.as<Blk>() + (size_t)p * ni + i0;
    r.dltsft = pl->dlt.as<Blk>() + ((size_t)p * ni + i0) * d.levels * d.nk;
    r.aes_t0 = pl->aes_t0.as<uint32_t>();
    r.aes_rks = pl->aes_rks.as<uint8_t>();
    r.host_ex = party_dpf_exchange;
    r.host_user = &px;
    r.partials = pl->part.as<uint64_t>();
    r.prof = &pl->ctx->prof;
    if (fused) {
      r.mode = kLeafGather;
      r.g_acc = G + (size_t)b0 * d.N;
      r.g_mask = pl->ct.as<uint64_t>();
      r.mask_words = gather_mask_words(d);
      r.t = d.t;
      r.B = d.B;
      r.N_log = d.L;
      r.sets = gb;
      r.parties = 1;
    } else {
      r.mode = kLeafConvert;
      r.ct = pl->ct.as<uint64_t>();
    }
    dpf_expand(st, r, pl->dpf);
    launch_fold_sums(st, pl->part.as<uint64_t>(), (int)gi, nparts, dsums, dsums + gi);
    hip_check(hipMemcpyAsync(sums.data(), dsums, 2 * gi * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, st), "ole party sums");
    for (int k = 0; k < 4; ++k)  // beta, r, s, rs of this party
      hip_check(hipMemcpyAsync(bt.data() + k * gi, BT + (size_t)k * 2 * ni + (size_t)p * ni + i0,
                               gi * sizeof(uint64_t), hipMemcpyDeviceToHost, st),
                "ole party beaver shares");
    hip_check(hipStreamSynchronize(st), "ole party sums");
    const uint64_t* sumC = sums.data();
    const uint64_t* sumT = sums.data() + gi;
    const uint64_t *beta = bt.data(), *rr = bt.data() + gi, *ss = bt.data() + 2 * gi,
                   *rs = bt.data() + 3 * gi;
    // blocks [bb, bb + nbb) of the group: d, e of every instance ([block][d | e]),
    // one "beaver_de" and one "beaver_cw" exchange
    const int nbb = pl->exchange_mode == PCGSQ_EXCHANGE_BATCHED ? gb : 1;
    for (int bb = 0; bb < gb; bb += nbb) {
      const size_t o0 = (size_t)bb * d.B, n = (size_t)nbb * d.B;
      for (int q = 0; q < nbb; ++q)
        for (int k = 0; k < d.B; ++k) {
          const size_t o = o0 + (size_t)q * d.B + k;
          uint64_t A = freduce64(sumT[o]);
          if (p == 1) A = fneg(A);
          uint64_t Bs = sumC[o];
          if (p == 0) Bs = fneg(Bs);
          Bs = fadd(Bs, beta[o]);
          de_my[2 * (size_t)q * d.B + k] = fsub(A, rr[o]);
          de_my[(2 * (size_t)q + 1) * d.B + k] = fsub(Bs, ss[o]);
        }
      party_call(px, "beaver_de", b0 + bb, (int)n, de_my.data(), 2 * n * sizeof(uint64_t),
                 de_peer.data(), 2 * n * sizeof(uint64_t));
      for (int q = 0; q < nbb; ++q)
        for (int k = 0; k < d.B; ++k) {
          const size_t o = o0 + (size_t)q * d.B + k;
          const size_t id = 2 * (size_t)q * d.B + k, ie = id + d.B;
          const uint64_t dv = fadd(de_my[id], de_peer[id]);
          const uint64_t ev = fadd(de_my[ie], de_peer[ie]);
          uint64_t c = fadd(fadd(fmul(dv, ss[o]), fmul(ev, rr[o])), rs[o]);
          if (p == 0) c = fadd(c, fmul(dv, ev));
          cw_my[(size_t)q * d.B + k] = c;
        }
      party_call(px, "beaver_cw", b0 + bb, (int)n, cw_my.data(), n * sizeof(uint64_t),
                 cw_peer.data(), n * sizeof(uint64_t));
      for (size_t k = 0; k < n; ++k) cw_full[o0 + k] = fadd(cw_my[k], cw_peer[k]);
    }
    hip_check(hipMemcpyAsync(pl->cw.as<uint64_t>() + i0, cw_full.data(), gi * sizeof(uint64_t),
                             hipMemcpyHostToDevice, st), "ole party cw");
    if (fused) {
      GatherFixArgs fa;
      fa.g = G + (size_t)b0 * d.N;
      fa.mask = pl->ct.as<uint64_t>();
      fa.mask_words = gather_mask_words(d);
      fa.CW = pl->cw.as<uint64_t>() + i0;
      fa.sets = gb;
      fa.B = d.B;
      fa.t = d.t;
      fa.N_log = d.L;
      fa.bin_log = d.D_log - 1;
      fa.parties = 1;
      fa.first_party = p;
      fa.prof = &pl->ctx->prof;
      launch_gather_fix(st, fa);
    } else {
      GatherArgs ga;
      ga.ct = pl->ct.as<uint64_t>();
      ga.set_src_stride = (size_t)d.B * d.D;
      ga.CW = pl->cw.as<uint64_t>() + i0;
      ga.cw_stride = (size_t)d.B;
      ga.g = G + (size_t)b0 * d.N;
      ga.nsets = gb;
      ga.sets_per_party = gb;
      ga.first_party = p;
      ga.B = d.B;
      ga.t = d.t;
      ga.N = d.N;
      ga.D_log = d.D_log;
      ga.prof = &pl->ctx->prof;
      launch_gather_g(st, ga);
    }
  }
  // step 4 for this party: x = sum_i a_i e_i, z = sum_ij a_i a_j g_ij
  uint64_t* E = pl->fw.as<uint64_t>() + (size_t)2 * d.nblk * d.N + (size_t)p * d.c * d.N;
  const size_t ne_words = (size_t)d.c * d.N;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((ne_words + 255) / 256)), dim3(256), 0, st, E,
                     ne_words);
  const int ne = 2 * d.c * d.t;
  hipLaunchKernelGGL(build_e_kernel, dim3((d.t + 255) / 256, d.c), dim3(256), 0, st, E,
                     pl->noise.as<uint64_t>() + (size_t)p * d.c * d.t,
                     pl->noise.as<uint64_t>() + ne + (size_t)p * d.c * d.t, d.c, d.t, d.N);
  hip_check(hipGetLastError(), "launch build_e");
  pl->ctx->ntt.prof = &pl->ctx->prof;
  pl->ctx->ntt.forward(st, G, d.nblk, d.N, kNttFourStep);
  pl->ctx->ntt.forward(st, E, d.c, d.N, kNttFourStep);
  launch_step4_dot(st, pl->ahat.as<uint64_t>(), pl->fw.as<uint64_t>(),
                   pl->fw.as<uint64_t>() + (size_t)2 * d.nblk * d.N, pl->out.as<uint64_t>(), d.c,
                   d.N, p, 1);
  pl->ctx->ntt.inverse(st, pl->out.as<uint64_t>() + (size_t)2 * p * d.N, 2, d.N, kNttFourStep);
  hip_check(hipStreamSynchronize(st), "ole party run");
}

}  // namespace

extern "C" {

int pcgsq_ole_plan_create(pcgsq_ctx* ctx, const pcgsq_params* p, uint64_t pub_seed,
                          size_t hbm_budget_bytes, pcgsq_ole_plan** out) {
  return guarded([&] {
    require(ctx && p && out, "ole_plan_create: null argument");
    CtxGuard g(ctx);
    std::unique_ptr<pcgsq_ole_plan> pl(new pcgsq_ole_plan);
    pl->ctx = ctx;
    pl->d = make_dims(*p);
    pl->pub_seed = pub_seed;
    plan_alloc(pl.get(), hbm_budget_bytes);
    *out = pl.release();
  });
}

int pcgsq_ole_party_create(pcgsq_ctx* ctx, const pcgsq_params* p, int party, uint64_t pub_seed,
                           size_t hbm_budget_bytes, pcgsq_ole_plan** out) {
  return guarded([&] {
    require(ctx && p && out, "ole_party_create: null argument");
    require(party == 0 || party == 1, "ole_party_create: party must be 0 or 1");
    CtxGuard g(ctx);
    std::unique_ptr<pcgsq_ole_plan> pl(new pcgsq_ole_plan);
    pl->ctx = ctx;
    pl->d = make_dims(*p);
    pl->pub_seed = pub_seed;
    pl->party = party;
    plan_alloc(pl.get(), hbm_budget_bytes);
    *out = pl.release();
  });
}

int pcgsq_ole_party_run(pcgsq_ole_plan* plan, pcgsq_ole_exchange_fn fn, void* user) {
  return guarded([&] {
    require(plan && fn, "ole_party_run: null argument");
    CtxGuard g(plan->ctx);
    plan_run_party(plan, fn, user);
  });
}

int pcgsq_ole_party_outputs(pcgsq_ole_plan* plan, uint64_t* x, uint64_t* z) {
  return guarded([&] {
    require(plan, "ole_party_outputs: null plan");
    require(plan->party == 0 || plan->party == 1, "ole_party_outputs: not a one-party plan");
    CtxGuard g(plan->ctx);
    const size_t N = (size_t)plan->d.N;
    const uint64_t* o = plan->out.as<uint64_t>() + (size_t)2 * plan->party * N;
    if (x)
      hip_check(hipMemcpyAsync(x, o, N * 8, hipMemcpyDeviceToHost, plan->ctx->stream),
                "ole party outputs");
    if (z)
      hip_check(hipMemcpyAsync(z, o + N, N * 8, hipMemcpyDeviceToHost, plan->ctx->stream),
                "ole party outputs");
    hip_check(hipStreamSynchronize(plan->ctx->stream), "ole party outputs");
  });
}

int pcgsq_ole_plan_destroy(pcgsq_ole_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    CtxGuard g(plan->ctx);
    hip_check(hipStreamSynchronize(plan->ctx->stream), "ole destroy");
    delete plan;
  });
}

size_t pcgsq_ole_plan_estimate(const pcgsq_params* p, size_t hbm_budget_bytes) {
  size_t r = 0;
  const int rc = guarded([&] {
    require(p, "ole_plan_estimate: null params");
    r = plan_estimate(make_dims(*p), hbm_budget_bytes);
  });
  return rc == PCGSQ_OK ? r : 0;
}

size_t pcgsq_ole_plan_bytes(const pcgsq_ole_plan* pl) {
  if (!pl) return 0;
  size_t b = 0;
  for (const DevBuf* x : {&pl->roots, &pl->dlt, &pl->bt, &pl->noise, &pl->aes_t0, &pl->aes_rks,
                          &pl->ahat, &pl->ct, &pl->part, &pl->cw, &pl->fw, &pl->out, &pl->tri,
                          &pl->dpf.fh[0], &pl->dpf.fh[1], &pl->dpf.cw, &pl->dpf.peer})
    b += x->bytes;
  return b;
}

int pcgsq_ole_deal(pcgsq_ole_plan* plan, uint64_t seed) {
  return guarded([&] {
    require(plan, "ole_deal: null plan");
    CtxGuard g(plan->ctx);
    plan_deal(plan, seed);
  });
}

int pcgsq_ole_deal_gmw(pcgsq_ole_plan* plan, uint64_t seed) {
  return guarded([&] {
    require(plan, "ole_deal_gmw: null plan");
    CtxGuard g(plan->ctx);
    plan_deal_gmw(plan, seed);
  });
}

int pcgsq_ole_party_deal_own(pcgsq_ole_plan* plan, uint64_t own_seed) {
  return guarded([&] {
    require(plan, "ole_party_deal_own: null plan");
    require(plan->party == 0 || plan->party == 1, "ole_party_deal_own: not a one-party plan");
    (void)make_pre_shape(plan->d.N, plan->d.c, plan->d.t, plan->d.w);   // shape check
    plan->own_seed = own_seed;
    plan->own_deal = true;
    plan->dealt = true;
  });
}

int pcgsq_ole_party_set_exchange_mode(pcgsq_ole_plan* plan, int mode) {
  return guarded([&] {
    require(plan, "ole_party_set_exchange_mode: null plan");
    require(mode == PCGSQ_EXCHANGE_PER_BLOCK || mode == PCGSQ_EXCHANGE_BATCHED,
            "ole_party_set_exchange_mode: unknown mode");
    plan->exchange_mode = mode;
  });
}

int pcgsq_ole_preprocessing_read(pcgsq_ole_plan* plan, int party, int what, void* host,
                                 size_t bytes, size_t* needed) {
  return guarded([&] {
    require(plan && needed, "ole_preprocessing_read: null argument");
    require(party == 0 || party == 1, "ole_preprocessing_read: party must be 0 or 1");
    require(plan->party < 0 || plan->party == party,
            "ole_preprocessing_read: this plan does not hold that party");
    const PreRun& r = plan->pre[party];
    require(r.stage >= 0, "ole_preprocessing_read: no preprocessing ran for that party");
    const PreShape& sh = r.sh;
    const size_t ni = (size_t)sh.ninst, le = (size_t)sh.levels * sh.nk;
    const size_t mr = sh.m[1 - party], ms = sh.m[party];
    const DevBuf* src = nullptr;
    size_t n = 0;
    switch (what) {
      case PCGSQ_PRE_CHOICE: src = &r.choice; n = mr / 8; break;
      case PCGSQ_PRE_RECEIVED: src = &r.mc; n = mr * sizeof(Blk); break;
      case PCGSQ_PRE_SENT0: src = &r.m0; n = ms * sizeof(Blk); break;
      case PCGSQ_PRE_SENT1: src = &r.m1; n = ms * sizeof(Blk); break;
      case PCGSQ_PRE_ADDER_TRIPLES: src = &r.tri_add; n = ni * sh.ops * 3 * 8; break;
      case PCGSQ_PRE_DELTA_TRIPLES: src = &r.tri_dlt; n = ni * le * 5 * 8; break;
      case PCGSQ_PRE_TREE_TRIPLES: src = &r.tri_tree; n = sh.w >= 2 ? ni * le * 4 : 0; break;
      default: require(false, "ole_preprocessing_read: unknown item");
    }
    *needed = n;
    if (!host) return;
    require(bytes >= n, "ole_preprocessing_read: host buffer too small");
    CtxGuard g(plan->ctx);
    if (n)
      hip_check(hipMemcpyAsync(host, src->ptr, n, hipMemcpyDeviceToHost, plan->ctx->stream),
                "ole preprocessing read");
    hip_check(hipStreamSynchronize(plan->ctx->stream), "ole preprocessing read");
  });
}

int pcgsq_base_ots_run(int party, uint64_t seed, pcgsq_ole_exchange_fn fn, void* user,
                       pcgsq_base_ots* out) {
  return guarded([&] {
    require(fn && out, "base_ots_run: null argument");
    require(party == 0 || party == 1, "base_ots_run: party must be 0 or 1");
    static_assert(sizeof(pcgsq_base_ots) == sizeof(BaseOts), "pcgsq_base_ots layout");
    std::unique_ptr<BaseOts> b(new BaseOts);
    base_ots_exchange(party, seed, fn, user, b.get());
    memcpy(out, b.get(), sizeof(BaseOts));
  });
}

int pcgsq_ole_dealt_dltsft(pcgsq_ole_plan* plan, pcgsq_blk* dltsft) {
  return guarded([&] {
    require(plan && dltsft, "ole_dealt_dltsft: null argument");
    require(plan->dealt, "ole_dealt_dltsft: nothing dealt");
    CtxGuard g(plan->ctx);
    const OleDims& d = plan->d;
    const size_t n = (size_t)2 * d.ninst * d.levels * d.nk;
    hip_check(hipMemcpyAsync(dltsft, plan->dlt.ptr, n * sizeof(Blk), hipMemcpyDeviceToHost,
                             plan->ctx->stream),
              "ole dltsft");
    hip_check(hipStreamSynchronize(plan->ctx->stream), "ole dltsft");
  });
}

int pcgsq_ole_dealt_products(pcgsq_ole_plan* plan, uint64_t* out) {
  return guarded([&] {
    require(plan && out, "ole_dealt_products: null argument");
    require(plan->dealt, "ole_dealt_products: nothing dealt");
    CtxGuard g(plan->ctx);
    const size_t n = (size_t)8 * plan->d.ninst;
    hip_check(hipMemcpyAsync(out, plan->bt.ptr, n * sizeof(uint64_t), hipMemcpyDeviceToHost,
                             plan->ctx->stream),
              "ole products");
    hip_check(hipStreamSynchronize(plan->ctx->stream), "ole products");
  });
}

int pcgsq_ole_run(pcgsq_ole_plan* plan) {
  return guarded([&] {
    require(plan, "ole_run: null plan");
    CtxGuard g(plan->ctx);
    plan_run_blocks(plan);
    plan_step4(plan);
  });
}

int pcgsq_ole_run_blocks(pcgsq_ole_plan* plan) {
  return guarded([&] {
    require(plan, "ole_run_blocks: null plan");
    CtxGuard g(plan->ctx);
    plan_run_blocks(plan);
  });
}

int pcgsq_ole_outputs(pcgsq_ole_plan* plan, uint64_t* x0, uint64_t* z0, uint64_t* x1,
                      uint64_t* z1) {
  return guarded([&] {
    require(plan, "ole_outputs: null plan");
    require(plan->party < 0, "ole_outputs: a one-party plan has pcgsq_ole_party_outputs");
    CtxGuard g(plan->ctx);
    const size_t N = (size_t)plan->d.N;
    uint64_t* dst[4] = {x0, z0, x1, z1};
    for (int i = 0; i < 4; ++i)
      if (dst[i])
        hip_check(hipMemcpyAsync(dst[i], plan->out.as<uint64_t>() + i * N, N * 8,
                                 hipMemcpyDeviceToHost, plan->ctx->stream), "ole outputs");
    hip_check(hipStreamSynchronize(plan->ctx->stream), "ole outputs");
  });
}

const uint64_t* pcgsq_ole_outputs_dev(pcgsq_ole_plan* plan) {
  return plan ? plan->out.as<uint64_t>() : nullptr;
}

int pcgsq_triples_generate(pcgsq_ole_plan* plan, uint64_t seed1, uint64_t seed2, uint64_t* a0,
                           uint64_t* b0, uint64_t* c0, uint64_t* a1, uint64_t* b1,
                           uint64_t* c1) {
  return guarded([&] {
    require(plan, "triples_generate: null plan");
    CtxGuard g(plan->ctx);
    const OleDims& d = plan->d;
    const size_t n = (size_t)d.N;
    hipStream_t st = plan->ctx->stream;
    plan->tri.ensure(10 * n * sizeof(uint64_t), "hipMalloc triples");
    uint64_t* O1 = plan->tri.as<uint64_t>();
    uint64_t* T = O1 + 4 * n;
    ProfZone pz(&plan->ctx->prof, st, "triples");
    plan_deal(plan, seed1);
    plan_run_blocks(plan);
    plan_step4(plan);
    hip_check(hipMemcpyAsync(O1, plan->out.ptr, 4 * n * 8, hipMemcpyDeviceToDevice, st),
              "triples copy");
    plan_deal(plan, seed2);
    plan_run_blocks(plan);
    plan_step4(plan);
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(triple_compose_kernel, grid, dim3(256), 0, st, O1,
                       plan->out.as<uint64_t>(), T, d.N);
    plan->ctx->ntt.forward(st, T, 6, d.N, kNttFourStep);
    hipLaunchKernelGGL(triple_finish_kernel, grid, dim3(256), 0, st, T, d.N);
    hip_check(hipGetLastError(), "launch triples");
    uint64_t* dst[6] = {a0, b0, a1, b1, c0, c1};
    for (int k = 0; k < 6; ++k)
      if (dst[k])
        hip_check(hipMemcpyAsync(dst[k], T + k * n, n * 8, hipMemcpyDeviceToHost, st),
                  "triples out");
    hip_check(hipStreamSynchronize(st), "triples");
  });
}

int pcgsq_ole_g(pcgsq_ole_plan* plan, uint64_t* g_host) {
  return guarded([&] {
    require(plan && g_host, "ole_g: null argument");
    CtxGuard g(plan->ctx);
    const size_t n = (size_t)2 * plan->d.nblk * plan->d.N;
    hip_check(hipMemcpyAsync(g_host, plan->fw.ptr, n * 8, hipMemcpyDeviceToHost, plan->ctx->stream),
              "ole g");
    hip_check(hipStreamSynchronize(plan->ctx->stream), "ole g");
  });
}

int pcgsq_ole_expand(pcgsq_ctx* ctx, const pcgsq_params* p, uint64_t seed, uint64_t pub_seed,
                     uint64_t* x0, uint64_t* z0, uint64_t* x1, uint64_t* z1) {
  pcgsq_ole_plan* pl = nullptr;
  int rc = pcgsq_ole_plan_create(ctx, p, pub_seed, 0, &pl);
  if (rc) return rc;
  rc = pcgsq_ole_deal(pl, seed);
  if (!rc) rc = pcgsq_ole_run(pl);
  if (!rc) rc = pcgsq_ole_outputs(pl, x0, z0, x1, z1);
  const int rc2 = pcgsq_ole_plan_destroy(pl);
  return rc ? rc : rc2;
}

uint64_t pcgsq_dealer_u64(uint64_t seed, uint32_t domain, uint64_t a, uint64_t b) {
  return dealer_u64(seed, domain, a, b);
}

}  // extern "C"
