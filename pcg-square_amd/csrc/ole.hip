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
// Preprocessing as the parties run it (pcgsq_ole_deal_gmw): the same draws,
// except that
//   * alpha and DltSft come from the parties' Kogge-Stone adder and
//     F_DeltaShiftShare circuits under GMW, with masks from U(seed,14,g,j)
//     and every AND triple made from two OTs (spec at gmw_and below);
//   * beta and rs come from Gilboa OT products (spec at gilboa below).
// The OTs are ideal (the receiver gets the chosen message) where the
// reference runs FerretCOT; oracle/orc_gmw.c restates all of it.
#include <string.h>

#include <algorithm>
#include <vector>

#include "context.hpp"
#include "pcgsq.h"

namespace pcgsq {

namespace {

PCGSQ_HD uint64_t sm64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

PCGSQ_HD uint64_t dealer_hd(uint64_t seed, uint32_t domain, uint64_t a, uint64_t b) {
  return sm64(sm64(sm64(seed ^ ((uint64_t)domain * 0xD6E8FEB86659FD93ULL)) ^ a) ^ b);
}

constexpr uint64_t kWlo = 0x0123456789abcdefULL, kWhi = 0xdeadbeefcafe0001ULL;

// ---- GMW preprocessing (pcg_ole_impl.h:114-199), both parties per thread ----
// XOR-shared words: bit i of (s0, s1) are the two parties' shares of wire i
// of a bit-sliced gate layer.  Randomness is dealer domain 14, word
// M(g, j) = U(seed, 14, g, j) for global instance g: input masks j = 0..7,
// AND op o takes a0, a1, b0, b1, c0 = M(8 + 5o + 0..4), c1 = (a0^a1)&(b0^b1)^c0
// (a gate in bit i of a layer uses bit i of each word).
struct Sh {
  uint64_t s0, s1;
};

// GMW AND (bool_circuit/circuit.cpp:245-256): open d = x ^ a, e = y ^ b;
// party 0 gets d&e ^ d&b0 ^ e&a0 ^ c0, party 1 d&b1 ^ e&a1 ^ c1.  The
// triple is made from two OTs as TripleGen::generate (circuit.cpp:30-85)
// makes it: each party draws its a, b; party 0, the sender, draws r, s and
// offers (r, r ^ a0) and (s, s ^ b0); party 1 chooses with b1 and a1.  Words
// of op o: a0, a1, b0, b1, r, s = M(8 + 6o + 0..5).
__device__ __forceinline__ Sh gmw_and(Sh x, Sh y, uint64_t seed, uint64_t g, uint32_t op) {
  const uint64_t j = 8 + 6 * (uint64_t)op;
  const uint64_t a0 = dealer_hd(seed, 14, g, j), a1 = dealer_hd(seed, 14, g, j + 1);
  const uint64_t b0 = dealer_hd(seed, 14, g, j + 2), b1 = dealer_hd(seed, 14, g, j + 3);
  const uint64_t r = dealer_hd(seed, 14, g, j + 4), s = dealer_hd(seed, 14, g, j + 5);
  const uint64_t ot1 = r ^ (a0 & b1), ot2 = s ^ (b0 & a1);   // what party 1 receives
  const uint64_t c0 = (a0 & b0) ^ r ^ s, c1 = (a1 & b1) ^ ot1 ^ ot2;
  const uint64_t d = x.s0 ^ a0 ^ x.s1 ^ a1, e = y.s0 ^ b0 ^ y.s1 ^ b1;
  return Sh{(d & e) ^ (d & b0) ^ (e & a0) ^ c0, (d & b1) ^ (e & a1) ^ c1};
}

// The Kogge-Stone adder (bool_circuit/adder.h:21-63) on nb-bit offsets owned
// by party 0 (off0) and party 1 (off1), one bit-sliced layer per AND depth:
// op 0 the generate ANDs, prefix step s (distance d) ops 2s-1 = P & G<<d and
// 2s = P & P<<d over bits >= d.  Returns XOR shares of the nb+1 sum bits.
__device__ Sh gmw_adder(uint64_t off0, uint64_t off1, int nb, uint64_t seed, uint64_t g) {
  const uint64_t mask = (1ULL << nb) - 1;
  const uint64_t m0 = dealer_hd(seed, 14, g, 0) & mask, m1 = dealer_hd(seed, 14, g, 1) & mask;
  const Sh A{off0 ^ m0, m0}, B{m1, off1 ^ m1};  // inputs shared by their owners
  Sh G = gmw_and(A, B, seed, g, 0);
  G.s0 &= mask;
  G.s1 &= mask;
  Sh P{A.s0 ^ B.s0, A.s1 ^ B.s1};
  const Sh P0 = P;
  uint32_t s = 1;
  for (int d = 1; d < nb; d <<= 1, ++s) {
    const uint64_t hi = mask & ~((1ULL << d) - 1);
    const Sh X = gmw_and(P, Sh{G.s0 << d, G.s1 << d}, seed, g, 2 * s - 1);
    const Sh Y = gmw_and(P, Sh{P.s0 << d, P.s1 << d}, seed, g, 2 * s);
    G.s0 ^= X.s0 & hi;
    G.s1 ^= X.s1 & hi;
    P.s0 = (P.s0 & ~hi) | (Y.s0 & hi);
    P.s1 = (P.s1 & ~hi) | (Y.s1 & hi);
  }
  return Sh{P0.s0 ^ (G.s0 << 1), P0.s1 ^ (G.s1 << 1)};
}

// The F_DeltaShiftShare circuit (half_tree_dpf/dpf.cpp:172-240) on the alpha
// shares and the Delta shares (each party's input shared by its masks),
// writing both parties' DltSft shares [levels][nk].  Level i, child k uses
// ops base = aops + (i*nk + k)*(w+1): its AND-tree gates base + j (bit 0),
// its 128 Delta ANDs base + w-1 (bits 0..63) and base + w (bits 64..127).
__device__ void gmw_dltsft(Sh alpha, Blk d0, Blk d1, int n, int w, int aops, uint64_t seed,
                           uint64_t g, Blk* __restrict__ out0, Blk* __restrict__ out1) {
  const int nk = (1 << w) - 1, levels = n / w;
  const uint64_t m2 = dealer_hd(seed, 14, g, 2), m3 = dealer_hd(seed, 14, g, 3);
  const Sh A{(alpha.s0 ^ m2) ^ m3, m2 ^ (alpha.s1 ^ m3)};
  const uint64_t m4 = dealer_hd(seed, 14, g, 4), m5 = dealer_hd(seed, 14, g, 5);
  const uint64_t m6 = dealer_hd(seed, 14, g, 6), m7 = dealer_hd(seed, 14, g, 7);
  const Sh Xlo{d0.lo ^ m4 ^ m6, m4 ^ d1.lo ^ m6}, Xhi{d0.hi ^ m5 ^ m7, m5 ^ d1.hi ^ m7};
  for (int i = 0; i < levels; ++i)
    for (int k = 0; k < nk; ++k) {
      const uint32_t base = (uint32_t)(aops + (i * nk + k) * (w + 1));
      Sh f[4];
      for (int c = 0; c < w; ++c) {
        const int pos = n - w * (i + 1) + c;
        f[c] = Sh{(A.s0 >> pos) & 1, (A.s1 >> pos) & 1};
        if (!((k >> c) & 1)) f[c].s0 ^= 1;  // NOT: party 0 flips its share
      }
      int nf = w, tree = 0;
      while (nf > 1) {  // pairwise, odd one carried (dpf.cpp:219-226)
        int nn = 0;
        for (int j = 0; j + 1 < nf; j += 2) {
          const Sh z = gmw_and(f[j], f[j + 1], seed, g, base + (uint32_t)tree++);
          f[nn++] = Sh{z.s0 & 1, z.s1 & 1};
        }
        if (nf & 1) f[nn++] = f[nf - 1];
        nf = nn;
      }
      const Sh I{0ULL - f[0].s0, 0ULL - f[0].s1};  // the indicator in every bit
      const Sh lo = gmw_and(I, Xlo, seed, g, base + (uint32_t)w - 1);
      const Sh hi = gmw_and(I, Xhi, seed, g, base + (uint32_t)w);
      out0[i * nk + k] = Blk{lo.s0, hi.s0};
      out1[i * nk + k] = Blk{lo.s1, hi.s1};
    }
}

// Gilboa's OT product (TripleGen::zp_multiply, circuit.cpp:91-153): party 0,
// the sender, holds x; party 1 holds y < P.  For bit j < 62 of y the sender
// offers (r_j, r_j + x 2^j mod P), r_j = U(seed, 15, g, 64q + j) % P for
// product q of instance g, and the receiver takes the one its bit selects.
// Shares: party 0 P - sum r_j, party 1 the sum of what it took.
__device__ Sh gilboa(uint64_t x, uint64_t y, uint64_t seed, uint64_t g, uint32_t q) {
  uint64_t rsum = 0, got = 0, xs = x;
  for (int j = 0; j < 62; ++j) {
    const uint64_t r = dealer_hd(seed, 15, g, 64 * (uint64_t)q + (uint64_t)j) % kP;
    rsum = fadd(rsum, r);
    got = fadd(got, (y >> j) & 1 ? fadd(r, xs) : r);
    xs = fadd(xs, xs);
  }
  return Sh{fsub(0, rsum), got};
}

// AND ops of the nb-bit adder: the generate layer and two per prefix step.
int gmw_adder_ops(int nb) {
  int s = 0;
  for (int d = 1; d < nb; d <<= 1) ++s;
  return 1 + 2 * s;
}

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
                                 int D_log, uint64_t bin, int ninst, int gmw, int gmw_aops) {
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
  if (gmw) {
    // alpha and DltSft by the parties' GMW circuits (pcg_ole_impl.h:114-199)
    const Sh alpha = gmw_adder(p0 & (bin - 1), p1 & (bin - 1), D_log - 1, seed, G);
    gmw_dltsft(alpha, d0, d1, D_log, w, gmw_aops, seed, G, dlt + (size_t)g * levels * nk,
               dlt + (ni + g) * levels * nk);
  } else
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
  if (gmw) {
    // beta by a Gilboa product (pcg_ole_impl.h:158-171); the Beaver triple's
    // rs by zp_triple's two cross products (circuit.cpp:156-200)
    const Sh b = gilboa(pay0, pay1, seed, G, 0);
    BETA[g] = b.s0;
    BETA[ni + g] = b.s1;
    const Sh x1 = gilboa(r0, s1, seed, G, 1), x2 = gilboa(s0, r1, seed, G, 2);
    RS[g] = fadd(fadd(fmul(r0, s0), x1.s0), x2.s0);
    RS[ni + g] = fadd(fadd(fmul(r1, s1), x1.s1), x2.s1);
  } else {
    BETA[g] = dealer_hd(seed, 5, G, 0) % kP;
    BETA[ni + g] = fsub(fmul(pay0, pay1), BETA[g]);
    RS[g] = dealer_hd(seed, 8, G, 0) % kP;
    RS[ni + g] = fsub(fmul(fadd(r0, r1), fadd(s0, s1)), RS[g]);
  }
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

// X_s = sum_i A_i E_{s,i};  Z_s = sum_i A_i sum_j A_j G_{s,ij}  (NTT domain)
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

// parties: 2 when one plan runs both parties, 1 for one party.
size_t plan_group_bytes(const OleDims& d, int gb, int parties) {
  const size_t trees = (size_t)parties * gb * d.B;
  const size_t maxS = (size_t)1 << (d.w * (d.levels - 1));
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
  dpf_reserve(pl->dpf, d.D_log, d.w, gtrees, parties == 1);
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

// gmw: alpha and DltSft by the GMW circuits instead of the dealer's formula
// (same Delta shares and positions; DltSft shares differ, their XOR agrees).
void plan_deal(pcgsq_ole_plan* pl, uint64_t seed, bool gmw = false) {
  const OleDims& d = pl->d;
  hipStream_t st = pl->ctx->stream;
  const int t_log = ilog((uint64_t)d.t);
  require(t_log >= 0, "ole: t must be a power of two");
  hipLaunchKernelGGL(deal_noise_kernel, dim3((unsigned)((d.c * d.t + 255) / 256), 2), dim3(256),
                     0, st, pl->noise.as<uint64_t>(), seed, d.c, t_log, (uint64_t)d.bin);
  hipLaunchKernelGGL(deal_inst_kernel, dim3((unsigned)((d.B + 255) / 256), d.c, d.c), dim3(256),
                     0, st, pl->noise.as<uint64_t>(), pl->roots.as<Blk>(), pl->dlt.as<Blk>(),
                     pl->bt.as<uint64_t>(), seed, d.c, t_log, d.levels, d.nk, d.w, d.D_log,
                     (uint64_t)d.bin, d.ninst, gmw ? 1 : 0, gmw_adder_ops(d.D_log - 1));
  hip_check(hipGetLastError(), "launch dealer");
  pl->dealt = true;
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
  uint64_t* E = G + (size_t)2 * d.nblk * d.N;                          // [2][c][N]
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
  hipLaunchKernelGGL(step4_dot_kernel, dim3((d.N + 255) / 256, 2), dim3(256), 0, st,
                     pl->ahat.as<uint64_t>(), G, E, pl->out.as<uint64_t>(), d.c, d.N, 0);
  hip_check(hipGetLastError(), "launch step4_dot");
  pl->ctx->ntt.inverse(st, pl->out.as<uint64_t>(), 4, d.N, kNttFourStep);
}

// ---- one party (pcg_ole_impl.h:46-490 as run by one process) --------------

struct PartyExchange {
  pcgsq_ole_exchange_fn fn;
  void* user;
  int B, nk, gb;
};

// The DPF's per-level exchange of a group of blocks, one call per block as
// the reference exchanges one DPF batch per block (dpf.cpp:535-551).
int party_dpf_exchange(void* user, int level, const Blk* mine, Blk* peer, size_t count) {
  const PartyExchange* x = static_cast<const PartyExchange*>(user);
  const size_t per = (size_t)x->B * x->nk;
  if (count != per * (size_t)x->gb) return -1;
  for (int bb = 0; bb < x->gb; ++bb) {
    const int rc = x->fn(x->user, "dpf_gen", level, x->B, mine + bb * per, peer + bb * per,
                         per * sizeof(Blk));
    if (rc != 0) return rc;
  }
  return 0;
}

void party_exchange(const PartyExchange& x, const char* phase, int level, const void* mine,
                    void* peer, size_t bytes) {
  if (x.fn(x.user, phase, level, x.B, mine, peer, bytes) != 0)
    throw Error(kErrCallback, std::string("ole: exchange callback failed at ") + phase);
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
  PartyExchange px{fn, user, d.B, d.nk, gb};
  std::vector<uint64_t> sums(2 * gi), bt(4 * gi), cw_full(gi);
  std::vector<uint64_t> de_my(2 * (size_t)d.B), de_peer(2 * (size_t)d.B);
  std::vector<uint64_t> cw_my(d.B), cw_peer(d.B);
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
    for (int bb = 0; bb < gb; ++bb) {
      const size_t o = (size_t)bb * d.B;
      for (int k = 0; k < d.B; ++k) {
        uint64_t A = freduce64(sumT[o + k]);
        if (p == 1) A = fneg(A);
        uint64_t Bs = sumC[o + k];
        if (p == 0) Bs = fneg(Bs);
        Bs = fadd(Bs, beta[o + k]);
        de_my[k] = fsub(A, rr[o + k]);
        de_my[d.B + k] = fsub(Bs, ss[o + k]);
      }
      party_exchange(px, "beaver_de", b0 + bb, de_my.data(), de_peer.data(),
                     de_my.size() * sizeof(uint64_t));
      for (int k = 0; k < d.B; ++k) {
        const uint64_t dv = fadd(de_my[k], de_peer[k]);
        const uint64_t ev = fadd(de_my[d.B + k], de_peer[d.B + k]);
        uint64_t c = fadd(fadd(fmul(dv, ss[o + k]), fmul(ev, rr[o + k])), rs[o + k]);
        if (p == 0) c = fadd(c, fmul(dv, ev));
        cw_my[k] = c;
      }
      party_exchange(px, "beaver_cw", b0 + bb, cw_my.data(), cw_peer.data(),
                     cw_my.size() * sizeof(uint64_t));
      for (int k = 0; k < d.B; ++k) cw_full[o + k] = fadd(cw_my[k], cw_peer[k]);
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
  hipLaunchKernelGGL(step4_dot_kernel, dim3((d.N + 255) / 256, 1), dim3(256), 0, st,
                     pl->ahat.as<uint64_t>(), pl->fw.as<uint64_t>(),
                     pl->fw.as<uint64_t>() + (size_t)2 * d.nblk * d.N, pl->out.as<uint64_t>(),
                     d.c, d.N, p);
  hip_check(hipGetLastError(), "launch step4_dot");
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
    require(plan->d.D_log >= 1 && plan->d.D_log <= 63, "ole_deal_gmw: domain out of range");
    CtxGuard g(plan->ctx);
    plan_deal(plan, seed, true);
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
