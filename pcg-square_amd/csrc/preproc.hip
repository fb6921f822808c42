// One party's 2PC preprocessing on the device (see preproc.hpp): own draws,
// IKNP random OTs, AND triples, Gilboa products and the GMW evaluation of
// the position adder and the DltSft circuit, message by message.
//
// Own draws (party p, own seed S; the dealer's formulas with S as the seed,
// so that two parties whose seeds are equal draw what pcgsq_ole_deal(S)
// deals them):
//   pos[p][i][k] = k*bin + U(S,1,p,i*t+k) % bin, pay[p][i][k] = U(S,2,p,i*t+k) % (P-1) + 1,
//   Delta_p = (U(S,3,p,2g), U(S,3,p,2g+1)) with lsb 1 (p = 0) / 0 (p = 1),
//   r_p = U(S,6,g,p) % P, s_p = U(S,7,g,p) % P,
//   the receiver's random choice words U(S,16,dir,w) (its a-bits of the
//   AND triples), the base-OT scalars and choice bits (domain 17).
//
// AND triples from random OTs, one per cross term of
// (a0 ^ a1)(b0 ^ b1) = a0 b0 ^ a1 b1 ^ a0 b1 ^ a1 b0 (the construction of
// TripleGen::generate, circuit.cpp:30-85, with a random OT in place of each
// chosen-message OT): in direction p party p is the sender of random
// messages (m0, m1) and defines b_p = m0 ^ m1; party 1 - p chooses with its
// random bit a_{1-p} and receives m0 ^ a_{1-p} b_p.  So each party holds
//   c_p = a_p b_p ^ m0 (its sender output) ^ m_{a_p} (its receiver output).
// The 128 ANDs I & Delta_j of one DltSft child share their input I: they
// use one triple (a, b_j, c_j) with a common a, made from one OT per
// direction with 128-bit messages, so their d = I ^ a is opened once.
//
// Gilboa (zp_multiply, circuit.cpp:91-153): party 0 holds x, party 1 holds y;
// for bit j < 62 of y a random OT of direction 0 gives party 0 (m0, m1) and
// party 1 m_{y_j}; party 0 takes r_j = m0 and sends delta_j = r_j + x 2^j - m1
// so that party 1 holds r_j + y_j x 2^j; shares -sum r_j and the sum of what
// party 1 holds add up to x y.  Products: beta = pay0 pay1, and zp_triple's
// cross terms r0 s1 and s0 r1 (circuit.cpp:156-200).
//
// GMW (circuit.cpp:245-380): XOR shares; a party's own input is shared as
// (its value, 0); NOT flips party 0's share; an AND opens d = x ^ a and
// e = y ^ b and sets z_p = [p = 0] d e ^ d b_p ^ e a_p ^ c_p.  One message per
// AND depth: the adder's generate layer and its prefix steps, then the
// DltSft AND tree (w >= 2) and the Delta ANDs, each split per DPF block as
// the reference evaluates one circuit per block.
#include <algorithm>

#include "preproc.hpp"

namespace pcgsq {

namespace {

constexpr int kTpb = 256;

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Instance g = blk*B + idx, blk = ii*c + jj, idx = kk*t + ll.
struct Inst {
  int ii, jj, kk, ll;
};
__device__ __forceinline__ Inst inst_of(size_t g, int c, int t_log) {
  const int B_log = 2 * t_log;
  const int blk = (int)(g >> B_log), idx = (int)(g & ((1ULL << B_log) - 1));
  return Inst{blk / c, blk % c, idx >> t_log, idx & ((1 << t_log) - 1)};
}

__device__ __forceinline__ uint64_t noise_pos(const uint64_t* noise, int ct, int p, int poly, int k,
                                              int t) {
  return noise[(size_t)p * ct + (size_t)poly * t + k];
}
__device__ __forceinline__ uint64_t noise_pay(const uint64_t* noise, int ct, int p, int poly, int k,
                                              int t) {
  return noise[(size_t)2 * ct + (size_t)p * ct + (size_t)poly * t + k];
}

__device__ __forceinline__ Blk own_delta(uint64_t seed, int p, uint64_t g) {
  Blk d{dealer_hd(seed, 3, (uint64_t)p, 2 * g), dealer_hd(seed, 3, (uint64_t)p, 2 * g + 1)};
  if (p == 0) d.lo |= 1ULL;   // dpf.cpp:78-82
  else d.lo &= ~1ULL;
  return d;
}

__device__ __forceinline__ uint64_t own_r(uint64_t seed, int p, uint64_t g) {
  return dealer_hd(seed, 6, g, (uint64_t)p) % kP;
}
__device__ __forceinline__ uint64_t own_s(uint64_t seed, int p, uint64_t g) {
  return dealer_hd(seed, 7, g, (uint64_t)p) % kP;
}

// n < 64 bits of a bit array starting at bit `pos`.
__device__ __forceinline__ uint64_t bits_at(const uint64_t* w, size_t pos, int n) {
  const size_t i = pos >> 6;
  const int sh = (int)(pos & 63);
  uint64_t v = w[i] >> sh;
  if (sh != 0 && sh + n > 64) v |= w[i + 1] << (64 - sh);
  return n >= 64 ? v : v & ((1ULL << n) - 1);
}

__device__ __forceinline__ uint64_t and_share(int p, uint64_t D, uint64_t E, uint64_t a, uint64_t b,
                                              uint64_t c) {
  return (p == 0 ? (D & E) : 0ULL) ^ (D & b) ^ (E & a) ^ c;
}

// ---- own draws ------------------------------------------------------------------

__global__ void pre_noise_kernel(uint64_t* __restrict__ noise, uint64_t seed, int p, int c,
                                 int t_log, uint64_t bin) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;   // i*t + k
  const int ct = c << t_log;
  if (r >= ct) return;
  const int k = r & ((1 << t_log) - 1);
  noise[(size_t)p * ct + r] =
      (uint64_t)k * bin + (dealer_hd(seed, 1, (uint64_t)p, (uint64_t)r) & (bin - 1));
  noise[(size_t)2 * ct + (size_t)p * ct + r] =
      dealer_hd(seed, 2, (uint64_t)p, (uint64_t)r) % (kP - 1) + 1;
}

__global__ void pre_own_kernel(Blk* __restrict__ roots, uint64_t* __restrict__ bt, uint64_t seed,
                               int p, int ninst) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ninst) return;
  const size_t ni = (size_t)ninst, o = (size_t)p * ni + (size_t)g;
  const Blk d = own_delta(seed, p, (uint64_t)g);
  roots[o] = Blk{d.lo ^ kWlo, d.hi ^ kWhi};
  bt[2 * ni + o] = own_r(seed, p, (uint64_t)g);   // R
  bt[4 * ni + o] = own_s(seed, p, (uint64_t)g);   // S
}

// Gilboa operands: party 0's x_q = (pay0, r0, s0), party 1's y_q = (pay1, s1, r1).
__device__ __forceinline__ uint64_t gilboa_operand(const uint64_t* noise, uint64_t seed, int p,
                                                   size_t g, int q, int c, int t_log) {
  const Inst in = inst_of(g, c, t_log);
  const int t = 1 << t_log, ct = c << t_log;
  if (q == 0)
    return p == 0 ? noise_pay(noise, ct, 0, in.ii, in.kk, t) : noise_pay(noise, ct, 1, in.jj, in.ll, t);
  if (q == 1) return p == 0 ? own_r(seed, 0, g) : own_s(seed, 1, g);
  return p == 0 ? own_s(seed, 0, g) : own_r(seed, 1, g);
}

// Choice words of the receiver of direction dir (party 1 - dir): random
// below sec_gil; in direction 0 the Gilboa section holds party 1's y_q.
__global__ void pre_choice_kernel(uint64_t* __restrict__ choice, const uint64_t* __restrict__ noise,
                                  uint64_t seed, int dir, size_t nwords, size_t gil_word,
                                  int ninst, int c, int t_log) {
  const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  if (dir != 0 || w < gil_word) {
    choice[w] = dealer_hd(seed, 16, (uint64_t)dir, (uint64_t)w);
    return;
  }
  const size_t x = w - gil_word, g = x / 3;
  const int q = (int)(x % 3);
  choice[w] = g < (size_t)ninst ? gilboa_operand(noise, seed, 1, g, q, c, t_log) : 0;
}

// ---- triples ----------------------------------------------------------------------

struct TriArgs {
  const uint64_t* choice;   // receiving direction (1 - p)
  const Blk* mc;
  const Blk* m0;            // sending direction (p)
  const Blk* m1;
  uint64_t* tri_add;        // [g][op][a, b, c]
  uint32_t* tri_tree;       // [g][e]: a bits 0..3, b bits 4..7, c bits 8..11
  uint64_t* tri_dlt;        // [g][e][a, b.lo, b.hi, c.lo, c.hi]
  size_t sec_add, sec_tree, sec_delta;
  int ninst, ops, nb, le, w;   // le = levels * nk
};

__global__ void pre_tri_add_kernel(TriArgs A) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // g*ops + op
  if (x >= (size_t)A.ninst * A.ops) return;
  const size_t pos = A.sec_add + x * (size_t)A.nb;
  const uint64_t a = bits_at(A.choice, pos, A.nb);
  uint64_t rb = 0, s0 = 0, s1 = 0;
  for (int i = 0; i < A.nb; ++i) {
    rb |= (A.mc[pos + i].lo & 1ULL) << i;
    s0 |= (A.m0[pos + i].lo & 1ULL) << i;
    s1 |= (A.m1[pos + i].lo & 1ULL) << i;
  }
  const uint64_t b = s0 ^ s1;
  A.tri_add[3 * x] = a;
  A.tri_add[3 * x + 1] = b;
  A.tri_add[3 * x + 2] = (a & b) ^ s0 ^ rb;
}

__global__ void pre_tri_dlt_kernel(TriArgs A) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // g*le + e
  if (x >= (size_t)A.ninst * A.le) return;
  if (A.w >= 2) {
    uint32_t packed = 0;
    for (int j = 0; j < A.w - 1; ++j) {
      const size_t pos = A.sec_tree + x * (size_t)(A.w - 1) + (size_t)j;
      const uint32_t a = (uint32_t)(A.choice[pos >> 6] >> (pos & 63)) & 1u;
      const uint32_t rb = (uint32_t)A.mc[pos].lo & 1u;
      const uint32_t s0 = (uint32_t)A.m0[pos].lo & 1u, s1 = (uint32_t)A.m1[pos].lo & 1u;
      const uint32_t b = s0 ^ s1;
      packed |= (a << j) | (b << (4 + j)) | ((((a & b) ^ s0 ^ rb) & 1u) << (8 + j));
    }
    A.tri_tree[x] = packed;
  }
  const size_t pos = A.sec_delta + x;
  const bool a = (A.choice[pos >> 6] >> (pos & 63)) & 1ULL;
  const Blk s0 = A.m0[pos], s1 = A.m1[pos], rv = A.mc[pos];
  const Blk b{s0.lo ^ s1.lo, s0.hi ^ s1.hi};
  uint64_t* o = A.tri_dlt + 5 * x;
  o[0] = a ? ~0ULL : 0ULL;
  o[1] = b.lo;
  o[2] = b.hi;
  o[3] = (a ? b.lo : 0ULL) ^ s0.lo ^ rv.lo;
  o[4] = (a ? b.hi : 0ULL) ^ s0.hi ^ rv.hi;
}

// ---- Gilboa -------------------------------------------------------------------------

__global__ void pre_gil_send_kernel(const Blk* __restrict__ m0, const Blk* __restrict__ m1,
                                    const uint64_t* __restrict__ noise, uint64_t seed,
                                    size_t sec_gil, int ninst, int c, int t_log,
                                    uint64_t* __restrict__ delta, uint64_t* __restrict__ share) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // g*3 + q
  if (x >= (size_t)ninst * 3) return;
  const uint64_t xv = gilboa_operand(noise, seed, 0, x / 3, (int)(x % 3), c, t_log);
  const size_t pos = sec_gil + x * 64;
  uint64_t rsum = 0, xs = xv;
  for (int j = 0; j < 62; ++j) {
    const uint64_t r = freduce64(m0[pos + j].lo), v1 = freduce64(m1[pos + j].lo);
    delta[x * 64 + j] = fsub(fadd(r, xs), v1);
    rsum = fadd(rsum, r);
    xs = fadd(xs, xs);
  }
  delta[x * 64 + 62] = 0;
  delta[x * 64 + 63] = 0;
  share[x] = fneg(rsum);
}

__global__ void pre_gil_recv_kernel(const Blk* __restrict__ mc, const uint64_t* __restrict__ delta,
                                    const uint64_t* __restrict__ noise, uint64_t seed,
                                    size_t sec_gil, int ninst, int c, int t_log,
                                    uint64_t* __restrict__ share) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (size_t)ninst * 3) return;
  const uint64_t y = gilboa_operand(noise, seed, 1, x / 3, (int)(x % 3), c, t_log);
  const size_t pos = sec_gil + x * 64;
  uint64_t v = 0;
  for (int j = 0; j < 62; ++j) {
    uint64_t m = freduce64(mc[pos + j].lo);
    if ((y >> j) & 1ULL) m = fadd(m, delta[x * 64 + j]);
    v = fadd(v, m);
  }
  share[x] = v;
}

// beta_p and rs_p = r_p s_p + (r0 s1)_p + (s0 r1)_p.
__global__ void pre_bt_kernel(const uint64_t* __restrict__ share, uint64_t* __restrict__ bt,
                              uint64_t seed, int p, int ninst) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ninst) return;
  const size_t ni = (size_t)ninst, o = (size_t)p * ni + (size_t)g;
  bt[o] = share[3 * (size_t)g];
  const uint64_t rs = fmul(own_r(seed, p, (uint64_t)g), own_s(seed, p, (uint64_t)g));
  bt[6 * ni + o] = fadd(fadd(rs, share[3 * (size_t)g + 1]), share[3 * (size_t)g + 2]);
}

// ---- GMW: the position adder ----------------------------------------------------------
// state[g] = {P, G, P0, alpha}.  Step k = 0 opens the generate layer (op 0:
// A & B); step 1 <= k <= S finishes step k - 1 and opens prefix step k
// (distance d = 2^(k-1): ops 2k-1 = P & G<<d and 2k = P & P<<d, kept on bits
// >= d); step S + 1 finishes step S and forms alpha = P0 ^ (G << 1).
struct AdderArgs {
  const uint64_t* noise;
  const uint64_t* tri;       // [g][op][a, b, c]
  uint64_t* state;
  const uint64_t* mine_prev;
  const uint64_t* peer;
  uint64_t* mine_new;
  int p, k, S, ops, nb, ninst, c, t_log;
  uint64_t bin;
};

__global__ void pre_adder_kernel(AdderArgs A) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.ninst) return;
  const uint64_t mask = (1ULL << A.nb) - 1;
  uint64_t* st = A.state + 4 * (size_t)g;
  const uint64_t* tri = A.tri + (size_t)g * A.ops * 3;
  auto fin = [&](int op, int slot, int wpi) {
    const uint64_t* me = A.mine_prev + (size_t)g * wpi + 2 * slot;
    const uint64_t* pe = A.peer + (size_t)g * wpi + 2 * slot;
    return and_share(A.p, me[0] ^ pe[0], me[1] ^ pe[1], tri[3 * op], tri[3 * op + 1],
                     tri[3 * op + 2]) & mask;
  };
  uint64_t P, G, P0;
  if (A.k == 0) {
    const Inst in = inst_of((size_t)g, A.c, A.t_log);
    const int t = 1 << A.t_log, ct = A.c << A.t_log;
    const uint64_t off = (A.p == 0 ? noise_pos(A.noise, ct, 0, in.ii, in.kk, t)
                                   : noise_pos(A.noise, ct, 1, in.jj, in.ll, t)) & (A.bin - 1);
    const uint64_t X = A.p == 0 ? off : 0, Y = A.p == 1 ? off : 0;
    P = X ^ Y;
    P0 = P;
    G = 0;
    uint64_t* o = A.mine_new + 2 * (size_t)g;
    o[0] = X ^ tri[0];
    o[1] = Y ^ tri[1];
  } else {
    P = st[0];
    G = st[1];
    P0 = st[2];
    if (A.k == 1) {
      G = fin(0, 0, 2);
    } else {
      const int s = A.k - 1, d = 1 << (s - 1);
      const uint64_t hi = mask & ~((1ULL << d) - 1);
      const uint64_t X = fin(2 * s - 1, 0, 4), Y = fin(2 * s, 1, 4);
      G ^= X & hi;
      P = (P & ~hi) | (Y & hi);
    }
    if (A.k <= A.S) {
      const int d = 1 << (A.k - 1), o1 = 2 * A.k - 1, o2 = 2 * A.k;
      uint64_t* o = A.mine_new + 4 * (size_t)g;
      o[0] = P ^ tri[3 * o1];
      o[1] = ((G << d) & mask) ^ tri[3 * o1 + 1];
      o[2] = P ^ tri[3 * o2];
      o[3] = ((P << d) & mask) ^ tri[3 * o2 + 1];
    } else {
      st[3] = P0 ^ (G << 1);   // alpha, nb + 1 bits
    }
  }
  st[0] = P;
  st[1] = G;
  st[2] = P0;
}

// ---- GMW: the F_DeltaShiftShare circuit ------------------------------------------------
// One thread per (g, e), e = i*nk + k.  fac[g][e] holds the current list of
// indicator factors (bits 0..3) and its length (bits 8..11).  Layer kinds:
// 1, 2 = AND-tree depth 1, 2 (pairs of the list, the odd one carried, as
// dpf.cpp:219-226 builds them); 3 = the Delta ANDs.
enum { kDltNone = 0, kDltTree1 = 1, kDltTree2 = 2, kDltDelta = 3 };

struct DltArgs {
  const uint64_t* state;    // alpha at state[4g + 3]
  const uint32_t* tri_tree;
  const uint64_t* tri_dlt;
  uint32_t* fac;
  const uint64_t* mine_prev;
  const uint64_t* peer;
  uint64_t* mine_new;
  Blk* dlt;                 // this party's DltSft [ninst][levels][nk]
  uint64_t seed;
  int p, fin, open, ninst, le, nk, w, n;
};

__global__ void pre_dlt_kernel(DltArgs A) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // g*le + e
  if (x >= (size_t)A.ninst * A.le) return;
  const size_t g = x / A.le;
  const int e = (int)(x % A.le), i = e / A.nk, k = e % A.nk;
  uint32_t f, nf;
  if (A.fin == kDltNone) {
    const uint64_t alpha = A.state[4 * g + 3];
    f = 0;
    for (int cc = 0; cc < A.w; ++cc) {
      const int pos = A.n - A.w * (i + 1) + cc;
      uint32_t b = (uint32_t)(alpha >> pos) & 1u;
      if (A.p == 0 && !((k >> cc) & 1)) b ^= 1u;   // NOT: party 0 flips its share
      f |= b << cc;
    }
    nf = (uint32_t)A.w;
  } else {
    f = A.fac[x] & 0xFu;
    nf = (A.fac[x] >> 8) & 0xFu;
  }
  const uint32_t tt = A.w >= 2 ? A.tri_tree[x] : 0u;
  if (A.fin == kDltTree1 || A.fin == kDltTree2) {
    const int base = A.fin == kDltTree1 ? 0 : A.w / 2;
    const uint64_t me = A.mine_prev[x], pe = A.peer[x];
    const int np = (int)nf / 2;
    uint32_t nfac = 0;
    for (int j = 0; j < np; ++j) {
      const int gi = base + j;
      const uint64_t D = ((me ^ pe) >> (2 * j)) & 1ULL, E = ((me ^ pe) >> (2 * j + 1)) & 1ULL;
      const uint64_t z = and_share(A.p, D, E, (tt >> gi) & 1u, (tt >> (4 + gi)) & 1u,
                                   (tt >> (8 + gi)) & 1u) & 1ULL;
      nfac |= (uint32_t)z << j;
    }
    if (nf & 1u) nfac |= ((f >> (nf - 1)) & 1u) << np;
    f = nfac;
    nf = (uint32_t)np + (nf & 1u);
  } else if (A.fin == kDltDelta) {
    const uint64_t* t5 = A.tri_dlt + 5 * x;
    const uint64_t* me = A.mine_prev + 3 * x;
    const uint64_t* pe = A.peer + 3 * x;
    const uint64_t D = ((me[0] ^ pe[0]) & 1ULL) ? ~0ULL : 0ULL;
    const uint64_t Elo = me[1] ^ pe[1], Ehi = me[2] ^ pe[2];
    A.dlt[g * A.le + e] = Blk{and_share(A.p, D, Elo, t5[0], t5[1], t5[3]),
                              and_share(A.p, D, Ehi, t5[0], t5[2], t5[4])};
    return;
  }
  if (A.open == kDltTree1 || A.open == kDltTree2) {
    const int base = A.open == kDltTree1 ? 0 : A.w / 2;
    uint64_t word = 0;
    for (int j = 0; j < (int)nf / 2; ++j) {
      const int gi = base + j;
      word |= (uint64_t)(((f >> (2 * j)) ^ (tt >> gi)) & 1u) << (2 * j);
      word |= (uint64_t)(((f >> (2 * j + 1)) ^ (tt >> (4 + gi))) & 1u) << (2 * j + 1);
    }
    A.mine_new[x] = word;
  } else if (A.open == kDltDelta) {
    const uint64_t* t5 = A.tri_dlt + 5 * x;
    const Blk d = own_delta(A.seed, A.p, (uint64_t)g);
    uint64_t* o = A.mine_new + 3 * x;
    o[0] = (uint64_t)(f & 1u) ^ (t5[0] & 1ULL);
    o[1] = d.lo ^ t5[1];
    o[2] = d.hi ^ t5[2];
  }
  A.fac[x] = f | (nf << 8);
}

unsigned grid_of(size_t n) { return (unsigned)((n + kTpb - 1) / kTpb); }

// The DltSft layers in order (after the adder).
int dlt_layers(int w, int* kinds) {
  int n = 0;
  if (w >= 2) kinds[n++] = kDltTree1;
  if (w >= 3) kinds[n++] = kDltTree2;
  kinds[n++] = kDltDelta;
  return n;
}

}  // namespace

PreShape make_pre_shape(int N, int c, int t, int w) {
  PreShape s;
  s.c = c;
  s.t = t;
  s.t_log = ilog2_exact((uint64_t)t);
  require(s.t_log >= 0, "preprocessing: t must be a power of two");
  s.B = t * t;
  s.nblk = c * c;
  s.ninst = s.nblk * s.B;
  s.bin = (uint64_t)N / (uint64_t)t;
  s.D_log = ilog2_exact(2 * s.bin);
  s.w = w;
  s.levels = s.D_log / w;
  s.nk = (1 << w) - 1;
  s.nb = s.D_log - 1;
  require(s.nb >= 1 && s.nb <= 62, "preprocessing: N/t must be in [2, 2^62]", kErrUnsupported);
  s.S = 0;
  for (int d = 1; d < s.nb; d <<= 1) ++s.S;
  s.ops = 1 + 2 * s.S;
  const size_t ni = (size_t)s.ninst, le = (size_t)s.levels * s.nk;
  s.sec_add = 0;
  s.sec_tree = round_up(ni * s.ops * s.nb, kOtChunk);
  s.sec_delta = s.sec_tree + round_up(ni * le * (size_t)(w - 1), kOtChunk);
  s.sec_gil = s.sec_delta + round_up(ni * le, kOtChunk);
  s.m[1] = s.sec_gil;
  s.m[0] = s.sec_gil + round_up(ni * 3 * 64, kOtChunk);
  return s;
}

void pre_prepare(PreRun& r, const PreShape& sh, int party, const BaseOts& base) {
  require(party == 0 || party == 1, "preprocessing: party must be 0 or 1");
  r.party = party;
  r.sh = sh;
  const size_t ni = (size_t)sh.ninst, le = (size_t)sh.levels * sh.nk;
  const size_t mr = sh.m[1 - party], ms = sh.m[party];
  r.base.ensure(3 * kKappa * sizeof(Blk), "preprocessing base OTs");
  hip_check(hipMemcpy(r.base.as<Blk>(), base.k0, kKappa * sizeof(Blk), hipMemcpyHostToDevice),
            "upload base OTs");
  hip_check(hipMemcpy(r.base.as<Blk>() + kKappa, base.k1, kKappa * sizeof(Blk),
                      hipMemcpyHostToDevice), "upload base OTs");
  hip_check(hipMemcpy(r.base.as<Blk>() + 2 * kKappa, base.ks, kKappa * sizeof(Blk),
                      hipMemcpyHostToDevice), "upload base OTs");
  r.base_s = base.s;
  r.choice.ensure(mr / 64 * 8, "preprocessing choice bits");
  r.u_mine.ensure(mr * 16, "preprocessing IKNP matrix");
  r.u_peer.ensure(ms * 16, "preprocessing IKNP matrix");
  r.mc.ensure(mr * sizeof(Blk), "preprocessing OT outputs");
  r.m0.ensure(ms * sizeof(Blk), "preprocessing OT outputs");
  r.m1.ensure(ms * sizeof(Blk), "preprocessing OT outputs");
  r.tri_add.ensure(ni * sh.ops * 3 * 8, "preprocessing triples");
  r.tri_tree.ensure(std::max<size_t>(1, ni * le) * 4, "preprocessing triples");
  r.tri_dlt.ensure(ni * le * 5 * 8, "preprocessing triples");
  r.state.ensure(ni * 4 * 8, "preprocessing state");
  r.fac.ensure(ni * le * 4, "preprocessing state");
  const size_t wmax = ni * std::max<size_t>(4, 3 * le) * 8;
  r.open_mine[0].ensure(wmax, "preprocessing opens");
  r.open_mine[1].ensure(wmax, "preprocessing opens");
  r.open_peer.ensure(wmax, "preprocessing opens");
  r.gil.ensure(ni * 3 * 64 * 8, "preprocessing Gilboa");
  r.gshare.ensure(ni * 3 * 8, "preprocessing Gilboa");
  r.stage = -1;
}

void pre_begin(PreRun& r, uint64_t seed, uint64_t sid, const PreOut& out, hipStream_t st) {
  r.seed = seed;
  r.sid = sid;
  r.out = out;
  r.st = st;
  r.stage = 0;
}

namespace {

// Layer bookkeeping kept in PreRun.stage: 0 ot_ext, 1 gilboa, 2 adder step
// 0, 3 + k - 1 adder step k (1 <= k <= S), then S + 3 + j the j-th DltSft
// layer message, and past the last one the finishing launch.
void launch_adder(PreRun& r, int k) {
  const PreShape& sh = r.sh;
  AdderArgs A;
  A.noise = r.out.noise;
  A.tri = r.tri_add.as<uint64_t>();
  A.state = r.state.as<uint64_t>();
  A.mine_prev = r.open_mine[(k + 1) & 1].as<uint64_t>();
  A.peer = r.open_peer.as<uint64_t>();
  A.mine_new = r.open_mine[k & 1].as<uint64_t>();
  A.p = r.party;
  A.k = k;
  A.S = sh.S;
  A.ops = sh.ops;
  A.nb = sh.nb;
  A.ninst = sh.ninst;
  A.c = sh.c;
  A.t_log = sh.t_log;
  A.bin = sh.bin;
  hipLaunchKernelGGL(pre_adder_kernel, dim3(grid_of((size_t)sh.ninst)), dim3(kTpb), 0, r.st, A);
  hip_check(hipGetLastError(), "launch pre_adder_kernel");
}

// layer index L counts the opens made so far (adder: S + 1 of them).
void launch_dlt(PreRun& r, int fin, int open, int L) {
  const PreShape& sh = r.sh;
  DltArgs A;
  A.state = r.state.as<uint64_t>();
  A.tri_tree = r.tri_tree.as<uint32_t>();
  A.tri_dlt = r.tri_dlt.as<uint64_t>();
  A.fac = r.fac.as<uint32_t>();
  A.mine_prev = r.open_mine[(L + 1) & 1].as<uint64_t>();
  A.peer = r.open_peer.as<uint64_t>();
  A.mine_new = r.open_mine[L & 1].as<uint64_t>();
  A.dlt = r.out.dlt + (size_t)r.party * sh.ninst * sh.levels * sh.nk;
  A.seed = r.seed;
  A.p = r.party;
  A.fin = fin;
  A.open = open;
  A.ninst = sh.ninst;
  A.le = sh.levels * sh.nk;
  A.nk = sh.nk;
  A.w = sh.w;
  A.n = sh.D_log;
  hipLaunchKernelGGL(pre_dlt_kernel, dim3(grid_of((size_t)sh.ninst * A.le)), dim3(kTpb), 0, r.st, A);
  hip_check(hipGetLastError(), "launch pre_dlt_kernel");
}

void circuit_msg(PreRun& r, PreMsg* msg, int level, int gates_per_inst, size_t words_per_inst,
                 int L) {
  const PreShape& sh = r.sh;
  msg->phase = "circuit_and";
  msg->level = level;
  msg->batch = sh.B * gates_per_inst;
  msg->d_send = r.open_mine[L & 1].ptr;
  msg->send_bytes = (size_t)sh.ninst * words_per_inst * 8;
  msg->d_recv = r.open_peer.ptr;
  msg->recv_bytes = msg->send_bytes;
  msg->nblocks = sh.nblk;
}

}  // namespace

bool pre_step(PreRun& r, PreMsg* msg) {
  const PreShape& sh = r.sh;
  const int p = r.party;
  const size_t ni = (size_t)sh.ninst, le = (size_t)sh.levels * sh.nk;
  const size_t mr = sh.m[1 - p], ms = sh.m[p];
  hipStream_t st = r.st;
  *msg = PreMsg{};
  const int S = sh.S;
  int kinds[3];
  const int nd = dlt_layers(sh.w, kinds);
  if (r.stage == 0) {
    hipLaunchKernelGGL(pre_noise_kernel, dim3(grid_of((size_t)sh.c * sh.t)), dim3(kTpb), 0, st,
                       r.out.noise, r.seed, p, sh.c, sh.t_log, sh.bin);
    hipLaunchKernelGGL(pre_own_kernel, dim3(grid_of(ni)), dim3(kTpb), 0, st, r.out.roots,
                       r.out.bt, r.seed, p, sh.ninst);
    hipLaunchKernelGGL(pre_choice_kernel, dim3(grid_of(mr / 64)), dim3(kTpb), 0, st,
                       r.choice.as<uint64_t>(), r.out.noise, r.seed, 1 - p, mr / 64,
                       sh.sec_gil / 64, sh.ninst, sh.c, sh.t_log);
    hip_check(hipGetLastError(), "launch preprocessing draws");
    const Blk* base = r.base.as<Blk>();
    iknp_receive(st, base, base + kKappa, r.sid, 1 - p, r.choice.as<uint64_t>(), mr,
                 r.u_mine.as<uint64_t>(), r.mc.as<Blk>());
    msg->phase = "ot_ext";
    msg->level = 0;
    msg->batch = (int)std::min<size_t>(mr, (size_t)INT32_MAX);
    msg->d_send = r.u_mine.ptr;
    msg->send_bytes = mr * 16;
    msg->d_recv = r.u_peer.ptr;
    msg->recv_bytes = ms * 16;
    r.stage = 1;
    return true;
  }
  if (r.stage == 1) {
    iknp_send(st, r.base.as<Blk>() + 2 * kKappa, r.base_s, r.sid, p, r.u_peer.as<uint64_t>(), ms,
              r.m0.as<Blk>(), r.m1.as<Blk>());
    TriArgs T;
    T.choice = r.choice.as<uint64_t>();
    T.mc = r.mc.as<Blk>();
    T.m0 = r.m0.as<Blk>();
    T.m1 = r.m1.as<Blk>();
    T.tri_add = r.tri_add.as<uint64_t>();
    T.tri_tree = r.tri_tree.as<uint32_t>();
    T.tri_dlt = r.tri_dlt.as<uint64_t>();
    T.sec_add = sh.sec_add;
    T.sec_tree = sh.sec_tree;
    T.sec_delta = sh.sec_delta;
    T.ninst = sh.ninst;
    T.ops = sh.ops;
    T.nb = sh.nb;
    T.le = (int)le;
    T.w = sh.w;
    hipLaunchKernelGGL(pre_tri_add_kernel, dim3(grid_of(ni * sh.ops)), dim3(kTpb), 0, st, T);
    hipLaunchKernelGGL(pre_tri_dlt_kernel, dim3(grid_of(ni * le)), dim3(kTpb), 0, st, T);
    if (p == 0)
      hipLaunchKernelGGL(pre_gil_send_kernel, dim3(grid_of(ni * 3)), dim3(kTpb), 0, st,
                         r.m0.as<Blk>(), r.m1.as<Blk>(), r.out.noise, r.seed, sh.sec_gil,
                         sh.ninst, sh.c, sh.t_log, r.gil.as<uint64_t>(), r.gshare.as<uint64_t>());
    hip_check(hipGetLastError(), "launch preprocessing triples");
    msg->phase = "gilboa";
    msg->level = 0;
    msg->batch = (int)(ni * 3);
    if (p == 0) {
      msg->d_send = r.gil.ptr;
      msg->send_bytes = ni * 3 * 64 * 8;
    } else {
      msg->d_recv = r.gil.ptr;
      msg->recv_bytes = ni * 3 * 64 * 8;
    }
    r.stage = 2;
    return true;
  }
  if (r.stage == 2) {
    if (p == 1)
      hipLaunchKernelGGL(pre_gil_recv_kernel, dim3(grid_of(ni * 3)), dim3(kTpb), 0, st,
                         r.mc.as<Blk>(), r.gil.as<uint64_t>(), r.out.noise, r.seed, sh.sec_gil,
                         sh.ninst, sh.c, sh.t_log, r.gshare.as<uint64_t>());
    hipLaunchKernelGGL(pre_bt_kernel, dim3(grid_of(ni)), dim3(kTpb), 0, st,
                       r.gshare.as<uint64_t>(), r.out.bt, r.seed, p, sh.ninst);
    hip_check(hipGetLastError(), "launch preprocessing Gilboa");
    launch_adder(r, 0);
    circuit_msg(r, msg, 1, sh.nb, 2, 0);
    r.stage = 3;
    return true;
  }
  const int k = r.stage - 2;   // adder step, 1 <= k <= S + 1; then DltSft layers
  if (k <= S) {
    launch_adder(r, k);
    circuit_msg(r, msg, k + 1, 2 * (sh.nb - (1 << (k - 1))), 4, k);
    ++r.stage;
    return true;
  }
  const int j = k - (S + 1);   // DltSft layer j opens now (j < nd), or finish (j == nd)
  if (j > nd) return false;
  if (j == 0) launch_adder(r, S + 1);   // alpha
  const int L = S + 1 + j;              // opens made so far before this launch
  launch_dlt(r, j == 0 ? kDltNone : kinds[j - 1], j < nd ? kinds[j] : kDltNone, L);
  ++r.stage;
  if (j == nd) return false;
  const int kind = kinds[j];
  const int gates = kind == kDltTree1 ? (int)le * (sh.w / 2)
                                      : kind == kDltTree2 ? (int)le : (int)le * 128;
  circuit_msg(r, msg, j + 1, gates, kind == kDltDelta ? 3 * le : le, L);
  return true;
}

}  // namespace pcgsq
