// Device-side building blocks shared by the PCG-OLE kernels (gfx950):
//   * F_p arithmetic for P = 2^62 - 3*2^25 + 1 on the 32-bit VALU
//     (lazy-reduced add/sub, Shoup multiplication by precomputed constants,
//     and a general product reduced through the special form of P),
//   * the ChaCha8 block function held entirely in VGPRs,
//   * 64-lane wave reductions.
// All arithmetic is integer and bit-exact; nothing here uses floating point.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PCGSQ_HD __host__ __device__ __forceinline__

namespace pcgsq {

constexpr uint64_t kP = 4611686018326724609ULL;      // 2^62 - 3*2^25 + 1
constexpr uint64_t kTwoP = 2 * kP;
constexpr uint64_t kC62 = (3ULL << 25) - 1;          // 2^62 mod P
constexpr uint64_t kOutDomain = 0x6f75745f68617368ULL; // "out_hash" nonce

struct alignas(16) Blk {
  uint64_t lo, hi;
};

PCGSQ_HD Blk bxor(Blk a, Blk b) { return Blk{a.lo ^ b.lo, a.hi ^ b.hi}; }

// ---- field ----------------------------------------------------------------

PCGSQ_HD uint64_t fadd(uint64_t a, uint64_t b) {  // a,b < P
  uint64_t s = a + b;
  return s >= kP ? s - kP : s;
}
PCGSQ_HD uint64_t fsub(uint64_t a, uint64_t b) {  // a,b < P
  return a >= b ? a - b : a + kP - b;
}
PCGSQ_HD uint64_t fneg(uint64_t a) { return a ? kP - a : 0; }

// x mod P for any 64-bit x: x - (x>>62)*P lands in [0, 2P), one subtract.
PCGSQ_HD uint64_t freduce64(uint64_t x) {
  uint64_t q = x >> 62;
  uint64_t r = (x & ((1ULL << 62) - 1)) + q * kC62;
  return r >= kP ? r - kP : r;
}

// High 64 bits of a 64x64 product as a chain of 32x32+64 multiply-adds
// (v_mad_u64_u32): t and u never overflow 64 bits.
PCGSQ_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
  const uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32);
  const uint32_t bl = (uint32_t)b, bh = (uint32_t)(b >> 32);
  const uint64_t t = (uint64_t)ah * bl + (((uint64_t)al * bl) >> 32);
  const uint64_t u = (uint64_t)al * bh + (uint32_t)t;
  return (uint64_t)ah * bh + ((t >> 32) + (u >> 32));
}

// q*P mod 2^64 using the shape of P (shifts and adds only).
PCGSQ_HD uint64_t mulP(uint64_t q) {
  return (q << 62) - (q << 26) - (q << 25) + q;
}

// ---- the seeded dealer's draws ----------------------------------------------
// U(seed, domain, a, b) = sm(sm(sm(seed ^ domain*K) ^ a) ^ b), sm = the
// splitmix64 finalizer, K = 0xD6E8FEB86659FD93: every seeded draw of the
// dealer and of a party's own randomness (the oracle restates it).
PCGSQ_HD uint64_t sm64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

PCGSQ_HD uint64_t dealer_hd(uint64_t seed, uint32_t domain, uint64_t a, uint64_t b) {
  return sm64(sm64(sm64(seed ^ ((uint64_t)domain * 0xD6E8FEB86659FD93ULL)) ^ a) ^ b);
}

// The public block W of the DPF roots, root_s = Delta_s ^ W (the reference
// draws it from emp::PRG(&F_RAND_SEED), half_tree_dpf/dpf.cpp:25-26,291; it
// cancels between the parties, so any public constant gives the same CWs).
constexpr uint64_t kWlo = 0x0123456789abcdefULL, kWhi = 0xdeadbeefcafe0001ULL;

// Shoup: x*w mod P given wq = floor(w * 2^64 / P), w < P.  Any x < 2^64;
// result in [0, 2P).  x*w - q*P is formed as (x*w + 3q*2^25) - (q + q*2^62).
PCGSQ_HD uint64_t shoup_lazy(uint64_t x, uint64_t w, uint64_t wq) {
  const uint64_t q = mulhi64(x, wq);
  return (x * w + ((q * 3) << 25)) - (q + (q << 62));
}
// The same two functions in the form the NTT butterflies use on the device.
// mulhi64_dev splits the middle sum as u = al*bh + t with t the full 64-bit
// ah*bl + hi(al*bl): u can exceed 2^64, and its carry comes from
// v_mad_u64_u32's own carry-out (so the high product is ah*bh + u/2^32 +
// carry*2^32, with no zero-extended 32-bit addends to pair up in registers).
// shoup_lazy_dev forms x*w - q*P as x*w + q*(2^64 - P) mod 2^64 in 32-bit
// pieces.  Both equal mulhi64 / shoup_lazy for every input (16 instead of 19
// VALU instructions per multiply).
__device__ __forceinline__ uint64_t mulhi64_dev(uint64_t a, uint64_t b) {
  const uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32);
  const uint32_t bl = (uint32_t)b, bh = (uint32_t)(b >> 32);
  const uint64_t t = (uint64_t)ah * bl + (uint64_t)__umulhi(al, bl);
  uint64_t u;
  uint32_t c;
  asm("v_mad_u64_u32 %0, vcc, %2, %3, %4\n\t"
      "v_cndmask_b32_e64 %1, 0, 1, vcc"
      : "=&v"(u), "=v"(c)
      : "v"(al), "v"(bh), "v"(t)
      : "vcc");
  const uint64_t q = (uint64_t)ah * bh + (u >> 32);
  return ((uint64_t)((uint32_t)(q >> 32) + c) << 32) | (uint32_t)q;
}
__device__ __forceinline__ uint64_t shoup_lazy_dev(uint64_t x, uint64_t w, uint64_t wq) {
  constexpr uint64_t kNegP = 0ULL - kP;
  constexpr uint32_t kNl = (uint32_t)kNegP, kNh = (uint32_t)(kNegP >> 32);
  const uint64_t q = mulhi64_dev(x, wq);
  const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
  const uint32_t wl = (uint32_t)w, wh = (uint32_t)(w >> 32);
  const uint32_t ql = (uint32_t)q, qh = (uint32_t)(q >> 32);
  const uint64_t lo = (uint64_t)ql * kNl + (uint64_t)xl * wl;
  const uint32_t hi = (uint32_t)(lo >> 32) + xl * wh + xh * wl + ql * kNh + qh * kNl;
  return ((uint64_t)hi << 32) | (uint32_t)lo;
}

// floor(w 2^64 / P) or one less, for w < P: floor(2^128 / P) = 2^66 + 16 c
// with c = 3 2^25 - 1 gives 4w + floor(16 c w / 2^64), at most one below.  A
// Shoup product (shoup_lazy) with this constant in place of the exact one lies
// in [0, 3P) for any x < 2^64 (the quotient is at most two below
// floor(x w / P) instead of one).
PCGSQ_HD uint64_t shoup_w_approx(uint64_t w) { return 4 * w + mulhi64(w, 16 * kC62); }

PCGSQ_HD uint64_t shoup(uint64_t x, uint64_t w, uint64_t wq) {
  uint64_t r = shoup_lazy(x, w, wq);
  return r >= kP ? r - kP : r;
}

// General a*b mod P for a, b < 2^64 with a*b < 2^124 (both < 2^62 suffices):
// fold the 124-bit product twice through 2^62 = 3*2^25 - 1 (mod P).
PCGSQ_HD uint64_t fmul(uint64_t a, uint64_t b) {
  const uint64_t lo = a * b;
  const uint64_t hi = mulhi64(a, b);                 // < 2^60
  // x = hi*2^64 + lo = x1*2^62 + x0
  const uint64_t x0 = lo & ((1ULL << 62) - 1);
  const uint64_t x1 = (hi << 2) | (lo >> 62);        // < 2^62
  // y = x1*c + x0, c < 2^27: x1*c as (x1h*c)<<32 + x1l*c
  const uint64_t pl = (uint64_t)(uint32_t)x1 * kC62;             // < 2^59
  const uint64_t ph = (uint64_t)(uint32_t)(x1 >> 32) * kC62;     // < 2^57
  // y = ph*2^32 + pl + x0 ; split at 62
  const uint64_t s = pl + x0;                                    // < 2^63
  const uint64_t y0 = (s & ((1ULL << 62) - 1)) + ((ph & ((1ULL << 30) - 1)) << 32);
  const uint64_t y1 = (s >> 62) + (ph >> 30);                    // < 2^28
  // y0 < 2^62 + 2^62 ; fold y1 and the carry of y0
  uint64_t r = (y0 & ((1ULL << 62) - 1)) + ((y1 + (y0 >> 62)) * kC62);  // < 2^62 + 2^56
  r = r >= kP ? r - kP : r;
  return r >= kP ? r - kP : r;
}

// ---- ChaCha8 ----------------------------------------------------------------

__device__ __forceinline__ uint32_t rotl32(uint32_t v, int n) {
  return __builtin_amdgcn_alignbit(v, v, 32 - n);
}

#define PCGSQ_QR(a, b, c, d)            \
  a += b; d ^= a; d = rotl32(d, 16);    \
  c += d; b ^= c; b = rotl32(b, 12);    \
  a += b; d ^= a; d = rotl32(d, 8);     \
  c += d; b ^= c; b = rotl32(b, 7);

// One ChaCha8 block with key = (k0..k3) || 0^128, 64-bit counter ctr and
// 64-bit nonce (CPU_baseline/chacha8.h:23-150 word placement).  Writes the
// first NW output words (NW <= 16); unused words are eliminated by the
// compiler, so the tree PRG (first 4 words) and the out-hash (first 2 words)
// only pay for what they read.
template <int NW>
__device__ __forceinline__ void chacha8(uint32_t k0, uint32_t k1, uint32_t k2, uint32_t k3,
                                        uint64_t ctr, uint64_t nonce, uint32_t* out) {
  const uint32_t c0 = 0x61707865u, c1 = 0x3320646eu, c2 = 0x79622d32u, c3 = 0x6b206574u;
  const uint32_t n12 = (uint32_t)ctr, n13 = (uint32_t)(ctr >> 32);
  const uint32_t n14 = (uint32_t)nonce, n15 = (uint32_t)(nonce >> 32);
  uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
  uint32_t x4 = k0, x5 = k1, x6 = k2, x7 = k3;
  uint32_t x8 = 0, x9 = 0, x10 = 0, x11 = 0;
  uint32_t x12 = n12, x13 = n13, x14 = n14, x15 = n15;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    PCGSQ_QR(x0, x4, x8, x12)
    PCGSQ_QR(x1, x5, x9, x13)
    PCGSQ_QR(x2, x6, x10, x14)
    PCGSQ_QR(x3, x7, x11, x15)
    PCGSQ_QR(x0, x5, x10, x15)
    PCGSQ_QR(x1, x6, x11, x12)
    PCGSQ_QR(x2, x7, x8, x13)
    PCGSQ_QR(x3, x4, x9, x14)
  }
  const uint32_t o[16] = {x0 + c0, x1 + c1, x2 + c2, x3 + c3, x4 + k0, x5 + k1, x6 + k2,
                          x7 + k3, x8, x9, x10, x11, x12 + n12, x13 + n13, x14 + n14,
                          x15 + n15};
#pragma unroll
  for (int i = 0; i < NW; ++i) out[i] = o[i];
}

// Tree PRG child k of a parent (ChaCha8TreePRG, half_tree_dpf/tree_prg.h:80-111):
// 16-byte chunk k of the keystream keyed by parent || 0^128, nonce 0.
__device__ __forceinline__ void chacha8_words_of(Blk p, uint32_t& k0, uint32_t& k1,
                                                 uint32_t& k2, uint32_t& k3) {
  k0 = (uint32_t)p.lo; k1 = (uint32_t)(p.lo >> 32);
  k2 = (uint32_t)p.hi; k3 = (uint32_t)(p.hi >> 32);
}

// lo64 of ChaCha8OutHash(leaf) (half_tree_dpf/tree_prg.h:128-148).
__device__ __forceinline__ uint64_t out_hash_lo64(Blk leaf) {
  uint32_t k0, k1, k2, k3, o[2];
  chacha8_words_of(leaf, k0, k1, k2, k3);
  chacha8<2>(k0, k1, k2, k3, 0, kOutDomain, o);
  return (uint64_t)o[0] | ((uint64_t)o[1] << 32);
}

// ---- wave reductions (wave64) --------------------------------------------------

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

// Workgroups of 256 threads.  Reduce per-thread values over "segments": a
// segment is the set of threads of one workgroup that belong to one tree (S = 2^S_log parents per tree;
// trees are aligned to S threads, workgroups to 256).  Word i combines with
// op(i, x, y).  Returns true on exactly one thread per segment, which then
// holds the reduced words in v and the segment's first thread in *first.
template <int NW, class Op>
__device__ __forceinline__ bool segment_reduce(uint64_t* v, int S_log, Op op, uint64_t* lds,
                                               int* first) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int width = S_log >= 6 ? 64 : (1 << S_log);
#pragma unroll
  for (int i = 0; i < NW; ++i)
    for (int m = width >> 1; m > 0; m >>= 1) v[i] = op(i, v[i], shfl_xor_u64(v[i], m));
  if (S_log <= 6) {
    *first = (int)threadIdx.x & ~(width - 1);
    return (lane & (width - 1)) == 0;
  }
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NW; ++i) lds[wave * NW + i] = v[i];
  __syncthreads();
  const int wps = S_log >= 8 ? 4 : (1 << (S_log - 6));  // waves per segment
  const int nseg = 4 / wps;
  if ((int)threadIdx.x >= nseg) return false;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    uint64_t r = lds[(threadIdx.x * wps) * NW + i];
    for (int w = 1; w < wps; ++w) r = op(i, r, lds[(threadIdx.x * wps + w) * NW + i]);
    v[i] = r;
  }
  *first = (int)threadIdx.x * wps * 64;
  return true;
}

}  // namespace pcgsq
