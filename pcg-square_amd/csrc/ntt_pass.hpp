// One pass of the negacyclic NTT over F_p, as a device function shared by the
// per-length launches (ntt.hip) and the persistent batched kernel
// (ntt_batch.hip).
//
// A pass runs C independent sub-transforms ("columns") of length n = 2^NLOG
// through all NLOG radix-2 butterfly stages.  The stages are cut into up to
// three register groups of RA, RB, RC stages (forward stage order).  Each
// thread owns one set of 2^r elements of a column per group: the first group
// in time loads its set straight from HBM into registers, the last one stores
// straight to HBM, and consecutive groups exchange through one padded,
// swizzled LDS tile.  A length-1024 pass (5 + 5 stages) therefore touches LDS
// once per element in each direction.
//
// Forward passes are Cooley-Tukey with merged roots (natural in, bit-reversed
// positions), values lazily reduced in [0, 4P) (Harvey); inverse passes are
// Gentleman-Sande in the mirrored order, values in [0, 2P).  The stage table
// holds (w, floor(w 2^64 / P)) pairs: stage s, butterfly block j uses entry
// 2^s + j.
#pragma once

#include <stdint.h>

#include "device_common.hpp"

namespace pcgsq {

// Where a pass reads and writes.  Strides are powers of two (log2, in
// elements); a column's element at position pos lives at row p = pos, or at
// p = bitreverse(pos) on the bit-reversed side of the pass (see PassKind).
struct NttPassDesc {
  const uint64_t* in;
  const uint64_t* in2;       // *Prod kinds: the load multiplies in[] by in2[] (mod P)
  uint64_t* out;
  const uint64_t* boff;      // word offset of each transform of the launch (in, in2 and out
                             // alike); nullptr: back to back, transform b at b * 2^N_log
  const uint64_t* tab;       // stage table [n][2]
  const uint64_t* tw;        // first pass of a four-step transform: its twiddle w, one word
                             // per element, at the element's word offset in the pass's
                             // output [N] (so a tile's twiddle loads coalesce as its stores
                             // do); the Shoup constant is formed from w (shoup_w_approx)
  uint64_t scale_w, scale_wq;  // whole-transform inverse: N^-1
  int c_log;                 // C = 2^c_log columns per tile
  int in_ps_log, in_cs_log, out_ps_log, out_cs_log;
  int N_log;                 // whole transform length (batch stride, twiddle modulus 2N)
  int ncols;                 // valid columns (a ragged last tile skips the rest)
};

// The kinds of pass, fixed at compile time so that the per-element load and
// store code is straight-line:
//   kFwdCols  four-step pass 1: natural rows in, bit-reversed positions out
//             with the twiddle psi^(c (2p + 1)), left in [0, 2P) (the first
//             passes, STORE == 1, do not reduce fully); lanes along the columns.
//   kFwdRows  four-step pass 2: contiguous sub-transforms in (lanes along the
//             positions), transposed store.
//   kInvRows / kInvRowsProd  inverse pass 1 (optionally multiplying two
//             forward-domain inputs on load), twiddle psi^-(p (2c + 1)) N^-1.
//   kInvCols  inverse pass 2.
//   kFwdWhole / kInvWhole / kInvWholeProd  a whole transform per column
//             (one pass; the inverse scales by N^-1).
enum PassKind : int {
  kFwdCols = 0, kFwdRows, kInvRows, kInvRowsProd, kInvCols, kFwdWhole, kInvWhole, kInvWholeProd
};

namespace nttpass {

__device__ __forceinline__ uint64_t red2p(uint64_t x) { return x >= kTwoP ? x - kTwoP : x; }
__device__ __forceinline__ uint64_t red_full(uint64_t x) {
  x = x >= kTwoP ? x - kTwoP : x;
  return x >= kP ? x - kP : x;
}

__device__ __forceinline__ uint32_t brv(uint32_t x, int bits) {
  return bits ? __brev(x) >> (32 - bits) : 0u;
}

// Device memory reached through a pointer the compiler cannot trace to a
// kernel argument (a work-list descriptor) would otherwise be accessed with
// flat instructions; the pass names its buffers as global memory explicitly.
#define PCGSQ_GLOBAL __attribute__((address_space(1)))
typedef const PCGSQ_GLOBAL uint64_t* gcptr;
typedef PCGSQ_GLOBAL uint64_t* gptr;

struct Pair {
  uint64_t w, wq;
};
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ Pair ld_pair(gcptr tab, uint32_t i) {
  const u64x2 t = reinterpret_cast<const PCGSQ_GLOBAL u64x2*>(tab)[i];
  return Pair{t.x, t.y};
}

// Two LDS layouts of a tile, chosen by the number of register groups:
//  * one or two groups: columns padded so that column c starts on bank slot
//    c * 32 / C, and the low five position bits XORed with the next five;
//  * three groups: one pad word after every 8 positions (word pos + pos / 8)
//    and 16 / C pad words per column for C >= 4.  A set's positions are its
//    base with the element index in disjoint bits, so every element's word is
//    the set's word plus a compile-time constant (an instruction offset, no
//    address arithmetic per element), and the conflicts of all three
//    groups' access orders stay at two-way or less.
__host__ __device__ constexpr int lds_pad(int C) { return C >= 2 ? (C >= 32 ? 1 : 32 / C) : 0; }

__host__ __device__ constexpr uint32_t lds_col_stride(int nlog, int C, int groups) {
  return groups == 3 ? (1u << nlog) + (1u << nlog >> 3) + (C >= 4 ? 16u / (uint32_t)C : 0u)
                     : (1u << nlog) + (uint32_t)lds_pad(C);
}

// Three-group layout: word of position pos within its column.
__host__ __device__ constexpr uint32_t lds_lin(uint32_t pos) { return pos + (pos >> 3); }

// One/two-group layout: LDS word of (column c, position pos).
template <int NLOG>
__device__ __forceinline__ uint32_t lds_at(uint32_t c, uint32_t pos, uint32_t col_stride) {
  uint32_t x = pos;
  if constexpr (NLOG >= 6) {
    constexpr uint32_t k = NLOG - 5 < 5 ? NLOG - 5 : 5;
    x ^= (pos >> 5) & ((1u << k) - 1);
  }
  return c * col_stride + x;
}

// r butterfly stages [S0, S0 + R) on one set held in registers.  Hh is the
// set's block index (its butterfly blocks in stage S0 + q are (Hh << q) | j).
// CYC: the table is a cyclic sub-transform's, whose block 0 of every stage
// has the twiddle 1 (entry 2^s); in the group that starts at stage 0 the
// block index is known at compile time, and those butterflies skip the
// multiply (about one butterfly per element of a length-2^n cyclic pass:
// 2^n - 1 of its n 2^(n-1)).
template <bool INV, bool CYC, int S0, int R>
__device__ __forceinline__ void group_butterflies(uint64_t (&v)[1 << R], uint32_t Hh,
                                                  gcptr __restrict__ tab) {
  constexpr int M = 1 << R;
  if constexpr (S0 == 0) Hh = 0;
  if constexpr (!INV) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int bit = R - 1 - q;
      // keep each stage's twiddle loads in that stage (bounds register use)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int w = 0; w < M; ++w) {
        if ((w >> bit) & 1) continue;
        const int w2 = w | (1 << bit);
        const uint64_t U = red2p(v[w]);
        uint64_t T;
        if (CYC && S0 == 0 && (w >> (R - q)) == 0) {
          T = red2p(v[w2]);  // twiddle 1: [0, 4P) -> [0, 2P), as a Shoup product leaves it
        } else {
          const Pair t = ld_pair(tab, (1u << (S0 + q)) + ((Hh << q) | (uint32_t)(w >> (R - q))));
          T = shoup_lazy_dev(v[w2], t.w, t.wq);
        }
        v[w] = U + T;
        v[w2] = U - T + kTwoP;
      }
    }
  } else {
#pragma unroll
    for (int q = R - 1; q >= 0; --q) {
      const int bit = R - 1 - q;
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int w = 0; w < M; ++w) {
        if ((w >> bit) & 1) continue;
        const int w2 = w | (1 << bit);
        const uint64_t U = v[w], V = v[w2];  // both < 2P
        v[w] = red2p(U + V);
        if (CYC && S0 == 0 && (w >> (R - q)) == 0) {
          v[w2] = red2p(U - V + kTwoP);  // twiddle 1
        } else {
          const Pair t = ld_pair(tab, (1u << (S0 + q)) + ((Hh << q) | (uint32_t)(w >> (R - q))));
          v[w2] = shoup_lazy_dev(U - V + kTwoP, t.w, t.wq);
        }
      }
    }
  }
}

// The same stages with the set's 2^R - 1 stage twiddles already in registers
// (stage q, block j at tw[2^q - 1 + j]): for a group after the first, whose
// twiddles depend on the set, all of them are loaded before the set's gather
// so that one wait covers their latency instead of one per stage.
template <bool INV, int R>
__device__ __forceinline__ void group_butterflies_pre(uint64_t (&v)[1 << R],
                                                      const Pair (&tw)[(1 << R) - 1]) {
  constexpr int M = 1 << R;
  if constexpr (!INV) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int bit = R - 1 - q;
#pragma unroll
      for (int w = 0; w < M; ++w) {
        if ((w >> bit) & 1) continue;
        const int w2 = w | (1 << bit);
        const Pair t = tw[(1 << q) - 1 + (w >> (R - q))];
        const uint64_t U = red2p(v[w]);
        const uint64_t T = shoup_lazy_dev(v[w2], t.w, t.wq);
        v[w] = U + T;
        v[w2] = U - T + kTwoP;
      }
    }
  } else {
#pragma unroll
    for (int q = R - 1; q >= 0; --q) {
      const int bit = R - 1 - q;
#pragma unroll
      for (int w = 0; w < M; ++w) {
        if ((w >> bit) & 1) continue;
        const int w2 = w | (1 << bit);
        const Pair t = tw[(1 << q) - 1 + (w >> (R - q))];
        const uint64_t U = v[w], V = v[w2];
        v[w] = red2p(U + V);
        v[w2] = shoup_lazy_dev(U - V + kTwoP, t.w, t.wq);
      }
    }
  }
}

// The 2^R - 1 stage twiddles of a set of a group starting at stage S0
// (stage q, block j at stw[2^q - 1 + j]), block index Hh.
template <int S0, int R>
__device__ __forceinline__ void load_stage_twiddles(Pair (&stw)[(1 << R) - 1],
                                                    gcptr __restrict__ tab, uint32_t Hh) {
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int j = 0; j < (1 << q); ++j)
      stw[(1 << q) - 1 + j] = ld_pair(tab, (1u << (S0 + q)) + ((Hh << q) | (uint32_t)j));
}

// The same with a wave-uniform Hh, read through the scalar cache: the stage
// table is written when the plan is made and never during a pass, so it is
// read as constant memory (scalar loads, off the vector memory path).
template <int S0, int R>
__device__ __forceinline__ void load_stage_twiddles_uniform(Pair (&stw)[(1 << R) - 1],
                                                            const uint64_t* tab, uint32_t Hh) {
  typedef const __attribute__((address_space(4))) u64x2* cptr2;
  const cptr2 t = (cptr2)(uintptr_t)tab;
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int j = 0; j < (1 << q); ++j) {
      const u64x2 v = t[(1u << (S0 + q)) + ((Hh << q) | (uint32_t)j)];
      stw[(1 << q) - 1 + j] = Pair{v.x, v.y};
    }
}

template <int K>
struct KindTraits {
  static constexpr bool INV = K == kInvRows || K == kInvRowsProd || K == kInvCols ||
                              K == kInvWhole || K == kInvWholeProd;
  static constexpr bool IN_BRV = INV;
  static constexpr bool OUT_BRV = !INV;
  static constexpr bool PROD = K == kInvRowsProd || K == kInvWholeProd;
  static constexpr int STORE = (K == kFwdCols || K == kInvRows || K == kInvRowsProd) ? 1
                               : (K == kInvWhole || K == kInvWholeProd)             ? 2
                                                                                    : 0;
  static constexpr bool TW_SWAP = K == kInvRows || K == kInvRowsProd;
  // the four-step row passes are cyclic sub-transforms (twiddle-1 blocks)
  static constexpr bool CYC = K == kFwdRows || K == kInvRows || K == kInvRowsProd;
  // lanes along the columns (unit column stride in HBM) or along the positions
  static constexpr bool LOAD_CFAST = K == kFwdCols || K == kInvRows || K == kInvRowsProd ||
                                     K == kInvCols;
  static constexpr bool STORE_CFAST = K == kFwdCols || K == kFwdRows || K == kInvCols;
};

// Role of a register group in the pass (kCopy: move the set, no stages).
enum : int { kFromHbm = 1, kToHbm = 2, kCopy = 4 };

// Shape facts of group `S0, R` of a pass.
template <int K, int NLOG, int S0, int R, int ROLE, bool TWPRE>
struct GroupTraits {
  using Tr = KindTraits<K>;
  static constexpr int TMIN = NLOG - S0 - R;  // log2 of the set stride
  // lanes run along the columns when the HBM side of this group has unit
  // column stride, else along the positions
  static constexpr bool cfast = (ROLE & kFromHbm) ? Tr::LOAD_CFAST
                                : (ROLE & kToHbm) ? Tr::STORE_CFAST
                                                  : true;
  // a later group: the set's stage twiddles, loaded before its gather (in
  // the batch kernel, whose registers are those of its largest shape, for
  // groups of up to 3 stages: 7 pairs)
  static constexpr bool STW = S0 > 0 && !(ROLE & kCopy) && (TWPRE || R <= 3);
};

// Group `S0, R` of a pass over a tile: for each set owned by this thread,
// gather (from HBM or LDS), run the stages, scatter (to LDS or HBM).
// UNI: the set's stage twiddles are wave-uniform and come by scalar loads.
template <int K, int NLOG, int S0, int R, int RMAX, int ROLE, int G, bool TWPRE, bool UNI>
__device__ __forceinline__ void run_group_sets(const NttPassDesc& d, uint64_t* __restrict__ lds,
                                               uint32_t t, uint32_t T, uint32_t C, uint32_t c0,
                                               gcptr __restrict__ inb, gcptr __restrict__ in2b,
                                               gptr __restrict__ outb, gcptr __restrict__ twb) {
  using GT = GroupTraits<K, NLOG, S0, R, ROLE, TWPRE>;
  constexpr int M = 1 << R;
  constexpr int TMIN = GT::TMIN;
  constexpr int SLOG = NLOG - R;              // log2 of sets per column
  constexpr int PER = 1 << (RMAX - R);        // sets per thread
  using Tr = KindTraits<K>;
  const uint32_t col_stride = lds_col_stride(NLOG, (int)C, G);
  constexpr bool cfast = GT::cfast;
#pragma unroll 1
  for (int k = 0; k < PER; ++k) {
    const uint32_t g = t + (uint32_t)k * T;
    uint32_t c, sigma;
    if constexpr (cfast) {
      c = g & (C - 1);
      sigma = g >> d.c_log;
    } else {
      sigma = g & ((1u << SLOG) - 1);
      c = g >> SLOG;
    }
    if (c0 + c >= (uint32_t)d.ncols) continue;
    const uint32_t L = sigma & ((1u << TMIN) - 1);
    const uint32_t Hh = sigma >> TMIN;
    const uint32_t base = (Hh << (TMIN + R)) | L;
    constexpr bool STW = GT::STW;
    Pair stw[STW ? M - 1 : 1];
    if constexpr (STW && UNI)
      load_stage_twiddles_uniform<S0, R>(stw, d.tab, __builtin_amdgcn_readfirstlane(Hh));
    else if constexpr (STW)
      load_stage_twiddles<S0, R>(stw, (gcptr)d.tab, Hh);
    // a storing first pass with sets of up to 8 elements (TWPRE): issue the
    // set's twiddle loads now (2 VGPRs per element), so that their latency
    // overlaps the gather and butterflies below
    constexpr bool TW = TWPRE && (ROLE & kToHbm) && Tr::STORE == 1 && M <= 8;
    uint64_t twa[TW ? M : 1];
    if constexpr (TW) {
#pragma unroll
      for (int w = 0; w < M; ++w) {
        const uint32_t pos = base | ((uint32_t)w << TMIN);
        const uint32_t p = Tr::OUT_BRV ? brv(pos, NLOG) : pos;
        twa[w] = twb[(p << d.out_ps_log) + (c << d.out_cs_log)];
      }
    }
    uint64_t v[M];
    if constexpr (ROLE & kFromHbm) {
#pragma unroll
      for (int w = 0; w < M; ++w) {
        if ((w & 7) == 0) __builtin_amdgcn_sched_barrier(0);
        const uint32_t pos = base | ((uint32_t)w << TMIN);
        const uint32_t p = Tr::IN_BRV ? brv(pos, NLOG) : pos;
        const uint32_t off = (p << d.in_ps_log) + (c << d.in_cs_log);
        v[w] = inb[off];
        if constexpr (Tr::PROD) v[w] = fmul(v[w], in2b[off]);
      }
    } else if constexpr (G == 3) {
      // base and w << TMIN have disjoint bits: lds_lin distributes over them
      const uint32_t sw = c * col_stride + lds_lin(base);
#pragma unroll
      for (int w = 0; w < M; ++w) v[w] = lds[sw + lds_lin((uint32_t)w << TMIN)];
    } else {
#pragma unroll
      for (int w = 0; w < M; ++w)
        v[w] = lds[lds_at<NLOG>(c, base | ((uint32_t)w << TMIN), col_stride)];
    }
    if constexpr (STW) group_butterflies_pre<Tr::INV, R>(v, stw);
    else if constexpr (!(ROLE & kCopy)) group_butterflies<Tr::INV, Tr::CYC, S0, R>(v, Hh, (gcptr)d.tab);
    if constexpr (ROLE & kToHbm) {
#pragma unroll
      for (int w = 0; w < M; ++w) {
        // bound the twiddle loads in flight (each is 4 VGPRs)
        if (!TW && (w & 3) == 0) __builtin_amdgcn_sched_barrier(0);
        const uint32_t pos = base | ((uint32_t)w << TMIN);
        const uint32_t p = Tr::OUT_BRV ? brv(pos, NLOG) : pos;
        const uint32_t off = (p << d.out_ps_log) + (c << d.out_cs_log);
        uint64_t x = v[w];
        if constexpr (Tr::STORE == 1) {
          // the four-step twiddle; with its approximate Shoup constant the
          // product lies in [0, 3P): a forward second pass takes it as it
          // is ([0, 4P)), an inverse one needs [0, 2P)
          const uint64_t tw = TW ? twa[TW ? w : 0] : twb[off];
          x = shoup_lazy_dev(x, tw, shoup_w_approx(tw));
          if constexpr (Tr::INV) x = red2p(x);
        } else if constexpr (Tr::STORE == 2) {
          x = shoup_lazy_dev(x, d.scale_w, d.scale_wq);
        }
        // a first pass hands its [0, 2P) values to the second as they are
        // (forward stages take [0, 4P), inverse stages [0, 2P)); only a
        // transform's final pass reduces to [0, P)
        outb[off] = Tr::STORE == 1 ? x : red_full(x);
      }
    } else if constexpr (G == 3) {
      const uint32_t sw = c * col_stride + lds_lin(base);
#pragma unroll
      for (int w = 0; w < M; ++w) lds[sw + lds_lin((uint32_t)w << TMIN)] = v[w];
    } else {
#pragma unroll
      for (int w = 0; w < M; ++w)
        lds[lds_at<NLOG>(c, base | ((uint32_t)w << TMIN), col_stride)] = v[w];
    }
  }
}

// A wave's 64 lanes hold sets of one block index when the lanes run along C
// columns with C 2^TMIN >= 64, or along the positions with 2^TMIN >= 64
// (waves and tiles are aligned to 64 threads): the stage twiddles are then
// wave-uniform and come by scalar loads, off the vector memory path.  The
// choice holds for the whole group, so it is made once, and each form of the
// set loop has its own registers.  Not in the batch kernel (TWPRE off),
// whose registers are those of its largest shape.
template <int K, int NLOG, int S0, int R, int RMAX, int ROLE, int G, bool TWPRE>
__device__ __forceinline__ void run_group(const NttPassDesc& d, uint64_t* __restrict__ lds,
                                          uint32_t t, uint32_t T, uint32_t C, uint32_t c0,
                                          gcptr __restrict__ inb, gcptr __restrict__ in2b,
                                          gptr __restrict__ outb, gcptr __restrict__ twb) {
  using GT = GroupTraits<K, NLOG, S0, R, ROLE, TWPRE>;
  constexpr bool ANY = GT::STW && TWPRE;
  constexpr bool ALWAYS = ANY && GT::TMIN >= 6;
  if constexpr (ALWAYS) {
    run_group_sets<K, NLOG, S0, R, RMAX, ROLE, G, TWPRE, true>(d, lds, t, T, C, c0, inb, in2b,
                                                               outb, twb);
  } else if constexpr (ANY && GT::cfast) {
    if ((C << GT::TMIN) >= 64u)
      run_group_sets<K, NLOG, S0, R, RMAX, ROLE, G, TWPRE, true>(d, lds, t, T, C, c0, inb, in2b,
                                                                 outb, twb);
    else
      run_group_sets<K, NLOG, S0, R, RMAX, ROLE, G, TWPRE, false>(d, lds, t, T, C, c0, inb, in2b,
                                                                  outb, twb);
  } else {
    run_group_sets<K, NLOG, S0, R, RMAX, ROLE, G, TWPRE, false>(d, lds, t, T, C, c0, inb, in2b,
                                                                outb, twb);
  }
}

// Stage layout of a pass: up to three groups RA, RB, RC (forward order).
template <int RA, int RB, int RC>
struct Part {
  static constexpr int NLOG = RA + RB + RC;
  static constexpr int G = RC ? 3 : (RB ? 2 : 1);
  static constexpr int RMAX = RA > RB ? (RA > RC ? RA : RC) : (RB > RC ? RB : RC);
  // threads a tile of C columns needs
  static __host__ __device__ constexpr int threads(int C) { return C << (NLOG - RMAX); }
};

// One tile (C columns starting at column tile `ct` of the transform that
// starts at word `base` of in, in2 and out).
// Threads t >= Part::threads(C) idle but take part in the barriers.
// TWPRE: prefetch a first pass's twiddles (more VGPRs; off where the
// registers of every shape add up, as in the persistent batch kernel).
template <int K, int RA, int RB, int RC, bool TWPRE = true>
__device__ __forceinline__ void pass_tile(const NttPassDesc& d, uint64_t* __restrict__ lds,
                                          size_t base, uint32_t ct) {
  using P = Part<RA, RB, RC>;
  constexpr int NLOG = P::NLOG, RMAX = P::RMAX;
  const uint32_t C = 1u << d.c_log;
  const uint32_t T = (uint32_t)P::threads((int)C);
  const uint32_t t = threadIdx.x;
  const bool active = t < T;
  const uint32_t c0 = ct << d.c_log;
  gcptr inb = (gcptr)(d.in + base + ((size_t)c0 << d.in_cs_log));
  gcptr in2b = KindTraits<K>::PROD ? (gcptr)(d.in2 + base + ((size_t)c0 << d.in_cs_log))
                                   : nullptr;
  gptr outb = (gptr)(d.out + base + ((size_t)c0 << d.out_cs_log));
  // the twiddle table is per length (shared by every transform of the launch)
  gcptr twb = KindTraits<K>::STORE == 1 ? (gcptr)(d.tw + ((size_t)c0 << d.out_cs_log))
                                        : nullptr;
  if constexpr (P::G == 1) {
    // one group: stage through LDS so that an in-place pass reads every
    // element before any bit-reversed store lands
    if (active) run_group<K, NLOG, 0, RA, RMAX, kFromHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
    __syncthreads();
    if (active)
      run_group<K, NLOG, 0, RA, RMAX, kToHbm | kCopy, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
  } else if constexpr (P::G == 2) {
    constexpr int SA = 0, SB = RA;
    if constexpr (!KindTraits<K>::INV) {
      if (active) run_group<K, NLOG, SA, RA, RMAX, kFromHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
      __syncthreads();
      if (active) run_group<K, NLOG, SB, RB, RMAX, kToHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
    } else {
      if (active) run_group<K, NLOG, SB, RB, RMAX, kFromHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
      __syncthreads();
      if (active) run_group<K, NLOG, SA, RA, RMAX, kToHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
    }
  } else {
    constexpr int SA = 0, SB = RA, SC = RA + RB;
    if constexpr (!KindTraits<K>::INV) {
      if (active) run_group<K, NLOG, SA, RA, RMAX, kFromHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
      __syncthreads();
      if (active) run_group<K, NLOG, SB, RB, RMAX, 0, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
      __syncthreads();
      if (active) run_group<K, NLOG, SC, RC, RMAX, kToHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
    } else {
      if (active) run_group<K, NLOG, SC, RC, RMAX, kFromHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
      __syncthreads();
      if (active) run_group<K, NLOG, SB, RB, RMAX, 0, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
      __syncthreads();
      if (active) run_group<K, NLOG, SA, RA, RMAX, kToHbm, P::G, TWPRE>(d, lds, t, T, C, c0, inb, in2b, outb, twb);
    }
  }
}

}  // namespace nttpass

// LDS bytes of a tile of C columns of length 2^nlog whose pass has `groups`
// register groups.
inline size_t ntt_tile_lds_bytes(int nlog, int C, int groups) {
  return (size_t)C * nttpass::lds_col_stride(nlog, C, groups) * sizeof(uint64_t);
}

}  // namespace pcgsq
