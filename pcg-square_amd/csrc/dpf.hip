// Half-tree DPF level kernels for gfx950 (see dpf.hpp for the protocol).
//
// Frontier layout per level: nodes F[T][j] and hashes H[T][j][k] (k < nk),
// T = tree, j = node index (MSB digit first, children of j at j*m + c, as in
// dpf.cpp:554-574).  One thread owns one parent: it rebuilds the m children
// from the parent's stored hashes and the level's CW, hashes each child with
// the tree PRG, writes the next frontier and XOR-accumulates the per-column
// share in registers.  Shares are reduced across the wave with lane shuffles,
// across the workgroup through LDS, and across workgroups with one 64-bit
// atomic XOR per (workgroup, word): XOR is associative and commutative, so
// the result is bit-exact whatever the order.
//
// The last tree level stores only its hashes, not its nodes: the leaf level
// rebuilds each parent node from the grandparent's stored node and hashes and
// the previous level's CW (XORs only, no PRG call).  That removes a third of
// the widest level's HBM traffic (16 of 48 bytes per node), and with it the
// power that held that level's clock down.
//
// The leaf level comes in three forms: store the leaves (drop-in entry
// points), write C|tau per leaf plus per-tree sums (convert), or, for the
// PCG-OLE expansion, do the output layer's gather as well (leaf_gather_kernel:
// per output coefficient, the sum of +-C and one tau bit per landing leaf).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "convert.hpp"
#include "dpf.hpp"
#include "prg.hpp"

namespace pcgsq {

namespace {

constexpr int kTpb = 256;
constexpr int kTpbLog = 8;
static_assert(kTpb == 1 << kTpbLog, "threads per workgroup");

constexpr int ilog2_constexpr(int x) { return x <= 1 ? 0 : 1 + ilog2_constexpr(x >> 1); }

// Tiles of kTpb parents per workgroup: a power of two (so a parent's partial
// slot is a shift, not a division), at most iters_max, at most S / kTpb.
int level_iters(size_t S, int iters_max) {
  if (S < (size_t)kTpb) return 1;
  size_t it = 1;
  while (it * 2 <= std::min<size_t>(S / kTpb, (size_t)std::max(1, iters_max))) it *= 2;
  return (int)it;
}

struct LevelArgs {
  const Blk* Fp;     // parents [ntrees][S] (unused when the parents are rebuilt)
  const Blk* Hp;     // parents' hashes [ntrees][S][nk]
  const Blk* Fg;     // grandparents [ntrees][S/m] and their hashes [ntrees][S/m][nk]
  const Blk* Hg;     //   (leaf modes that rebuild the parents)
  Blk* Fc;           // children [ntrees][S*m]        (hash mode: nullptr = not stored; store mode)
  Blk* Hc;           // children's hashes             (hash mode)
  uint64_t* ct;      // C|tau<<63 per leaf            (convert mode)
  uint64_t* part;    // partial sums [ntrees][nparts][2] (convert mode)
  Blk* cw;           // [levels][ntrees][nk] shares; level `level` accumulates
  const Blk* peer;   // [levels][npeer][nk]
  const uint32_t* aes_t0;
  const uint8_t* aes_rks;
  int peer_shift, npeer;
  int ntrees, S_log, level, iters, nparts;
  int chunk_log;     // log2(iters * threads): parents per workgroup (a power of two)
  // leaf level with the output layer's gather fused in (see leaf_gather_kernel)
  uint64_t* gA;
  size_t gA_party_stride;
  uint64_t* gmask;
  int mask_words, t, B, N_log, bin_log, sets, parties;
  // deep run: the subtree kernel's outputs are stored subtree-major (see
  // dpf_subtree_kernel); sub_n = nodes of level Ls + 1 (all trees)
  size_t sub_n;
};

enum Mode : int { kModeHash = 0, kModeStore = 1, kModeConvert = 2 };

// CW of tree T that builds the nodes of level `lev` + 1 from level `lev`.
template <int W>
__device__ __forceinline__ void load_cw(const LevelArgs& a, int T, int lev, Blk* CW) {
  constexpr int NK = (1 << W) - 1;
  const Blk* mine = a.cw + ((size_t)lev * a.ntrees + T) * NK;
  int pt = T + a.peer_shift;  // T < npeer and peer_shift < npeer
  if (pt >= a.npeer) pt -= a.npeer;
  const Blk* peer = a.peer + ((size_t)lev * a.npeer + pt) * NK;
  Blk fold{0, 0};
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    CW[k] = bxor(mine[k], peer[k]);
    fold = bxor(fold, CW[k]);
  }
  CW[NK] = fold;  // CW_{m-1} = XOR_k CW_k
}

// REBUILD (leaf modes): the parents' nodes were not stored; rebuild them
// from the grandparents as the level before did.
template <int W, int PRG, int MODE, bool REBUILD>
__global__ __launch_bounds__(kTpb) void dpf_level_kernel(LevelArgs a) {
  constexpr int M = 1 << W, NK = M - 1;
  __shared__ typename std::conditional<PRG == kPrgAes, AesLds, int>::type aes_lds_storage;
  __shared__ uint64_t red[4 * 2 * NK + 8];
  const AesLds* lds = nullptr;
  if constexpr (PRG == kPrgAes) {
    aes_stage(&aes_lds_storage, a.aes_t0, a.aes_rks, NK);
    lds = &aes_lds_storage;
  }
  const size_t S = (size_t)1 << a.S_log;
  const size_t total = (size_t)a.ntrees << a.S_log;
  const size_t base = (size_t)blockIdx.x * a.iters * kTpb;

  uint64_t acc[2 * NK];
#pragma unroll
  for (int i = 0; i < 2 * NK; ++i) acc[i] = 0;
  uint64_t accC = 0, accT = 0;

  // When the workgroup's parents all lie in one tree (S a multiple of the
  // chunk: every level that matters for time), the CWs are loaded once,
  // with scalar loads, instead of per parent.
  auto run = [&](auto one_tree) {
    constexpr bool ONE_TREE = decltype(one_tree)::value;
    Blk CWt[NK + 1], CWgt[NK + 1];
    if constexpr (ONE_TREE) {
      const int T0 = (int)(base >> a.S_log);
      load_cw<W>(a, T0, a.level - 1, CWt);
      if constexpr (REBUILD) load_cw<W>(a, T0, a.level - 2, CWgt);
    }
    for (int it = 0; it < a.iters; ++it) {
      const size_t f = base + (size_t)it * kTpb + threadIdx.x;
      if (f >= total) break;
      const int T = (int)(f >> a.S_log);
      Blk CW[NK + 1];
      if constexpr (ONE_TREE) {
#pragma unroll
        for (int k = 0; k <= NK; ++k) CW[k] = CWt[k];
      } else {
        load_cw<W>(a, T, a.level - 1, CW);
      }
      Blk parent;
      if constexpr (REBUILD) {
        // parent f is child c of grandparent f >> W (S is a multiple of m)
        const size_t g = f >> W;
        const int c = (int)(f & (M - 1));
        const Blk gp = a.Fg[g];
        Blk hg = {0, 0}, hgx = {0, 0};
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          const Blk v = a.Hg[g * NK + k];
          hgx = bxor(hgx, v);
          if (k == c) hg = v;
        }
        Blk CWg[NK + 1];
        if constexpr (ONE_TREE) {
#pragma unroll
          for (int k = 0; k <= NK; ++k) CWg[k] = CWgt[k];
        } else {
          load_cw<W>(a, T, a.level - 2, CWg);
        }
        parent = (c < NK) ? hg : bxor(hgx, gp);
        if ((gp.lo & 1ULL) != 0) parent = bxor(parent, CWg[c]);
      } else {
        parent = a.Fp[f];
      }
      Blk h[NK];
      Blk hx{0, 0};
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        h[k] = a.Hp[f * NK + k];
        hx = bxor(hx, h[k]);
      }
      const bool t = (parent.lo & 1ULL) != 0;
      if constexpr (MODE == kModeHash && W == 1) {
        // both children hashed before either is stored: the two 16-byte
        // nodes (and hashes) of a parent go out back to back, so that a
        // wave's stores cover whole lines instead of every other 16 bytes
        Blk node[2], hh[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          node[c] = c == 0 ? h[0] : bxor(hx, parent);
          if (t) node[c] = bxor(node[c], CW[c]);
          tree_hash<PRG, 1>(node[c], &hh[c], lds);
          acc[0] ^= hh[c].lo;
          acc[1] ^= hh[c].hi;
        }
        if (a.Fc) {
          a.Fc[2 * f] = node[0];
          a.Fc[2 * f + 1] = node[1];
        }
        a.Hc[2 * f] = hh[0];
        a.Hc[2 * f + 1] = hh[1];
        continue;
      }
#pragma unroll
      for (int c = 0; c < M; ++c) {
        Blk node = (c < NK) ? h[c < NK ? c : 0] : bxor(hx, parent);
        if (t) node = bxor(node, CW[c]);
        const size_t idx = f * M + c;
        if constexpr (MODE == kModeHash) {
          Blk hh[NK];
          tree_hash<PRG, NK>(node, hh, lds);
          if (a.Fc) a.Fc[idx] = node;
#pragma unroll
          for (int k = 0; k < NK; ++k) {
            a.Hc[idx * NK + k] = hh[k];
            acc[2 * k] ^= hh[k].lo;
            acc[2 * k + 1] ^= hh[k].hi;
          }
        } else if constexpr (MODE == kModeStore) {
          a.Fc[idx] = node;
        } else {
          const uint64_t C = freduce64(out_hash_lo64(node));
          const uint64_t tau = node.lo & 1ULL;
          a.ct[idx] = C | (tau << 63);
          accC = fadd(accC, C);
          accT += tau;
        }
      }
    }
  };
  if (a.S_log >= a.chunk_log)
    run(std::true_type{});
  else
    run(std::false_type{});

  int first = 0;
  if constexpr (MODE == kModeHash) {
    if (segment_reduce<2 * NK>(acc, a.S_log, [](int, uint64_t x, uint64_t y) { return x ^ y; },
                               red, &first)) {
      const int T = (int)((base + (size_t)first) >> a.S_log);
      if (T < a.ntrees) {
        Blk* dst = a.cw + ((size_t)a.level * a.ntrees + T) * NK;
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          atomicXor((unsigned long long*)&dst[k].lo, (unsigned long long)acc[2 * k]);
          atomicXor((unsigned long long*)&dst[k].hi, (unsigned long long)acc[2 * k + 1]);
        }
      }
    }
  } else if constexpr (MODE == kModeConvert) {
    // word 0: sum of C mod P; word 1: count of tau (plain)
    uint64_t v[2] = {accC, accT};
    if (segment_reduce<2>(v, a.S_log,
                          [](int i, uint64_t x, uint64_t y) { return i == 0 ? fadd(x, y) : x + y; },
                          red, &first)) {
      const size_t f = base + (size_t)first;
      const int T = (int)(f >> a.S_log);
      const size_t pidx = (a.S_log >= 8) ? ((f & (S - 1)) >> a.chunk_log) : 0;
      if (T < a.ntrees) {
        uint64_t* dst = a.part + ((size_t)T * a.nparts + pidx) * 2;
        dst[0] = v[0];
        dst[1] = v[1];
      }
    }
  }
}

// CW of tree T (the same on every lane) that builds the nodes of level
// `lev` + 1, read through the scalar cache: earlier launches wrote the
// shares, and nothing in the reading kernel writes them.
template <int W>
__device__ __forceinline__ void load_cw_uniform(const LevelArgs& a, int T, int lev, Blk* CW) {
  constexpr int NK = (1 << W) - 1;
  using cptr = const __attribute__((address_space(4))) uint64_t*;
  T = __builtin_amdgcn_readfirstlane(T);
  int pt = T + a.peer_shift;  // T < npeer and peer_shift < npeer
  if (pt >= a.npeer) pt -= a.npeer;
  const cptr mine = (cptr)(uintptr_t)(a.cw + ((size_t)lev * a.ntrees + T) * NK);
  const cptr peer = (cptr)(uintptr_t)(a.peer + ((size_t)lev * a.npeer + pt) * NK);
  Blk fold{0, 0};
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    CW[k] = Blk{mine[2 * k] ^ peer[2 * k], mine[2 * k + 1] ^ peer[2 * k + 1]};
    fold = bxor(fold, CW[k]);
  }
  CW[NK] = fold;
}

// The output layer's state in the leaf kernels, per thread: sums of +-C over
// the landing slots for its LPT output coefficients, their tau words, and the
// per-slot workgroup sums for the Beaver step.
//  * Sums: a leaf's out-hash word x = 2^62 h + l (h < 4) goes in raw, l into
//    A and h into Ah, so that A + Ah * (2^62 mod P) = sum of x mod P.  Three
//    l (each below 2^62) on an A below P stay below 2^64, so A is reduced
//    every third slot; Ah (at most 3 per slot) is folded in at the phase
//    switch and at the end.  The slots with h = 0 (+C) all come before those
//    with h = 1 (-C): at the switch the +C sums go to gA, and at the end
//    gA = (+C sums) - (-C sums).
//  * tau words: cur holds the current half (32 slots) of the current 64-slot
//    word; the low half goes out when the slots reach the high half, the rest
//    when the word is done (slots come in increasing order, one at a time).
//  * Slot sums: each thread's sum of its LPT C mod P and its tau count go
//    into LDS by atomics: the low and high 32 bits of the C sums separately
//    (below 2^40 and 2^38 over 256 threads), so the adds are exact.
struct SlotSums {
  unsigned long long c[kGatherMaxSlots][2];
  uint32_t t[kGatherMaxSlots];
  uint64_t at[kGatherMaxSlots];  // the partial each slot's sums go to
};

template <int LPT>
struct LeafOut {
  uint64_t A[LPT];
  uint32_t Ah[LPT], cur[LPT];
  int half = 0, nadd = 0, phase = 0, widx = 0;
  Blk* gdst;          // this thread's LPT sums in gA
  uint64_t* mask0;    // this thread's LPT words of tau word 0
  size_t word_stride;

  __device__ __forceinline__ LeafOut(Blk* g, uint64_t* m0, size_t ws)
      : gdst(g), mask0(m0), word_stride(ws) {
#pragma unroll
    for (int e = 0; e < LPT; ++e) {
      A[e] = 0;
      Ah[e] = cur[e] = 0;
    }
  }
  __device__ __forceinline__ void reduce() {
#pragma unroll
    for (int e = 0; e < LPT; ++e) A[e] = freduce64(A[e]);
    nadd = 0;
  }
  __device__ __forceinline__ void fold() {  // Ah * kC62 < 2^38 < P
#pragma unroll
    for (int e = 0; e < LPT; ++e) {
      A[e] = fadd(freduce64(A[e]), (uint64_t)Ah[e] * kC62);
      Ah[e] = 0;
    }
    nadd = 0;
  }
  __device__ __forceinline__ void store_half(int wi, int hf) {  // cur as half hf of word wi
    uint32_t* dst = reinterpret_cast<uint32_t*>(mask0 + (size_t)wi * word_stride) + hf;
#pragma unroll
    for (int e = 0; e < LPT; ++e) dst[2 * e] = cur[e];
  }
  __device__ __forceinline__ void finish_word(int wi) {  // zeros for an untouched high half
    if (half) {
      store_half(wi, 1);
    } else {
      uint64_t* dst = mask0 + (size_t)wi * word_stride;
#pragma unroll
      for (int e = 0; e < LPT; ++e) dst[e] = cur[e];
    }
  }
  // Before slot `slot` (sign h): word and half transitions, the phase switch,
  // the periodic reduction.
  __device__ __forceinline__ void begin_slot(int slot, int h) {
    if ((slot >> 6) != widx) {
      finish_word(widx);
      widx = slot >> 6;
      half = 0;
#pragma unroll
      for (int e = 0; e < LPT; ++e) cur[e] = 0;
    }
    if (!half && (slot & 32)) {
      store_half(widx, 0);
      half = 1;
#pragma unroll
      for (int e = 0; e < LPT; ++e) cur[e] = 0;
    }
    if (h != phase) {
      fold();
#pragma unroll
      for (int e = 0; e < LPT; e += 2) gdst[e / 2] = Blk{A[e], A[e + 1]};
#pragma unroll
      for (int e = 0; e < LPT; ++e) A[e] = 0;
      phase = h;
    } else if (nadd == 3) {
      reduce();
    }
    ++nadd;
  }
  // Leaf e of the slot: out-hash word x, tau bit; l goes to *L.
  __device__ __forceinline__ void leaf(int e, uint64_t x, uint32_t tau, int slot, uint64_t* L,
                                       uint32_t& sH, uint32_t& sT) {
    *L = x & ((1ULL << 62) - 1);
    const uint32_t hx = (uint32_t)(x >> 62);
    A[e] += *L;
    Ah[e] += hx;
    sH += hx;
    sT += tau;
    cur[e] |= tau << (slot & 31);
  }
  // After the slot's leaves: this thread's C sum into the slot's LDS sums.
  __device__ __forceinline__ void end_slot(SlotSums& r, int slot, const uint64_t (&L)[LPT],
                                           uint32_t sH, uint32_t sT) {
    uint64_t sC = (uint64_t)sH * kC62;  // sH <= 3 * LPT
#pragma unroll
    for (int g = 0; g < LPT; g += 4) {  // four l at a time stay below 2^64
      uint64_t s4 = L[g];
#pragma unroll
      for (int j = 1; j < 4 && g + j < LPT; ++j) s4 += L[g + j];
      sC = fadd(sC, freduce64(s4));
    }
    atomicAdd(&r.c[slot][0], (unsigned long long)(uint32_t)sC);
    atomicAdd(&r.c[slot][1], (unsigned long long)(sC >> 32));
    atomicAdd(&r.t[slot], sT);
  }
  // After the last slot: the remaining tau words and the sums.
  __device__ __forceinline__ void finish(int mask_words) {
    finish_word(widx);
    half = 0;
#pragma unroll
    for (int e = 0; e < LPT; ++e) cur[e] = 0;
    for (int wi = widx + 1; wi < mask_words; ++wi) finish_word(wi);
    fold();
    if (phase == 1) {  // gA holds the +C sums
#pragma unroll
      for (int e = 0; e < LPT; e += 2) {
        const Blk p = gdst[e / 2];
        gdst[e / 2] = Blk{fsub(p.lo, A[e]), fsub(p.hi, A[e + 1])};
      }
    } else {
#pragma unroll
      for (int e = 0; e < LPT; e += 2) gdst[e / 2] = Blk{A[e], A[e + 1]};
    }
  }
};

__device__ __forceinline__ void slot_sums_clear(SlotSums& r) {
  for (int i = threadIdx.x; i < kGatherMaxSlots; i += kTpb) {
    r.c[i][0] = r.c[i][1] = 0;
    r.t[i] = 0;
  }
  __syncthreads();
}

// After every thread's last slot (and a barrier): one partial per slot,
// (sum of C mod P, tau count), to a.part.
__device__ __forceinline__ void slot_sums_flush(const SlotSums& r, int nslots, uint64_t* part) {
  for (int sl = threadIdx.x; sl < nslots; sl += kTpb) {
    // low + 2^32 high, mod P (2^32 < P; both below 2^62 after freduce64)
    const uint64_t c = fadd(freduce64(r.c[sl][0]), fmul(freduce64(r.c[sl][1]), 1ULL << 32));
    uint64_t* dst = part + r.at[sl] * 2;
    dst[0] = c;
    dst[1] = r.t[sl];
  }
}

// Leaf level with the output layer's gather fused in.  A workgroup owns
// 256 * LPT consecutive output coefficients q of one (party, block) set, LPT
// per thread, all in one bin, and walks the leaves that land on them
// (for_each_landing: the same slots, in the same order, for every q of the
// tile).  For each landing slot it rebuilds the thread's parents from the
// grandparents (the last tree level stored only hashes), forms the thread's
// LPT leaves (consecutive, so the child indices are compile-time when
// LPT >= m), out-hashes them and adds +-C to the coefficients' sums.  tau
// goes into one bit per slot, because tau*CW needs the CW, and the CW needs
// every leaf's sums (the Beaver step).  A slot's leaves are 256 * LPT
// consecutive leaves of one tree, so their sums reduce over the workgroup
// into one partial per (tree, run of leaves).  This replaces a 2^30-word
// C|tau array written by the leaf level and read back by a separate gather.
template <int W, int LPT>
__global__ __launch_bounds__(kTpb) void leaf_gather_kernel(LevelArgs a) {
  constexpr int M = 1 << W, NK = M - 1;
  constexpr int TILE = kTpb * LPT;
  constexpr int CPP = LPT >= M ? M : LPT;  // the thread's leaves per parent
  constexpr int PPT = LPT / CPP;           // the thread's parents per slot
  __shared__ SlotSums red;
  slot_sums_clear(red);
  const int bb = blockIdx.y, party = blockIdx.z;  // party: index within this launch
  const uint32_t qt = blockIdx.x * TILE;
  const uint32_t q0 = qt + LPT * threadIdx.x;
  const size_t bin_mask = ((size_t)1 << a.bin_log) - 1;
  const size_t rr0 = q0 & bin_mask;  // a multiple of LPT
  const size_t S = (size_t)1 << a.S_log;
  const int nparts_log = a.S_log + W - (kTpbLog + ilog2_constexpr(LPT));
  const size_t set_off = ((size_t)party * a.sets + bb) << a.N_log;
  const size_t word_stride = ((size_t)a.parties * a.sets) << a.N_log;
  LeafOut<LPT> out(reinterpret_cast<Blk*>(a.gA + (size_t)party * a.gA_party_stride +
                                          ((size_t)bb << a.N_log) + q0),
                   a.gmask + set_off + q0, word_stride);
  const int nslots = for_each_landing(qt, a.N_log, a.bin_log, a.t, a.B, [&](int slot, int h,
                                                                             int ds, int inst) {
    out.begin_slot(slot, h);
    const int T = (party * a.sets + bb) * a.B + inst;
    Blk CW[NK + 1], CWg[NK + 1];
    load_cw_uniform<W>(a, T, a.level - 1, CW);
    load_cw_uniform<W>(a, T, a.level - 2, CWg);
    const size_t d0 = rr0 + ((size_t)ds << a.bin_log);  // a multiple of LPT
    uint64_t L[LPT];
    uint32_t sH = 0, sT = 0;
#pragma unroll
    for (int pp = 0; pp < PPT; ++pp) {
      const size_t dp = d0 + (size_t)pp * CPP;
      const size_t pi = (size_t)T * S + (dp >> W);  // parent (global index)
      const size_t gi = pi >> W;                    // its grandparent (S is a multiple of m)
      const int cp = (int)((dp >> W) & (M - 1));
      const Blk gp = a.Fg[gi];
      Blk hg[NK];
      Blk parent = gp, cwg = CWg[NK];  // child m-1: (xor of the hashes) ^ grandparent
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        hg[k] = a.Hg[gi * NK + k];
        parent = bxor(parent, hg[k]);
      }
#pragma unroll
      for (int k = 0; k < NK; ++k)
        if (k == cp) {
          parent = hg[k];
          cwg = CWg[k];
        }
      if ((gp.lo & 1ULL) != 0) parent = bxor(parent, cwg);
      Blk hp[NK];
      Blk hx{0, 0};
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        hp[k] = a.Hp[pi * NK + k];
        hx = bxor(hx, hp[k]);
      }
      const bool tp = (parent.lo & 1ULL) != 0;
#pragma unroll
      for (int e = 0; e < CPP; ++e) {
        Blk leaf, cwc;
        if constexpr (LPT >= M) {  // the leaves are children 0..m-1
          leaf = e < NK ? hp[e < NK ? e : 0] : bxor(hx, parent);
          cwc = CW[e];
        } else {  // children r*LPT + e of the parent, r = (dp mod m) / LPT
          const int r = (int)(dp & (M - 1)) / LPT;
          leaf = bxor(hx, parent);
          cwc = CW[NK];
#pragma unroll
          for (int rr = 0; rr < M / LPT; ++rr) {
            constexpr int kLast = NK;
            const int c = rr * LPT + e;  // compile-time
            if (rr == r && c != kLast) {
              leaf = hp[c < NK ? c : 0];
              cwc = CW[c];
            }
          }
        }
        if (tp) leaf = bxor(leaf, cwc);
        const int q = pp * CPP + e;
        out.leaf(q, out_hash_lo64(leaf), (uint32_t)leaf.lo & 1u, slot, &L[q], sH, sT);
      }
    }
    out.end_slot(red, slot, L, sH, sT);
    if (threadIdx.x == 0)
      red.at[slot] = ((uint64_t)T << nparts_log) +
                     (((qt & bin_mask) + ((size_t)ds << a.bin_log)) >> (kTpbLog + ilog2_constexpr(LPT)));
  });
  out.finish(a.mask_words);
  __syncthreads();
  slot_sums_flush(red, nslots, a.part);
}

// ---- both parties in one run: the last tree levels fused into the leaf level --
//
// With both parties' trees in one run, the correction words of the last
// three tree levels can be computed from the alpha path alone: nodes off the
// path are equal in the two parties, so their hashes cancel in the XOR of
// the two parties' level shares, and CW_L = H(path node of party 0) ^
// H(path node of party 1) ^ DltSft0_L ^ DltSft1_L, the value the fused
// protocol reduces to (dpf.cpp:520-551).  alpha's digit at level L is where
// DltSft0_L ^ DltSft1_L is nonzero ([digit_L(alpha) = k] Delta, Delta != 0).
// The leaf level then expands those three levels and the leaves in registers
// from the stored hashes of level levels - 4 (deep_leaf_kernel), and no wider
// level is ever stored.  Level levels - 4 stores only its hashes: its nodes
// are rebuilt from level levels - 5's nodes and hashes by XORs.  w = 1,
// ChaCha8 tree PRG.

__device__ __forceinline__ Blk child1(Blk parent, Blk h, int c, Blk cw) {
  Blk node = c ? bxor(h, parent) : h;   // w = 1: child 0 = h, child 1 = h ^ parent
  // the parent's t bit as one 32-bit mask for the four words of cw
  const uint32_t m = 0u - ((uint32_t)parent.lo & 1u);
  node.lo ^= ((uint64_t)((uint32_t)(cw.lo >> 32) & m) << 32) | ((uint32_t)cw.lo & m);
  node.hi ^= ((uint64_t)((uint32_t)(cw.hi >> 32) & m) << 32) | ((uint32_t)cw.hi & m);
  return node;
}

// One thread per instance (trees Ti and Ti + half).  Levels L0 - 1 and below
// are complete in cw; levels L0 .. levels - 1 still hold the seeded DltSft
// shares and receive CW_L in the first tree's slot (0 in the second's, so
// that mine ^ peer is CW_L for both).  Level L0 - 1 is stored in full (nodes
// F, hashes H).
__global__ __launch_bounds__(kTpb) void dpf_path_cw_kernel(const Blk* __restrict__ F,
                                                           const Blk* __restrict__ H,
                                                           Blk* __restrict__ cw,
                                                           const Blk* __restrict__ dlt0,
                                                           const Blk* __restrict__ dlt1,
                                                           int half, int levels, int L0) {
  const int Ti = blockIdx.x * kTpb + threadIdx.x;
  if (Ti >= half) return;
  const size_t ntrees = 2 * (size_t)half;
  auto digit = [&](int L) {
    const Blk a = dlt0[(size_t)Ti * levels + L], b = dlt1[(size_t)Ti * levels + L];
    return ((a.lo ^ b.lo) | (a.hi ^ b.hi)) ? 0 : 1;
  };
  size_t j = 0;   // path node of level L0 - 1
  for (int L = 0; L < L0 - 1; ++L) j = (j << 1) | (size_t)digit(L);
  const size_t S = (size_t)1 << (L0 - 1);
  const size_t i0 = (size_t)Ti * S + j, i1 = ((size_t)Ti + half) * S + j;
  Blk n0 = F[i0], n1 = F[i1];
  Blk h0 = H[i0], h1 = H[i1];
  const Blk* c0 = &cw[(size_t)(L0 - 1) * ntrees + Ti];
  Blk cwv = bxor(c0[0], c0[half]);
  int c = digit(L0 - 1);
  n0 = child1(n0, h0, c, cwv);
  n1 = child1(n1, h1, c, cwv);
  for (int L = L0; L < levels; ++L) {
    tree_hash<kPrgChaCha8, 1>(n0, &h0, nullptr);
    tree_hash<kPrgChaCha8, 1>(n1, &h1, nullptr);
    Blk* m = &cw[(size_t)L * ntrees + Ti];
    cwv = bxor(bxor(h0, h1), bxor(m[0], m[half]));
    m[0] = cwv;
    m[half] = Blk{0, 0};
    if (L + 1 < levels) {
      c = digit(L);
      n0 = child1(n0, h0, c, cwv);
      n1 = child1(n1, h1, c, cwv);
    }
  }
}

// Tree levels Ls + 1 .. Ls + 5 of both parties' run in one launch, their CWs
// known from the alpha path (no level sums).  A thread owns one node of level
// Ls + 1: it forms it from its stored parent (level Ls: a.Fp, a.Hp), then
// hashes its subtree depth-first in registers, 1 + 2 + 4 + 8 + 16 ChaCha8
// blocks, and stores level Ls + 4 in full (a.Fc, a.Hc) and the hashes of
// level Ls + 5 (a.Hg).  The stores are subtree-major, so that a wave's lanes
// write consecutive words: node s of the subtree of level-(Ls+1) node f
// (global index f * 2^d + s at depth d) sits at s * sub_n + f.  a.S_log =
// Ls.  UNIFORM: a workgroup's nodes lie in one tree (2^(Ls+1) >= 256), so
// the CWs are scalar loads.
template <bool UNIFORM>
__global__ __launch_bounds__(kTpb) void dpf_subtree_kernel(LevelArgs a) {
  const size_t f = (size_t)blockIdx.x * kTpb + threadIdx.x;  // level Ls + 1, global index
  const int L1 = a.S_log + 1;
  if (!UNIFORM && f >= ((size_t)a.ntrees << L1)) return;
  const int T = (int)(f >> L1);
  Blk cw[5][2];  // levels Ls .. Ls + 4 (w = 1: CW_0 and the fold are equal)
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    if constexpr (UNIFORM)
      load_cw_uniform<1>(a, T, a.S_log + l, cw[l]);
    else
      load_cw<1>(a, T, a.S_log + l, cw[l]);
  }
  Blk* __restrict__ F4 = a.Fc;
  Blk* __restrict__ H4 = a.Hc;
  Blk* __restrict__ H5 = const_cast<Blk*>(a.Hg);
  const Blk n1 = child1(a.Fp[f >> 1], a.Hp[f >> 1], (int)(f & 1), cw[0][0]);
  Blk h1;
  tree_hash<kPrgChaCha8, 1>(n1, &h1, nullptr);
#pragma unroll 1
  for (int x = 0; x < 2; ++x) {
    const Blk n2 = child1(n1, h1, x, cw[1][0]);
    Blk h2;
    tree_hash<kPrgChaCha8, 1>(n2, &h2, nullptr);
#pragma unroll 1
    for (int y = 0; y < 2; ++y) {
      const Blk n3 = child1(n2, h2, y, cw[2][0]);
      Blk h3;
      tree_hash<kPrgChaCha8, 1>(n3, &h3, nullptr);
#pragma unroll
      for (int z = 0; z < 2; ++z) {
        const Blk n4 = child1(n3, h3, z, cw[3][0]);
        Blk h4;
        tree_hash<kPrgChaCha8, 1>(n4, &h4, nullptr);
        const size_t s4 = ((size_t)x << 2) | ((size_t)y << 1) | (size_t)z;
        F4[s4 * a.sub_n + f] = n4;
        H4[s4 * a.sub_n + f] = h4;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const Blk n5 = child1(n4, h4, u, cw[4][0]);
          Blk h5;
          tree_hash<kPrgChaCha8, 1>(n5, &h5, nullptr);
          H5[((s4 << 1) | (size_t)u) * a.sub_n + f] = h5;
        }
      }
    }
  }
}

// The leaf level of both parties' run with the last three tree levels fused
// in (w = 1, 8 output coefficients per thread).  Per landing slot a thread's
// 8 consecutive leaves d0 .. d0 + 7 descend from one node of level
// levels - 3; the thread rebuilds that node's parent (level levels - 4) from
// its stored hash and its own parent's node and hashes (both written by
// dpf_subtree_kernel, subtree-major), forms the node, and expands 1 + 2 + 4
// tree hashes and 8 out-hashes in registers.  Every (tree, node) is expanded
// by exactly one thread, so the work is the same as hashing the levels one by
// one, without their frontiers.  The output layer's sums, tau bits and Beaver partials are
// those of leaf_gather_kernel.
__global__ __launch_bounds__(kTpb) void deep_leaf_kernel(LevelArgs a) {
  constexpr int LPT = 8;
  constexpr int TILE = kTpb * LPT;
  __shared__ SlotSums red;
  slot_sums_clear(red);
  const int bb = blockIdx.y, party = blockIdx.z;
  const uint32_t qt = blockIdx.x * TILE;
  const uint32_t q0 = qt + LPT * threadIdx.x;
  const size_t bin_mask = ((size_t)1 << a.bin_log) - 1;
  const size_t rr0 = q0 & bin_mask;  // a multiple of LPT
  const int levels = a.level;        // leaves are level `levels`
  const int Ltop = levels - 4;       // the last stored level (its hashes; nodes one level up)
  const size_t Stop = (size_t)1 << Ltop;
  const int nparts_log = levels - (kTpbLog + 3);
  const size_t set_off = ((size_t)party * a.sets + bb) << a.N_log;
  const size_t word_stride = ((size_t)a.parties * a.sets) << a.N_log;
  LeafOut<LPT> out(reinterpret_cast<Blk*>(a.gA + (size_t)party * a.gA_party_stride +
                                          ((size_t)bb << a.N_log) + q0),
                   a.gmask + set_off + q0, word_stride);
  const int nslots = for_each_landing(qt, a.N_log, a.bin_log, a.t, a.B, [&](int slot, int h,
                                                                             int ds, int inst) {
    out.begin_slot(slot, h);
    const int T = (party * a.sets + bb) * a.B + inst;
    Blk cw[5][2];   // levels Ltop - 1 .. levels - 1 (w = 1: CW_0 and the fold are equal)
#pragma unroll
    for (int l = 0; l < 5; ++l) load_cw_uniform<1>(a, T, Ltop - 1 + l, cw[l]);
    const size_t d0 = rr0 + ((size_t)ds << a.bin_log);  // a multiple of 8
    // level Ltop node gi and its parent, stored subtree-major under their
    // level Ltop - 4 ancestor (dpf_subtree_kernel)
    const size_t gi = (size_t)T * Stop + (d0 >> 4);      // level Ltop
    const size_t gp = gi >> 1;                           // level Ltop - 1
    const size_t ip = (gp & 7) * a.sub_n + (gp >> 3), ih = (gi & 15) * a.sub_n + (gi >> 4);
    const Blk n13 = child1(a.Fg[ip], a.Hg[ip], (int)(gi & 1), cw[0][0]);
    const Blk n14 = child1(n13, a.Hp[ih], (int)((d0 >> 3) & 1), cw[1][0]);
    Blk h14;
    tree_hash<kPrgChaCha8, 1>(n14, &h14, nullptr);
    Blk n15[2], h15[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      n15[c] = child1(n14, h14, c, cw[2][0]);
      tree_hash<kPrgChaCha8, 1>(n15[c], &h15[c], nullptr);
    }
    Blk n16[4], h16[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      n16[i] = child1(n15[i >> 1], h15[i >> 1], i & 1, cw[3][0]);
      tree_hash<kPrgChaCha8, 1>(n16[i], &h16[i], nullptr);
    }
    uint64_t L[LPT];
    uint32_t sH = 0, sT = 0;
#pragma unroll
    for (int e = 0; e < LPT; ++e) {
      const Blk leaf = child1(n16[e >> 1], h16[e >> 1], e & 1, cw[4][0]);
      out.leaf(e, out_hash_lo64(leaf), (uint32_t)leaf.lo & 1u, slot, &L[e], sH, sT);
    }
    out.end_slot(red, slot, L, sH, sT);
    if (threadIdx.x == 0)
      red.at[slot] = ((uint64_t)T << nparts_log) +
                     (((qt & bin_mask) + ((size_t)ds << a.bin_log)) >> (kTpbLog + 3));
  });
  out.finish(a.mask_words);
  __syncthreads();
  slot_sums_flush(red, nslots, a.part);
}

// Level 0: the frontier is the root.  Also moves DltSft from the reference
// layout [T][levels][nk] into the level-major share array [levels][T][nk]
// and seeds level 0 with H(root) (dpf.cpp:502, :522-532).
template <int W, int PRG>
__global__ __launch_bounds__(kTpb) void dpf_root_kernel(const Blk* __restrict__ roots,
                                                        const Blk* __restrict__ dlt,
                                                        const Blk* __restrict__ roots1,
                                                        const Blk* __restrict__ dlt1,
                                                        Blk* __restrict__ F0, Blk* __restrict__ H0,
                                                        Blk* __restrict__ cw, int ntrees,
                                                        int levels, const uint32_t* aes_t0,
                                                        const uint8_t* aes_rks) {
  constexpr int NK = (1 << W) - 1;
  __shared__ typename std::conditional<PRG == kPrgAes, AesLds, int>::type aes_lds_storage;
  const AesLds* lds = nullptr;
  if constexpr (PRG == kPrgAes) {
    aes_stage(&aes_lds_storage, aes_t0, aes_rks, NK);
    lds = &aes_lds_storage;
  }
  const int T = blockIdx.x * kTpb + threadIdx.x;
  if (T >= ntrees) return;
  const int half = ntrees / 2;
  const bool second = roots1 != nullptr && T >= half;
  const int Ti = second ? T - half : T;
  const Blk* dsrc = second ? dlt1 : dlt;
  for (int i = 0; i < levels; ++i)
    for (int k = 0; k < NK; ++k)
      cw[((size_t)i * ntrees + T) * NK + k] = dsrc[((size_t)Ti * levels + i) * NK + k];
  const Blk r = second ? roots1[Ti] : roots[T];
  F0[T] = r;
  Blk h[NK];
  tree_hash<PRG, NK>(r, h, lds);
  for (int k = 0; k < NK; ++k) {
    H0[(size_t)T * NK + k] = h[k];
    Blk* d = &cw[(size_t)T * NK + k];
    *d = bxor(*d, h[k]);
  }
}

template <int W, int PRG>
void launch_root(hipStream_t st, const DpfRun& r, Blk* F0, Blk* H0, Blk* cw) {
  const int levels = r.n / r.w;
  const int grid = (r.ntrees + kTpb - 1) / kTpb;
  hipLaunchKernelGGL((dpf_root_kernel<W, PRG>), dim3(grid), dim3(kTpb), 0, st, r.roots, r.dltsft,
                     r.roots1, r.dltsft1, F0, H0, cw, r.ntrees, levels, r.aes_t0, r.aes_rks);
  hip_check(hipGetLastError(), "launch dpf_root_kernel");
}

template <int W, int PRG, int MODE, bool REBUILD = false>
void launch_level(hipStream_t st, LevelArgs a, Profiler* prof) {
  char zone[32];
  if (MODE == kModeHash)
    snprintf(zone, sizeof zone, "dpf.level.%02d", a.level);
  else
    snprintf(zone, sizeof zone, "%s", MODE == kModeStore ? "dpf.leaf_store" : "dpf.leaf_convert");
  ProfZone pz(prof, st, zone);
  const size_t total = (size_t)a.ntrees << a.S_log;
  const size_t tiles = (total + kTpb - 1) / kTpb;
  const size_t grid = (tiles + a.iters - 1) / a.iters;
  require(grid < (1ULL << 31), "dpf: launch too large");
  hipLaunchKernelGGL((dpf_level_kernel<W, PRG, MODE, REBUILD>), dim3((unsigned)grid), dim3(kTpb), 0,
                     st, a);
  hip_check(hipGetLastError(), "launch dpf_level_kernel");
}

template <int W, int LPT>
void launch_leaf_gather_t(hipStream_t st, const LevelArgs& a) {
  const dim3 grid((unsigned)(((size_t)1 << a.N_log) / ((size_t)kTpb * LPT)), (unsigned)a.sets,
                  (unsigned)a.parties);
  hipLaunchKernelGGL((leaf_gather_kernel<W, LPT>), grid, dim3(kTpb), 0, st, a);
  hip_check(hipGetLastError(), "launch leaf_gather_kernel");
}

template <int W>
void launch_leaf_gather(hipStream_t st, const LevelArgs& a, Profiler* prof) {
  ProfZone pz(prof, st, "dpf.leaf_gather");
  switch (dpf_gather_lpt(a.bin_log + 1)) {
    case 8: return launch_leaf_gather_t<W, 8>(st, a);
    case 4: return launch_leaf_gather_t<W, 4>(st, a);
    default: return launch_leaf_gather_t<W, 2>(st, a);
  }
}

template <int W, int PRG>
void expand_t(hipStream_t st, const DpfRun& r, DpfScratch& s) {
  constexpr int M = 1 << W, NK = M - 1;
  const int n = r.n, levels = n / W;
  const bool deep = dpf_deep_leaf(r);
  dpf_reserve(s, n, W, (size_t)r.ntrees, r.host_ex != nullptr, deep);

  auto frontier = [&](int which, size_t S, Blk*& F, Blk*& H) {
    F = s.fh[which].as<Blk>();
    H = F + (size_t)r.ntrees * S;
  };

  Blk *F, *H;
  frontier(0, 1, F, H);
  {
    ProfZone pz(r.prof, st, "dpf.root");
    launch_root<W, PRG>(st, r, F, H, s.cw.as<Blk>());
  }

  std::vector<Blk> h_mine, h_peer;
  const size_t cw_level = (size_t)r.ntrees * NK;
  if (r.host_ex) {
    h_mine.resize(cw_level);
    h_peer.resize(cw_level);
  }

  int cur = 0;
  for (int i = 1; i <= levels; ++i) {
    // share of level i-1 is complete: exchange it
    if (r.host_ex) {
      Blk* dmine = s.cw.as<Blk>() + (size_t)(i - 1) * cw_level;
      Blk* dpeer = s.peer.as<Blk>() + (size_t)(i - 1) * cw_level;
      hip_check(hipMemcpyAsync(h_mine.data(), dmine, cw_level * sizeof(Blk),
                               hipMemcpyDeviceToHost, st), "D2H dpf cw shares");
      hip_check(hipStreamSynchronize(st), "sync dpf cw shares");
      const int rc = r.host_ex(r.host_user, i - 1, h_mine.data(), h_peer.data(), cw_level);
      if (rc != 0) throw Error(kErrCallback, "dpf exchange callback failed at level " +
                                                 std::to_string(i - 1));
      hip_check(hipMemcpyAsync(dpeer, h_peer.data(), cw_level * sizeof(Blk),
                               hipMemcpyHostToDevice, st), "H2D dpf peer shares");
    }
    const int S_log = W * (i - 1);
    const size_t S = (size_t)1 << S_log;
    LevelArgs a{};
    frontier(cur, S, F, H);
    a.Fp = F;
    a.Hp = H;
    a.cw = s.cw.as<Blk>();
    if (r.host_ex) {
      a.peer = s.peer.as<Blk>();
      a.npeer = r.ntrees;
      a.peer_shift = 0;
    } else {
      a.peer = s.cw.as<Blk>();
      a.npeer = r.ntrees;
      a.peer_shift = r.ntrees / 2;
    }
    a.aes_t0 = r.aes_t0;
    a.aes_rks = r.aes_rks;
    a.ntrees = r.ntrees;
    a.S_log = S_log;
    a.level = i;
    a.iters = level_iters(S, r.iters_max);
    a.chunk_log = ilog2_exact((uint64_t)a.iters) + ilog2_exact((uint64_t)kTpb);
    // the last tree level keeps only its hashes; the leaf level rebuilds
    // those parents from the grandparents (still in the other buffer)
    if (deep && i == levels - 8) {
      // Level Ls = levels - 9 is the last one the level kernels build (their
      // sums give CW_0 .. CW_Ls).  The CWs of the levels below come from the
      // alpha path; levels Ls + 1 .. levels - 4 are expanded in one launch
      // (level levels - 5 stored in full, levels - 4 hashes only), then the
      // leaf level with the last three tree levels fused in.
      const int Ls = levels - 9;
      const int half = r.ntrees / 2;
      const Blk* dlt1 = r.dltsft1 ? r.dltsft1 : r.dltsft + (size_t)half * levels;
      {
        ProfZone pz(r.prof, st, "dpf.path_cw");
        hipLaunchKernelGGL(dpf_path_cw_kernel, dim3((unsigned)((half + kTpb - 1) / kTpb)),
                           dim3(kTpb), 0, st, F, H, s.cw.as<Blk>(), r.dltsft, dlt1, half,
                           levels, Ls + 1);
        hip_check(hipGetLastError(), "launch dpf_path_cw_kernel");
      }
      // level Ls + 4 in full into the other buffer; level Ls + 5's hashes
      // into this one, behind level Ls (2^4 times wider, so no overlap)
      Blk *F4, *H4, *F5, *H5;
      frontier(cur ^ 1, S << (W * 4), F4, H4);
      frontier(cur, S << (W * 5), F5, H5);
      (void)F5;
      a.Fc = F4;
      a.Hc = H4;
      a.Hg = H5;
      a.sub_n = (size_t)r.ntrees << (Ls + 1);
      {
        ProfZone pz(r.prof, st, "dpf.subtree");
        const size_t threads = (size_t)r.ntrees << (Ls + 1);
        const dim3 grid((unsigned)((threads + kTpb - 1) / kTpb));
        if (Ls + 1 >= kTpbLog)
          hipLaunchKernelGGL(dpf_subtree_kernel<true>, grid, dim3(kTpb), 0, st, a);
        else
          hipLaunchKernelGGL(dpf_subtree_kernel<false>, grid, dim3(kTpb), 0, st, a);
        hip_check(hipGetLastError(), "launch dpf_subtree_kernel");
      }
      a.Fp = nullptr;
      a.Hp = H5;
      a.Fg = F4;
      a.Hg = H4;
      a.Fc = nullptr;
      a.Hc = nullptr;
      a.part = r.partials;
      a.gA = r.g_acc;
      a.gA_party_stride = r.g_party_stride;
      a.gmask = r.g_mask;
      a.mask_words = r.mask_words;
      a.t = r.t;
      a.B = r.B;
      a.N_log = r.N_log;
      a.bin_log = n - 1;
      a.sets = r.sets;
      a.parties = r.parties;
      a.level = levels;
      ProfZone pz(r.prof, st, "dpf.leaf_deep");
      const dim3 grid((unsigned)(((size_t)1 << a.N_log) / ((size_t)kTpb * 8)), (unsigned)a.sets,
                      (unsigned)a.parties);
      hipLaunchKernelGGL(deep_leaf_kernel, grid, dim3(kTpb), 0, st, a);
      hip_check(hipGetLastError(), "launch deep_leaf_kernel");
      break;
    }
    const bool rebuild = i == levels && levels >= 2;
    if (rebuild) {
      Blk *Fg, *Hg;
      frontier(cur ^ 1, S >> W, Fg, Hg);
      a.Fp = nullptr;
      a.Fg = Fg;
      a.Hg = Hg;
    }
    if (i < levels) {
      Blk *Fn, *Hn;
      frontier(cur ^ 1, S * M, Fn, Hn);
      // the last tree level keeps only its hashes (the leaf level rebuilds
      // its nodes by XORs)
      a.Fc = i == levels - 1 ? nullptr : Fn;
      a.Hc = Hn;
      launch_level<W, PRG, kModeHash>(st, a, r.prof);
      cur ^= 1;
    } else if (r.mode == kLeafStore) {
      a.Fc = r.leaves;
      if (rebuild)
        launch_level<W, PRG, kModeStore, true>(st, a, r.prof);
      else
        launch_level<W, PRG, kModeStore>(st, a, r.prof);
    } else if (r.mode == kLeafGather) {
      a.part = r.partials;
      a.gA = r.g_acc;
      a.gA_party_stride = r.g_party_stride;
      a.gmask = r.g_mask;
      a.mask_words = r.mask_words;
      a.t = r.t;
      a.B = r.B;
      a.N_log = r.N_log;
      a.bin_log = n - 1;
      a.sets = r.sets;
      a.parties = r.parties;
      launch_leaf_gather<W>(st, a, r.prof);
    } else {
      a.ct = r.ct;
      a.part = r.partials;
      a.nparts = dpf_leaf_nparts(n, W, r.iters_max);
      if (rebuild)
        launch_level<W, PRG, kModeConvert, true>(st, a, r.prof);
      else
        launch_level<W, PRG, kModeConvert>(st, a, r.prof);
    }
  }
  if (r.cw_out)
    hip_check(hipMemcpyAsync(r.cw_out, s.cw.as<Blk>(), (size_t)levels * r.ntrees * NK * sizeof(Blk),
                             hipMemcpyDeviceToDevice, st),
              "copy dpf cw");
}

}  // namespace

int dpf_gather_lpt(int n) {
  const size_t bin = (size_t)1 << (n - 1);
  return bin >= (size_t)kTpb * 8 ? 8 : bin >= (size_t)kTpb * 4 ? 4 : 2;
}

int dpf_gather_nparts(int n) {
  return (int)(((size_t)1 << n) / ((size_t)kTpb * dpf_gather_lpt(n)));
}

int dpf_leaf_nparts(int n, int w, int iters_max) {
  const int levels = n / w;
  const size_t S = (size_t)1 << (w * (levels - 1));
  if (S < (size_t)kTpb) return 1;
  return (int)(S / ((size_t)level_iters(S, iters_max) * kTpb));
}

bool dpf_deep_leaf(int n, int w, int prg, int mode, bool host_exchange, int parties) {
  return w == 1 && prg == kPrgChaCha8 && mode == kLeafGather && !host_exchange && parties == 2 &&
         n >= 5 && dpf_gather_lpt(n) == 8;
}

bool dpf_deep_leaf(const DpfRun& r) {
  return dpf_deep_leaf(r.n, r.w, r.prg, r.mode, r.host_ex != nullptr, r.parties) && !r.cw_out;
}

size_t dpf_frontier_nodes(int n, int w, bool deep) {
  // the widest stored level: the last parent level (2^(w(levels-1)) nodes
  // per tree), or level levels - 4 when the leaf level expands the rest
  return deep ? (size_t)1 << (n - 4) : (size_t)1 << (w * (n / w - 1));
}

void dpf_reserve(DpfScratch& s, int n, int w, size_t ntrees, bool host_exchange, bool deep) {
  const int nk = (1 << w) - 1, levels = n / w;
  // each stored node with its nk hashes
  const size_t maxS = dpf_frontier_nodes(n, w, deep);
  const size_t need = ntrees * maxS * (size_t)(1 + nk) * sizeof(Blk);
  s.fh[0].ensure(need, "hipMalloc dpf frontier");
  s.fh[1].ensure(need, "hipMalloc dpf frontier");
  const size_t cw_bytes = (size_t)levels * ntrees * (size_t)nk * sizeof(Blk);
  s.cw.ensure(cw_bytes, "hipMalloc dpf cw");
  if (host_exchange) s.peer.ensure(cw_bytes, "hipMalloc dpf peer");
}

void dpf_expand(hipStream_t st, const DpfRun& r, DpfScratch& s) {
  require(r.w >= 1 && r.w <= 4, "dpf: w must be in [1,4]");
  require(r.n >= r.w && r.n % r.w == 0, "dpf: w must divide n");
  require(r.n <= 30, "dpf: n too large");
  require(r.ntrees > 0, "dpf: no instances");
  require(r.prg == kPrgAes || r.prg == kPrgChaCha8, "dpf: unknown PRG");
  require(r.prg != kPrgAes || (r.aes_t0 && r.aes_rks), "dpf: AES PRG requires round keys");
  require(r.host_ex || r.ntrees % 2 == 0, "dpf: in-device exchange pairs trees T and T+ntrees/2");
  require(!r.roots1 || (r.dltsft1 && r.ntrees % 2 == 0), "dpf: second half needs roots1 and dltsft1");
  require(r.mode != kLeafStore || r.leaves, "dpf: leaves output missing");
  require(r.mode != kLeafConvert || (r.ct && r.partials), "dpf: convert outputs missing");
  require(r.mode != kLeafGather ||
              (r.g_acc && r.g_mask && r.partials && r.n / r.w >= 2 &&
               ((size_t)1 << (r.n - 1)) >= (size_t)kGatherTile && r.t >= 1 && r.t <= 64 &&
               r.B >= 1 && r.B <= r.t * r.t && 4 * r.t <= 64 * r.mask_words &&
               r.sets >= 1 && r.sets <= 65535 && (r.parties == 1 || r.parties == 2) &&
               r.ntrees == r.parties * r.sets * r.B && r.N_log >= r.n - 1 && r.N_log <= 30 &&
               (r.parties == 1 || !r.host_ex)),
          "dpf: fused gather needs bin >= 512, t <= 64, two levels and [party][sets][B] trees");
#define PCGSQ_DISPATCH(W)                                                        \
  if (r.w == W) {                                                                \
    if (r.prg == kPrgChaCha8) expand_t<W, kPrgChaCha8>(st, r, s);                \
    else expand_t<W, kPrgAes>(st, r, s);                                         \
    return;                                                                      \
  }
  PCGSQ_DISPATCH(1)
  PCGSQ_DISPATCH(2)
  PCGSQ_DISPATCH(3)
  PCGSQ_DISPATCH(4)
#undef PCGSQ_DISPATCH
}

// ---- host AES (FIPS-197) ------------------------------------------------------

namespace {
uint8_t gmul(uint8_t a, uint8_t b) {
  uint8_t r = 0;
  while (b) {
    if (b & 1) r ^= a;
    a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
    b >>= 1;
  }
  return r;
}
// S-box from the GF(2^8) inverse and the affine map (FIPS-197 Sec. 5.1.1).
struct SBox {
  uint8_t s[256];
  SBox() {
    for (int x = 0; x < 256; ++x) {
      uint8_t inv = 0;
      for (int y = 1; y < 256 && x; ++y)
        if (gmul((uint8_t)x, (uint8_t)y) == 1) { inv = (uint8_t)y; break; }
      uint8_t v = inv;
      for (int r = 1; r <= 4; ++r) v ^= (uint8_t)((inv << r) | (inv >> (8 - r)));
      s[x] = v ^ 0x63;
    }
  }
};
const SBox& sbox() {
  static const SBox sb;
  return sb;
}
}  // namespace

void aes128_expand_key_host(const uint8_t key[16], uint8_t rk[176]) {
  const uint8_t* S = sbox().s;
  uint8_t rcon = 1;
  memcpy(rk, key, 16);
  for (int i = 4; i < 44; ++i) {
    uint8_t t[4];
    memcpy(t, rk + 4 * (i - 1), 4);
    if (i % 4 == 0) {
      const uint8_t t0 = t[0];
      t[0] = (uint8_t)(S[t[1]] ^ rcon);
      t[1] = S[t[2]];
      t[2] = S[t[3]];
      t[3] = S[t0];
      rcon = gmul(rcon, 2);
    }
    for (int k = 0; k < 4; ++k) rk[4 * i + k] = rk[4 * (i - 4) + k] ^ t[k];
  }
}

void aes_ttable_host(uint32_t t0[256]) {
  const uint8_t* S = sbox().s;
  for (int x = 0; x < 256; ++x) {
    const uint8_t s = S[x];
    t0[x] = (uint32_t)gmul(s, 2) | ((uint32_t)s << 8) | ((uint32_t)s << 16) |
            ((uint32_t)gmul(s, 3) << 24);
  }
}

void aes128_encrypt_host(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]) {
  const uint8_t* S = sbox().s;
  uint8_t s[16], t[16];
  for (int i = 0; i < 16; ++i) s[i] = in[i] ^ rk[i];
  for (int r = 1; r <= 10; ++r) {
    for (int c = 0; c < 4; ++c)
      for (int j = 0; j < 4; ++j) t[4 * c + j] = S[s[4 * ((c + j) & 3) + j]];
    if (r < 10) {
      for (int c = 0; c < 4; ++c) {
        const uint8_t a0 = t[4 * c], a1 = t[4 * c + 1], a2 = t[4 * c + 2], a3 = t[4 * c + 3];
        s[4 * c + 0] = gmul(a0, 2) ^ gmul(a1, 3) ^ a2 ^ a3;
        s[4 * c + 1] = a0 ^ gmul(a1, 2) ^ gmul(a2, 3) ^ a3;
        s[4 * c + 2] = a0 ^ a1 ^ gmul(a2, 2) ^ gmul(a3, 3);
        s[4 * c + 3] = gmul(a0, 3) ^ a1 ^ a2 ^ gmul(a3, 2);
      }
    } else {
      memcpy(s, t, 16);
    }
    for (int i = 0; i < 16; ++i) s[i] ^= rk[16 * r + i];
  }
  memcpy(out, s, 16);
}

void aes_tree_round_keys_host(const uint8_t master[16], int nk, uint8_t* rks) {
  uint8_t mrk[176];
  aes128_expand_key_host(master, mrk);
  for (int k = 0; k < nk; ++k) {
    uint8_t tweak[16] = {0}, c[16], sub[16];
    for (int b = 0; b < 8; ++b) tweak[b] = (uint8_t)((uint64_t)k >> (8 * b));
    aes128_encrypt_host(mrk, tweak, c);
    for (int b = 0; b < 16; ++b) sub[b] = c[b] ^ tweak[b];
    aes128_expand_key_host(sub, rks + (size_t)k * 176);
  }
}

}  // namespace pcgsq
