// Batched half-tree DPF keys of uneven domains: dealer Gen plus sync-free
// full-domain Eval with the fused leaf convert, in one persistent launch.
//
// Gen (one thread per key pair, w = 1): roots root_s = Delta_s ^ W as in the
// fused protocol's setup (half_tree_dpf/dpf.cpp:338), and per level the CW
// that protocol's exchange reconstructs, which is the path-only value
// H(s0) ^ H(s1) ^ [bit_i(alpha) == 0] * Delta (dpf.cpp:358-385: off-path
// nodes are equal in both trees and cancel), with the path advanced by the
// children of dpf.cpp:391-397.  The output word cw_out makes
// y_s(j) = (-1)^s (C_s(j) + tau_s(j) cw_out), C = lo64(H'(leaf)) mod P,
// tau = lsb(leaf) -- the conversion of pcg_ole_impl.h:432-446 -- sum to
// beta * [j == alpha].
//
// Eval: with every CW known up front no level needs a reduction over the
// frontier, so nothing is kept in HBM but the outputs.  Work items are
// subtrees of 2^h leaves (h = min(n, 14)), one per wavefront: each lane walks
// from the key's root to its own subtree of 2^(h-6) leaves and expands that
// depth first in registers (a stack of pending right children), converts
// every leaf and writes 16 consecutive outputs at a time.  Lanes of a wave
// run the same tree shape, so the walk never diverges.  Wavefronts take items
// from one atomic counter, largest keys first; items are independent, so the
// queue drains without waits and the launch can be captured in a HIP graph.
#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "context.hpp"
#include "dpf.hpp"
#include "pcgsq.h"
#include "prg.hpp"

namespace pcgsq {
namespace {

constexpr int kThreads = 256;
constexpr int kItemLeavesLog = 14;   // leaves per work item (at most)
constexpr int kLaneLog = 6;          // 64 lanes split an item
constexpr int kChunk = 16;           // outputs a lane writes at once (128 bytes)
constexpr int kMinLog = kLaneLog + 4;
constexpr int kMaxLog = 30;

struct KeyDev {
  Blk root[2];
  uint64_t cw_out;
  uint64_t y_off;   // words into each party's output array
  uint32_t cw_off;  // first CW (level 0) of the key
  int n;
};

struct Item {
  uint32_t key;
  uint16_t party;
  uint16_t h;       // the subtree has 2^h leaves
  uint32_t sub;     // subtree index at depth n - h
  uint32_t pad;
};

template <int PRG>
__device__ __forceinline__ Blk hash1(Blk x, const AesLds* lds) {
  Blk h[1];
  tree_hash<PRG, 1>(x, h, lds);
  return h[0];
}

template <int PRG>
__global__ __launch_bounds__(kThreads) void dpf_gen_kernel(
    const uint64_t* __restrict__ alpha, const uint64_t* __restrict__ beta,
    const Blk* __restrict__ delta, const Blk* __restrict__ delta0, Blk W,
    KeyDev* __restrict__ keys, Blk* __restrict__ cw, int count, const uint32_t* aes_t0,
    const uint8_t* aes_rks) {
  __shared__ typename std::conditional<PRG == kPrgAes, AesLds, int>::type aes_lds_storage;
  const AesLds* lds = nullptr;
  if constexpr (PRG == kPrgAes) {
    aes_stage(&aes_lds_storage, aes_t0, aes_rks, 1);
    lds = &aes_lds_storage;
  }
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= count) return;
  const int n = keys[k].n;
  const Blk d = delta[k];
  const uint64_t a = alpha[k];
  Blk s0 = bxor(delta0[k], W);
  Blk s1 = bxor(bxor(delta0[k], d), W);
  keys[k].root[0] = s0;
  keys[k].root[1] = s1;
  Blk* c = cw + keys[k].cw_off;
  for (int i = 0; i < n; ++i) {
    const Blk h0 = hash1<PRG>(s0, lds), h1 = hash1<PRG>(s1, lds);
    const bool bit = (a >> (n - 1 - i)) & 1;
    Blk C = bxor(h0, h1);
    if (!bit) C = bxor(C, d);
    c[i] = C;
    Blk n0 = bit ? bxor(h0, s0) : h0;
    Blk n1 = bit ? bxor(h1, s1) : h1;
    if (s0.lo & 1) n0 = bxor(n0, C);
    if (s1.lo & 1) n1 = bxor(n1, C);
    s0 = n0;
    s1 = n1;
  }
  const uint64_t C0 = freduce64(out_hash_lo64(s0)), C1 = freduce64(out_hash_lo64(s1));
  const uint64_t v = fsub(fadd(beta[k], C1), C0);
  keys[k].cw_out = (s0.lo & 1) ? v : fneg(v);
}

template <int PRG>
__global__ __launch_bounds__(kThreads) void dpf_eval_kernel(
    const KeyDev* __restrict__ keys, const Blk* __restrict__ cw, const Item* __restrict__ items,
    uint32_t nitems, uint32_t* __restrict__ next, uint64_t* __restrict__ y, size_t total,
    const uint32_t* aes_t0, const uint8_t* aes_rks) {
  __shared__ typename std::conditional<PRG == kPrgAes, AesLds, int>::type aes_lds_storage;
  const AesLds* lds = nullptr;
  if constexpr (PRG == kPrgAes) {
    aes_stage(&aes_lds_storage, aes_t0, aes_rks, 1);
    lds = &aes_lds_storage;
  }
  const uint32_t lane = threadIdx.x & 63;
  for (;;) {
    uint32_t idx = 0;
    if (lane == 0) idx = atomicAdd(next, 1u);
    idx = __builtin_amdgcn_readfirstlane(idx);
    if (idx >= nitems) break;
    const Item it = items[idx];
    const KeyDev& K = keys[it.key];
    const int n = K.n, h = it.h, party = it.party;
    const Blk* C = cw + K.cw_off;
    // walk to this lane's subtree root at depth n - h + 6
    const uint32_t path = (it.sub << kLaneLog) | lane;
    const int plen = n - h + kLaneLog;
    Blk node = K.root[party];
    for (int dd = 0; dd < plen; ++dd) {
      const Blk hh = hash1<PRG>(node, lds);
      Blk c0 = hh;
      if (node.lo & 1) c0 = bxor(c0, C[dd]);
      node = ((path >> (plen - 1 - dd)) & 1) ? bxor(c0, node) : c0;
    }
    // depth-first over the lane's 2^D leaves
    const int D = h - kLaneLog;
    uint64_t* out = y + (size_t)party * total + K.y_off + ((size_t)path << D);
    const uint64_t cwo = K.cw_out;
    Blk stack[kItemLeavesLog - kLaneLog];
    uint64_t buf[kChunk];
    int lvl = 0;
    const uint32_t nleaves = 1u << D;
    for (uint32_t leaf = 0; leaf < nleaves; ++leaf) {
      while (lvl < D) {
        const Blk hh = hash1<PRG>(node, lds);
        Blk c0 = hh;
        if (node.lo & 1) c0 = bxor(c0, C[plen + lvl]);
        stack[lvl] = bxor(c0, node);  // right child, visited later
        node = c0;
        ++lvl;
      }
      uint64_t v = freduce64(out_hash_lo64(node));
      if (node.lo & 1) v = fadd(v, cwo);
      buf[leaf & (kChunk - 1)] = party ? fneg(v) : v;
      if ((leaf & (kChunk - 1)) == kChunk - 1) {
        uint64_t* dst = out + (leaf - (kChunk - 1));
#pragma unroll
        for (int q = 0; q < kChunk; q += 2)
          *reinterpret_cast<ulonglong2*>(dst + q) = ulonglong2{buf[q], buf[q + 1]};
      }
      if (leaf + 1 < nleaves) {
        const int k = __builtin_ctz(~leaf);  // the deepest decision that flips
        lvl = D - k;
        node = stack[D - 1 - k];
      }
    }
  }
}

__global__ void dpf_batch_reset(uint32_t* ctr) {
  if (threadIdx.x == 0) ctr[0] = 0;
}

}  // namespace
}  // namespace pcgsq

using namespace pcgsq;

struct pcgsq_dpf_batch {
  pcgsq_ctx* ctx = nullptr;
  int count = 0, prg = kPrgChaCha8;
  std::vector<int> logn;
  std::vector<uint64_t> off;
  size_t total = 0, cw_total = 0;
  DevBuf keys, cw, items, ctr, y, alpha, beta, delta, delta0, aes_t0, aes_rks;
  uint32_t nitems = 0;
  int grid = 0;
  double tree_blocks = 0, out_blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  void enqueue(hipStream_t st) {
    hipLaunchKernelGGL(dpf_batch_reset, dim3(1), dim3(64), 0, st, ctr.as<uint32_t>());
    if (prg == kPrgAes)
      hipLaunchKernelGGL(dpf_eval_kernel<kPrgAes>, dim3((unsigned)grid), dim3(kThreads), 0, st,
                         keys.as<KeyDev>(), cw.as<Blk>(), items.as<Item>(), nitems,
                         ctr.as<uint32_t>(), y.as<uint64_t>(), total, aes_t0.as<uint32_t>(),
                         aes_rks.as<uint8_t>());
    else
      hipLaunchKernelGGL(dpf_eval_kernel<kPrgChaCha8>, dim3((unsigned)grid), dim3(kThreads), 0,
                         st, keys.as<KeyDev>(), cw.as<Blk>(), items.as<Item>(), nitems,
                         ctr.as<uint32_t>(), y.as<uint64_t>(), total, nullptr, nullptr);
    hip_check(hipGetLastError(), "launch dpf_eval_kernel");
  }

  ~pcgsq_dpf_batch() { drop_graph(); }
};

namespace {

// Dealer draws of the batched keys (the seeded generator of the expansion,
// domains 10-13): alpha, beta, Delta (lsb 1), Delta_0 share (lsb 1).
void deal_points(uint64_t seed, const std::vector<int>& logn, std::vector<uint64_t>& alpha,
                 std::vector<uint64_t>& beta, std::vector<Blk>& delta,
                 std::vector<Blk>& delta0) {
  const size_t K = logn.size();
  alpha.resize(K);
  beta.resize(K);
  delta.resize(K);
  delta0.resize(K);
  for (size_t k = 0; k < K; ++k) {
    alpha[k] = dealer_u64(seed, 10, k, 0) & ((1ULL << logn[k]) - 1);
    beta[k] = dealer_u64(seed, 11, k, 0) % (kPrime - 1) + 1;
    delta[k] = Blk{dealer_u64(seed, 12, k, 0) | 1ULL, dealer_u64(seed, 12, k, 1)};
    delta0[k] = Blk{dealer_u64(seed, 13, k, 0) | 1ULL, dealer_u64(seed, 13, k, 1)};
  }
}

constexpr Blk kPublicW{0x0123456789abcdefULL, 0xdeadbeefcafe0001ULL};

void gen_keys(pcgsq_dpf_batch* b, const uint64_t* alpha, const uint64_t* beta, const Blk* delta,
              const Blk* delta0) {
  const size_t K = (size_t)b->count;
  hipStream_t st = b->ctx->stream;
  for (size_t k = 0; k < K; ++k) {
    require(alpha[k] < (1ULL << b->logn[k]), "dpf batch: alpha outside the domain");
    require((delta[k].lo & 1) == 1, "dpf batch: Delta needs lsb 1");
  }
  hip_check(hipMemcpyAsync(b->alpha.ptr, alpha, K * 8, hipMemcpyHostToDevice, st), "dpf gen");
  hip_check(hipMemcpyAsync(b->beta.ptr, beta, K * 8, hipMemcpyHostToDevice, st), "dpf gen");
  hip_check(hipMemcpyAsync(b->delta.ptr, delta, K * 16, hipMemcpyHostToDevice, st), "dpf gen");
  hip_check(hipMemcpyAsync(b->delta0.ptr, delta0, K * 16, hipMemcpyHostToDevice, st), "dpf gen");
  const dim3 grid((unsigned)((K + kThreads - 1) / kThreads));
  if (b->prg == kPrgAes)
    hipLaunchKernelGGL(dpf_gen_kernel<kPrgAes>, grid, dim3(kThreads), 0, st,
                       b->alpha.as<uint64_t>(), b->beta.as<uint64_t>(), b->delta.as<Blk>(),
                       b->delta0.as<Blk>(), kPublicW, b->keys.as<KeyDev>(), b->cw.as<Blk>(),
                       b->count, b->aes_t0.as<uint32_t>(), b->aes_rks.as<uint8_t>());
  else
    hipLaunchKernelGGL(dpf_gen_kernel<kPrgChaCha8>, grid, dim3(kThreads), 0, st,
                       b->alpha.as<uint64_t>(), b->beta.as<uint64_t>(), b->delta.as<Blk>(),
                       b->delta0.as<Blk>(), kPublicW, b->keys.as<KeyDev>(), b->cw.as<Blk>(),
                       b->count, nullptr, nullptr);
  hip_check(hipGetLastError(), "launch dpf_gen_kernel");
  hip_check(hipStreamSynchronize(st), "dpf gen");
}

}  // namespace

extern "C" {

int pcgsq_dpf_batch_create(pcgsq_ctx* ctx, const int* logn, int count, int prg,
                           pcgsq_dpf_batch** out) {
  return guarded([&] {
    require(ctx && logn && out, "dpf_batch_create: null argument");
    require(count >= 1 && count <= (1 << 20), "dpf_batch_create: count must be 1..2^20");
    require(prg == kPrgAes || prg == kPrgChaCha8, "dpf_batch_create: unknown PRG");
    CtxGuard g(ctx);
    std::unique_ptr<pcgsq_dpf_batch> b(new pcgsq_dpf_batch);
    b->ctx = ctx;
    b->count = count;
    b->prg = prg;
    std::vector<KeyDev> kd((size_t)count);
    for (int k = 0; k < count; ++k) {
      require(logn[k] >= kMinLog && logn[k] <= kMaxLog,
              "dpf_batch_create: domains must be 2^10..2^30", kErrUnsupported);
      b->logn.push_back(logn[k]);
      b->off.push_back(b->total);
      kd[(size_t)k] = KeyDev{};
      kd[(size_t)k].n = logn[k];
      kd[(size_t)k].y_off = b->total;
      kd[(size_t)k].cw_off = (uint32_t)b->cw_total;
      b->total += (size_t)1 << logn[k];
      b->cw_total += (size_t)logn[k];
      b->tree_blocks += 2.0 * (double)(((size_t)1 << logn[k]) - 1);
      b->out_blocks += 2.0 * (double)((size_t)1 << logn[k]);
    }
    // work items: both parties, subtrees of 2^h leaves, longest keys first
    std::vector<int> order((size_t)count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int u, int v) { return logn[u] > logn[v]; });
    std::vector<Item> it;
    for (int k : order) {
      const int h = std::min(logn[k], kItemLeavesLog);
      const uint32_t subs = 1u << (logn[k] - h);
      for (uint32_t s = 0; s < subs; ++s)
        for (uint16_t p = 0; p < 2; ++p) it.push_back(Item{(uint32_t)k, p, (uint16_t)h, s, 0});
    }
    b->nitems = (uint32_t)it.size();
    b->keys.ensure(kd.size() * sizeof(KeyDev), "hipMalloc dpf batch keys");
    b->cw.ensure(std::max<size_t>(1, b->cw_total) * sizeof(Blk), "hipMalloc dpf batch cw");
    b->items.ensure(it.size() * sizeof(Item), "hipMalloc dpf batch items");
    b->ctr.ensure(256, "hipMalloc dpf batch counter");
    b->y.ensure(2 * b->total * sizeof(uint64_t), "hipMalloc dpf batch outputs");
    b->alpha.ensure((size_t)count * 8, "hipMalloc dpf batch points");
    b->beta.ensure((size_t)count * 8, "hipMalloc dpf batch points");
    b->delta.ensure((size_t)count * 16, "hipMalloc dpf batch points");
    b->delta0.ensure((size_t)count * 16, "hipMalloc dpf batch points");
    hip_check(hipMemcpy(b->keys.ptr, kd.data(), kd.size() * sizeof(KeyDev), hipMemcpyHostToDevice),
              "upload dpf batch keys");
    hip_check(hipMemcpy(b->items.ptr, it.data(), it.size() * sizeof(Item), hipMemcpyHostToDevice),
              "upload dpf batch items");
    if (prg == kPrgAes) {
      // the FIPS-197 master key of the AES tree PRG (pcg_ole_impl.h:89-92)
      static const uint8_t master[16] = {0x2b, 0x7e, 0x15, 0x16, 0x28, 0xae, 0xd2, 0xa6,
                                         0xab, 0xf7, 0x15, 0x88, 0x09, 0xcf, 0x4f, 0x3c};
      uint8_t rks[176];
      aes_tree_round_keys_host(master, 1, rks);
      uint32_t t0[256];
      aes_ttable_host(t0);
      b->aes_t0.ensure(sizeof(t0), "hipMalloc aes table");
      b->aes_rks.ensure(sizeof(rks), "hipMalloc aes keys");
      hip_check(hipMemcpy(b->aes_t0.ptr, t0, sizeof(t0), hipMemcpyHostToDevice), "upload aes");
      hip_check(hipMemcpy(b->aes_rks.ptr, rks, sizeof(rks), hipMemcpyHostToDevice), "upload aes");
    }
    int per_cu = 0, cus = 0;
    if (prg == kPrgAes)
      hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dpf_eval_kernel<kPrgAes>,
                                                             kThreads, 0),
                "dpf batch occupancy");
    else
      hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, dpf_eval_kernel<kPrgChaCha8>, kThreads, 0),
                "dpf batch occupancy");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
              "dpf batch CU count");
    b->grid = std::max(1, per_cu * cus);
    *out = b.release();
  });
}

int pcgsq_dpf_batch_destroy(pcgsq_dpf_batch* b) {
  if (!b) return PCGSQ_OK;
  return guarded([&] {
    CtxGuard g(b->ctx);
    hip_check(hipStreamSynchronize(b->ctx->stream), "dpf batch sync");
    delete b;
  });
}

int pcgsq_dpf_batch_deal(pcgsq_dpf_batch* b, uint64_t seed) {
  return guarded([&] {
    require(b, "dpf_batch_deal: null batch");
    CtxGuard g(b->ctx);
    std::vector<uint64_t> alpha, beta;
    std::vector<Blk> delta, delta0;
    deal_points(seed, b->logn, alpha, beta, delta, delta0);
    gen_keys(b, alpha.data(), beta.data(), delta.data(), delta0.data());
  });
}

int pcgsq_dpf_batch_gen(pcgsq_dpf_batch* b, const uint64_t* alpha, const uint64_t* beta,
                        const pcgsq_blk* delta, const pcgsq_blk* delta0) {
  return guarded([&] {
    require(b && alpha && beta && delta && delta0, "dpf_batch_gen: null argument");
    CtxGuard g(b->ctx);
    gen_keys(b, alpha, beta, reinterpret_cast<const Blk*>(delta),
             reinterpret_cast<const Blk*>(delta0));
  });
}

int pcgsq_dpf_batch_points(pcgsq_dpf_batch* b, uint64_t* alpha, uint64_t* beta) {
  return guarded([&] {
    require(b, "dpf_batch_points: null batch");
    CtxGuard g(b->ctx);
    const size_t K = (size_t)b->count;
    if (alpha)
      hip_check(hipMemcpy(alpha, b->alpha.ptr, K * 8, hipMemcpyDeviceToHost), "dpf points");
    if (beta) hip_check(hipMemcpy(beta, b->beta.ptr, K * 8, hipMemcpyDeviceToHost), "dpf points");
  });
}

int pcgsq_dpf_batch_keys(pcgsq_dpf_batch* b, pcgsq_blk* roots, pcgsq_blk* cw, uint64_t* cw_out) {
  return guarded([&] {
    require(b, "dpf_batch_keys: null batch");
    CtxGuard g(b->ctx);
    std::vector<KeyDev> kd((size_t)b->count);
    hip_check(hipMemcpy(kd.data(), b->keys.ptr, kd.size() * sizeof(KeyDev),
                        hipMemcpyDeviceToHost), "dpf keys");
    for (size_t k = 0; k < kd.size(); ++k) {
      if (roots) {
        roots[k] = pcgsq_blk{kd[k].root[0].lo, kd[k].root[0].hi};
        roots[kd.size() + k] = pcgsq_blk{kd[k].root[1].lo, kd[k].root[1].hi};
      }
      if (cw_out) cw_out[k] = kd[k].cw_out;
    }
    if (cw)
      hip_check(hipMemcpy(cw, b->cw.ptr, b->cw_total * sizeof(Blk), hipMemcpyDeviceToHost),
                "dpf keys");
  });
}

int pcgsq_dpf_batch_run(pcgsq_dpf_batch* b, int use_graph) {
  return guarded([&] {
    require(b, "dpf_batch_run: null batch");
    CtxGuard g(b->ctx);
    hipStream_t st = b->ctx->stream;
    if (!use_graph) {
      ProfZone pz(&b->ctx->prof, st, "dpf.batch");
      b->enqueue(st);
      return;
    }
    if (!b->exec) {
      hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "dpf batch capture");
      b->enqueue(st);
      hip_check(hipStreamEndCapture(st, &b->graph), "dpf batch capture");
      hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0),
                "dpf batch graph instantiate");
    }
    ProfZone pz(&b->ctx->prof, st, "dpf.batch");
    hip_check(hipGraphLaunch(b->exec, st), "dpf batch graph launch");
  });
}

int pcgsq_dpf_batch_read(pcgsq_dpf_batch* b, uint64_t* y0, uint64_t* y1) {
  return guarded([&] {
    require(b, "dpf_batch_read: null batch");
    CtxGuard g(b->ctx);
    hip_check(hipStreamSynchronize(b->ctx->stream), "dpf batch sync");
    const size_t bytes = b->total * sizeof(uint64_t);
    if (y0) hip_check(hipMemcpy(y0, b->y.ptr, bytes, hipMemcpyDeviceToHost), "dpf batch read");
    if (y1)
      hip_check(hipMemcpy(y1, b->y.as<uint64_t>() + b->total, bytes, hipMemcpyDeviceToHost),
                "dpf batch read");
  });
}

int pcgsq_dpf_batch_dev(pcgsq_dpf_batch* b, uint64_t** y, size_t* total_words) {
  return guarded([&] {
    require(b, "dpf_batch_dev: null batch");
    if (y) *y = b->y.as<uint64_t>();
    if (total_words) *total_words = b->total;
  });
}

int pcgsq_dpf_batch_info(pcgsq_dpf_batch* b, int* nitems, int* grid, double* tree_blocks,
                         double* out_blocks) {
  return guarded([&] {
    require(b, "dpf_batch_info: null batch");
    if (nitems) *nitems = (int)b->nitems;
    if (grid) *grid = b->grid;
    if (tree_blocks) *tree_blocks = b->tree_blocks;
    if (out_blocks) *out_blocks = b->out_blocks;
  });
}

}  // extern "C"
