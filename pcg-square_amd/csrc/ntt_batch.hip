// Batched NTT of uneven lengths in one persistent launch.
//
// A batch holds `count` independent transforms of lengths 2^10..2^22, each in
// its own slice of the batch buffers.  The host turns every pass of every
// transform into tiles (the tiles the per-length launches use, shaped for
// 1024 threads) and writes them into work lists.  One persistent kernel
// drains the lists: a workgroup takes the next item with one atomic, waits
// until every tile of the pass the item depends on has finished (per-pass
// completion counters, a bounded spin), runs the tile and counts it done.
// Items are taken in list order and only ever wait on earlier items of the
// same list, so every list drains.  Transforms of different lengths share
// the GPU in one launch with no host round trips, and the launch (counter
// reset + persistent kernel) can be captured in a HIP graph and replayed.
//
// Each transform belongs to one XCD: the host spreads transforms over the
// XCDs (largest first onto the least loaded), keeps one work list per XCD,
// and a workgroup serves the list of the XCD it runs on (hardware register
// XCC_ID).  Every producer/consumer pair of a transform therefore shares one
// L2, so a finished tile needs only its stores acknowledged (s_waitcnt) before
// it is counted, never an L2 write-back or invalidate; and every buffer a
// pass reads was written earlier in the same launch and never read before it
// (separate scratch for the forward and the inverse direction), so no CU can
// hold a stale L1 line of it.
//
// Per transform: ops & 1 = forward (a -> f), ops & 2 = inverse (of f when
// both are asked, else of a) -> r.  Results equal the per-length launches
// bit for bit (same passes, same tables).
//
// The batch can instead run grouped by length (PCGSQ_NTT_BATCH_GROUPED): per
// length, the per-length pass kernels in launches of their own over all
// transforms of that length, found through an offset table (no regrouping
// copy), shaped as the context's per-length launches are (the autotuner's
// split, grouping and tile widths).  The autotuner picks the faster mode.
#include <algorithm>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "context.hpp"
#include "ntt_pass.hpp"
#include "ntt_plan.hpp"
#include "pcgsq.h"

namespace pcgsq {
namespace {

// 1024 threads: a 2^10 column pass gets 16 columns per tile, whole 128-byte
// lines of each row (with 8 columns the two halves of a line were read by
// different tiles, not always from L2: 1.22x the algorithmic bytes against
// 1.05x now), and the per-item bookkeeping (claim, dependency check,
// descriptor loads, store drain) is spread over 1024 threads' work; one such
// workgroup fills a CU's LDS (16 waves, as two 512-thread ones did).
constexpr int kBatchThreads = 1024;
constexpr int kMaxQueues = 16;      // work lists (one per XCD)
constexpr int kCtrLine = 32;        // counters of different lists sit 128 bytes apart
constexpr int kCtrErr = kMaxQueues * kCtrLine;
constexpr int kCtrDone = kCtrErr + kCtrLine;
constexpr int kHwRegXccId = (3 << 11) | 20;  // s_getreg: HW_REG_XCC_ID bits [3:0]

struct BatchPhase {
  NttPassDesc d;
  int kind;       // PassKind, -1 = no such phase
  int nlog;
  int tiles;
  int done_ctr;   // counter this phase's tiles increment
  int wait_ctr;   // counter of the phase this one waits for, -1 = none
  int wait_need;  // tiles of that phase
};

struct BatchItem {
  uint32_t xp;    // transform * 4 + phase
  uint32_t tile;
};

struct BatchQueues {
  uint32_t beg[kMaxQueues], end[kMaxQueues];
  uint32_t n;     // lists in use (the device's XCD count)
};

// Register grouping per pass length: the largest group and the number of
// groups of the shapes below.  The kernel takes the VGPRs of its largest
// shape, so every shape stays under 128 (four waves per SIMD): 2^8 runs as
// 3+3+2 stages (two 4-stage groups need 132 VGPRs), and the first passes do
// not prefetch their four-step twiddles.
__host__ __device__ constexpr int batch_rmax(int nlog) {
  return (nlog == 6 || nlog == 8 || nlog == 9) ? 3 : 4;
}
__host__ __device__ constexpr int batch_groups(int nlog) { return nlog >= 8 ? 3 : 2; }

// One-pass transforms up to 2^12; four-step passes of 2^6 (rows only) to 2^11.
#define PCGSQ_BATCH_SHAPES(X)                                                                 \
  X(kFwdWhole, 10, 4, 3, 3) X(kFwdWhole, 11, 4, 4, 3) X(kFwdWhole, 12, 4, 4, 4)              \
  X(kInvWhole, 10, 4, 3, 3) X(kInvWhole, 11, 4, 4, 3) X(kInvWhole, 12, 4, 4, 4)              \
  X(kFwdRows, 6, 3, 3, 0) X(kInvRows, 6, 3, 3, 0)                                            \
  X(kFwdCols, 7, 4, 3, 0) X(kFwdCols, 8, 3, 3, 2) X(kFwdCols, 9, 3, 3, 3)                    \
  X(kFwdCols, 10, 4, 3, 3) X(kFwdCols, 11, 4, 4, 3)                                          \
  X(kFwdRows, 7, 4, 3, 0) X(kFwdRows, 8, 3, 3, 2) X(kFwdRows, 9, 3, 3, 3)                    \
  X(kFwdRows, 10, 4, 3, 3) X(kFwdRows, 11, 4, 4, 3)                                          \
  X(kInvRows, 7, 4, 3, 0) X(kInvRows, 8, 3, 3, 2) X(kInvRows, 9, 3, 3, 3)                    \
  X(kInvRows, 10, 4, 3, 3) X(kInvRows, 11, 4, 4, 3)                                          \
  X(kInvCols, 7, 4, 3, 0) X(kInvCols, 8, 3, 3, 2) X(kInvCols, 9, 3, 3, 3)                    \
  X(kInvCols, 10, 4, 3, 3) X(kInvCols, 11, 4, 4, 3)

// Each shape is a separate (not inlined) function, so the persistent kernel
// needs only the registers of its largest shape rather than of all of them.
// The function reaches LDS through the kernel's dynamic LDS symbol (not a
// generic pointer, which would turn every LDS access into a flat one) and
// reads its descriptor through a wave-uniform pointer (scalar loads).
extern __shared__ uint64_t batch_lds[];

template <int K, int A, int B, int C>
__device__ __noinline__ void tile_fn(const NttPassDesc* d, uint32_t tile) {
  const uint64_t bits = (uint64_t)(uintptr_t)d;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)bits);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(bits >> 32));
  const __attribute__((address_space(4))) uint32_t* src =
      (const __attribute__((address_space(4))) uint32_t*)(uintptr_t)(((uint64_t)hi << 32) | lo);
  NttPassDesc desc;
  static_assert(sizeof(NttPassDesc) % 4 == 0, "descriptor is copied as words");
  uint32_t* dst = reinterpret_cast<uint32_t*>(&desc);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(NttPassDesc) / 4); ++k) dst[k] = src[k];
  nttpass::pass_tile<K, A, B, C, false>(desc, batch_lds, 0, __builtin_amdgcn_readfirstlane(tile));
}

// the host sizes tiles by batch_rmax / batch_groups: every shape must agree
#define PCGSQ_SHAPE_CHECK(K, N, A, B, C)                                                        \
  static_assert(nttpass::Part<A, B, C>::RMAX == batch_rmax(N) &&                               \
                    nttpass::Part<A, B, C>::G == batch_groups(N) && A + B + C == N,            \
                "batch shape disagrees with batch_rmax / batch_groups");
PCGSQ_BATCH_SHAPES(PCGSQ_SHAPE_CHECK)
#undef PCGSQ_SHAPE_CHECK

__device__ __forceinline__ void run_tile(int kind, int nlog, const NttPassDesc* d,
                                         uint32_t tile) {
  switch (kind * 32 + nlog) {
#define PCGSQ_CASE(K, N, A, B, C)                        \
  case K * 32 + N:                                      \
    tile_fn<K, A, B, C>(d, tile);                       \
    return;
    PCGSQ_BATCH_SHAPES(PCGSQ_CASE)
#undef PCGSQ_CASE
    default:
      return;
  }
}

__global__ void batch_reset_kernel(uint32_t* ctr, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    ctr[i] = 0;
}

// Thread 0 runs the only lane-divergent code (count the finished item, claim
// the item after the next one, wait for the next one's dependency); each such
// region sits between two block barriers, so no barrier is ever reached with
// part of a wave masked off.  Finishing item k is counted when item k+1
// starts.  A workgroup claims one item ahead, so the claim's round trip
// overlaps a tile; that cannot deadlock: a workgroup only ever waits on items
// claimed before the one it runs, and it claimed its pending item after that.
// Counters are touched only by atomics, and only by workgroups of one XCD.
__global__ __launch_bounds__(kBatchThreads) void ntt_batch_kernel(
    const BatchPhase* __restrict__ ph, const BatchItem* __restrict__ items, BatchQueues q,
    uint32_t* __restrict__ ctr, uint32_t spin_limit) {
  __shared__ uint32_t s_item;
  uint32_t xcc = __builtin_amdgcn_s_getreg(kHwRegXccId) & 15u;
  while (xcc >= q.n) xcc -= q.n;  // (no integer division: it would go through float)
  const uint32_t qb = q.beg[xcc], qe = q.end[xcc];
  uint32_t* next = ctr + xcc * kCtrLine;
  int finished = -1;      // counter of the item this workgroup just ran
  uint32_t pending = 0;   // (thread 0) the claimed item to run next
  if (threadIdx.x == 0) pending = qb + atomicAdd(next, 1u);
  for (;;) {
    __syncthreads();  // every wave's stores of the last tile are acknowledged; LDS is free
    if (threadIdx.x == 0) {
      if (finished >= 0) atomicAdd(ctr + finished, 1u);
      s_item = pending;
      if (pending < qe) pending = qb + atomicAdd(next, 1u);
    }
    __syncthreads();
    const uint32_t i = __builtin_amdgcn_readfirstlane(s_item);
    if (i >= qe) break;
    const BatchItem it = items[i];
    const uint32_t xp = __builtin_amdgcn_readfirstlane(it.xp);
    const uint32_t tile = __builtin_amdgcn_readfirstlane(it.tile);
    const int wait_ctr = ph[xp].wait_ctr;
    if (wait_ctr >= 0) {
      if (threadIdx.x == 0) {
        const uint32_t need = (uint32_t)ph[xp].wait_need;
        uint32_t spins = 0;
        while (atomicAdd(ctr + wait_ctr, 0u) < need) {
          if (++spins > spin_limit) {
            atomicOr(ctr + kCtrErr, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      __syncthreads();
    }
    run_tile(ph[xp].kind, ph[xp].nlog, &ph[xp].d, tile);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have reached L2
    finished = ph[xp].done_ctr;
  }
}

bool batch_supported(int L) { return L >= 10 && L <= 22; }

constexpr int kBatchSingleMax = 12;

int default_batch_split(int L) { return L <= kBatchSingleMax ? L : (L + 1) / 2; }

// n1_log = L (one pass) or a split whose passes the batch kernel has shapes for
bool batch_split_ok(int L, int s) {
  if (s == L) return L <= kBatchSingleMax;
  return s >= 7 && s <= 11 && L - s >= 6 && L - s <= 11;
}

// Algorithmic HBM bytes of one transform of 2^L in one direction: 16 bytes
// per element and HBM round trip, one round trip up to 2^13 (the transform
// fits one workgroup's LDS), two above.
double ntt_alg_bytes(int L) { return 16.0 * (double)(1u << L) * (L <= 13 ? 1.0 : 2.0); }

// Transforms of one length in the grouped mode: the per-length pass
// launches with an offset table (transform b of the launch at word offs[b]).
struct BatchGroup {
  int L = 0, k = 0;  // k <= 65535 transforms (one launch row each): a length with more
                     // transforms gets several groups
  DevBuf offs;
  PassSpec sp[4];
};

}  // namespace
}  // namespace pcgsq

using namespace pcgsq;

struct pcgsq_ntt_batch {
  pcgsq_ctx* ctx = nullptr;
  int count = 0, ops = 3;
  int mode = PCGSQ_NTT_BATCH_PERSISTENT;
  std::vector<std::unique_ptr<BatchGroup>> groups;  // grouped mode, longest first
  std::map<int, std::unique_ptr<NttPlan>> gplans;   // grouped mode: L -> plan
  std::vector<int> logN;
  std::vector<size_t> off;
  size_t total = 0;
  std::map<int, int> split;                       // L -> n1_log
  std::map<int, std::unique_ptr<NttPlan>> plans;  // L -> tables
  DevBuf a, f, r, tmp_f, tmp_i, phases, items, ctr;
  uint32_t nitems = 0;
  size_t lds = 0;
  int grid = 0, wgs_per_cu = 0, nctr = 0;
  uint32_t spin_limit = 1u << 22;
  double alg_bytes = 0;
  BatchQueues queues{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  // Grouped mode: one offset table and plan per length; each plan takes the
  // shape the context's per-length launches use (the NTT autotuner's).
  void build_grouped() {
    constexpr size_t kMaxRows = 65535;  // grid.y of one launch
    groups.clear();
    gplans.clear();
    std::map<int, std::vector<uint64_t>, std::greater<int>> by_len;
    for (int x = 0; x < count; ++x) by_len[logN[(size_t)x]].push_back(off[(size_t)x]);
    uint64_t* A = a.as<uint64_t>();
    uint64_t* F = f.as<uint64_t>();
    uint64_t* R = r.as<uint64_t>();
    uint64_t* TF = tmp_f.as<uint64_t>();
    uint64_t* TI = tmp_i.as<uint64_t>();
    nitems = 0;
    for (auto& kv : by_len) {
      const int L = kv.first;
      const NttPlan& tuned = ctx->ntt.plan(1 << L);
      std::unique_ptr<NttPlan>& pp = gplans[L];
      pp.reset(new NttPlan(1 << L, tuned.n1_log));
      pp->variant = tuned.variant;
      pp->cols1 = tuned.cols1;
      pp->cols2 = tuned.cols2;
      const NttPlan& pl = *pp;
      const bool one = pl.single();
      for (size_t b0 = 0; b0 < kv.second.size(); b0 += kMaxRows) {
        std::unique_ptr<BatchGroup> g(new BatchGroup);
        g->L = L;
        g->k = (int)std::min(kMaxRows, kv.second.size() - b0);
        upload(g->offs,
               std::vector<uint64_t>(kv.second.begin() + (ptrdiff_t)b0,
                                     kv.second.begin() + (ptrdiff_t)(b0 + (size_t)g->k)),
               "ntt batch offsets");
        if (ops & 1) {
          g->sp[0] = plan_pass(pl, kPhFwd1, A, nullptr, one ? F : TF, g->k);
          g->sp[1] = plan_pass(pl, kPhFwd2, TF, nullptr, F, g->k);
        }
        if (ops & 2) {
          const uint64_t* inv_in = (ops & 1) ? F : A;
          g->sp[2] = plan_pass(pl, kPhInv1, inv_in, nullptr, one ? R : TI, g->k);
          g->sp[3] = plan_pass(pl, kPhInv2, TI, nullptr, R, g->k);
        }
        for (PassSpec& sp : g->sp) {
          if (sp.kind < 0) continue;
          sp.d.boff = g->offs.as<uint64_t>();
          if (one) sp.ncols = 1;  // one transform per launch row
          ++nitems;
        }
        groups.push_back(std::move(g));
      }
    }
    grid = 0;
    lds = 0;
    nctr = 0;
  }

  // Tiles, work lists and launch shape from the current splits.
  void build() {
    drop_graph();
    alg_bytes = 0;
    for (int x = 0; x < count; ++x)
      alg_bytes += ((ops & 1 ? 1 : 0) + (ops & 2 ? 1 : 0)) * ntt_alg_bytes(logN[(size_t)x]);
    if (mode == PCGSQ_NTT_BATCH_GROUPED) {
      build_grouped();
      return;
    }
    groups.clear();
    gplans.clear();
    int xccs = 1, cus = 0;
    hip_check(hipDeviceGetAttribute(&xccs, hipDeviceAttributeNumberOfXccs, ctx->device),
              "ntt batch XCD count");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
              "ntt batch CU count");
    const int nq = std::max(1, std::min(xccs, kMaxQueues));
    // transforms onto XCDs: longest first onto the least loaded
    std::vector<int> order((size_t)count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int u, int v) { return logN[(size_t)u] > logN[(size_t)v]; });
    std::vector<double> load((size_t)nq, 0.0);
    std::vector<int> xq((size_t)count), local((size_t)count);
    std::vector<int> per_q((size_t)nq, 0);
    for (int x : order) {
      const int qi = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      xq[(size_t)x] = qi;
      local[(size_t)x] = per_q[(size_t)qi]++;
      load[(size_t)qi] += (double)logN[(size_t)x] * (double)(1 << logN[(size_t)x]);
    }
    // done counters: per list a 128-byte aligned region of 4 per transform
    std::vector<int> cbase((size_t)nq);
    int c = kCtrDone;
    for (int qi = 0; qi < nq; ++qi) {
      cbase[(size_t)qi] = c;
      c += (4 * per_q[(size_t)qi] + kCtrLine - 1) / kCtrLine * kCtrLine;
    }
    nctr = c;

    std::vector<BatchPhase> ph((size_t)count * 4);
    lds = 0;
    uint64_t* A = a.as<uint64_t>();
    uint64_t* F = f.as<uint64_t>();
    uint64_t* R = r.as<uint64_t>();
    uint64_t* TF = tmp_f.as<uint64_t>();
    uint64_t* TI = tmp_i.as<uint64_t>();
    for (int x = 0; x < count; ++x) {
      const int L = logN[(size_t)x];
      auto& pl = plans[L];
      if (!pl || pl->n1_log != split[L]) pl.reset(new NttPlan(1 << L, split[L]));
      pl->variant = 0;
      pl->cols1 = pl->cols2 = 0;
      const size_t o = off[(size_t)x];
      const uint64_t* inv_in = (ops & 1) ? F + o : A + o;
      PassSpec sp[4];
      if (ops & 1) {
        sp[0] = plan_pass(*pl, kPhFwd1, A + o, nullptr, pl->single() ? F + o : TF + o, 1);
        sp[1] = plan_pass(*pl, kPhFwd2, TF + o, nullptr, F + o, 1);
      }
      if (ops & 2) {
        sp[2] = plan_pass(*pl, kPhInv1, inv_in, nullptr, pl->single() ? R + o : TI + o, 1);
        sp[3] = plan_pass(*pl, kPhInv2, TI + o, nullptr, R + o, 1);
      }
      const int cb = cbase[(size_t)xq[(size_t)x]] + 4 * local[(size_t)x];
      int last = -1;
      for (int p = 0; p < 4; ++p) {
        BatchPhase& b = ph[(size_t)x * 4 + p];
        b.kind = sp[p].kind;
        b.tiles = 0;
        b.wait_ctr = -1;
        b.wait_need = 0;
        b.done_ctr = cb + p;
        if (b.kind < 0) continue;
        b.nlog = sp[p].nlog;
        b.d = sp[p].d;
        const int rmax = batch_rmax(b.nlog);
        int C = std::max(1, kBatchThreads >> (b.nlog - rmax));
        C = std::min(C, sp[p].ncols);
        b.d.c_log = ilog2_exact((uint64_t)C);
        b.d.ncols = sp[p].ncols;
        b.tiles = (sp[p].ncols + C - 1) / C;
        if (last >= 0) {
          b.wait_ctr = ph[(size_t)last].done_ctr;
          b.wait_need = ph[(size_t)last].tiles;
        }
        last = x * 4 + p;
        lds = std::max(lds, ntt_tile_lds_bytes(b.nlog, C, batch_groups(b.nlog)));
      }
    }
    // One pass-major work list per XCD, longest transforms first within a
    // pass.  In the first passes, which read their length's per-element
    // twiddle table (the same for every transform of that length), the
    // list's transforms of one length go tile-major: a group of tiles (those
    // sharing 128-byte lines) of each such transform in turn, so that the
    // group's slice of the table is read from HBM once per XCD, not once per
    // transform.
    std::vector<BatchItem> it;
    queues = BatchQueues{};
    queues.n = (uint32_t)nq;
    for (int qi = 0; qi < nq; ++qi) {
      queues.beg[qi] = (uint32_t)it.size();
      for (int p = 0; p < 4; ++p) {
        const bool first = p == kPhFwd1 || p == kPhInv1;
        for (size_t i = 0; i < order.size();) {
          // the run of this list's transforms of one length (order is sorted by length)
          const int L = logN[(size_t)order[i]];
          std::vector<int> xs;
          size_t j = i;
          for (; j < order.size() && logN[(size_t)order[j]] == L; ++j)
            if (xq[(size_t)order[j]] == qi) xs.push_back(order[j]);
          i = j;
          if (xs.empty()) continue;
          const BatchPhase& b0 = ph[(size_t)xs[0] * 4 + p];
          const int grp = (first && b0.kind >= 0 && b0.kind != kFwdWhole && b0.kind != kInvWhole)
                              ? std::max(1, 16 >> b0.d.c_log)
                              : b0.tiles;
          for (int g0 = 0; g0 < b0.tiles; g0 += std::max(1, grp))
            for (int x : xs) {
              const BatchPhase& b = ph[(size_t)x * 4 + p];
              for (int t = g0; t < std::min(b.tiles, g0 + std::max(1, grp)); ++t)
                it.push_back(BatchItem{(uint32_t)(x * 4 + p), (uint32_t)t});
            }
        }
      }
      queues.end[qi] = (uint32_t)it.size();
    }
    nitems = (uint32_t)it.size();
    phases.ensure(ph.size() * sizeof(BatchPhase), "hipMalloc ntt batch phases");
    items.ensure(std::max<size_t>(1, it.size()) * sizeof(BatchItem), "hipMalloc ntt batch items");
    ctr.ensure((size_t)nctr * sizeof(uint32_t), "hipMalloc ntt batch counters");
    hip_check(hipMemcpy(phases.ptr, ph.data(), ph.size() * sizeof(BatchPhase),
                        hipMemcpyHostToDevice), "upload batch phases");
    if (!it.empty())
      hip_check(hipMemcpy(items.ptr, it.data(), it.size() * sizeof(BatchItem),
                          hipMemcpyHostToDevice), "upload batch items");
    if (lds > 65536)
      hip_check(hipFuncSetAttribute((const void*)ntt_batch_kernel,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                "ntt batch: raise LDS limit");
    int per_cu = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ntt_batch_kernel,
                                                           kBatchThreads, lds),
              "ntt batch occupancy");
    require(per_cu >= 1, "ntt batch: kernel does not fit on a CU", kErrUnsupported);
    if (wgs_per_cu > 0) per_cu = std::min(per_cu, wgs_per_cu);
    // a multiple of the XCD count, so that every list has workgroups
    grid = std::max(nq, per_cu * cus / nq * nq);
  }

  void enqueue(hipStream_t st) {
    if (mode == PCGSQ_NTT_BATCH_GROUPED) {
      for (auto& g : groups)
        for (const PassSpec& sp : g->sp)
          if (sp.kind >= 0) ntt_launch_pass(st, sp, g->k, nullptr, "ntt.batch.pass");
      return;
    }
    hipLaunchKernelGGL(batch_reset_kernel, dim3((nctr + 255) / 256), dim3(256), 0, st,
                       ctr.as<uint32_t>(), nctr);
    hipLaunchKernelGGL(ntt_batch_kernel, dim3((unsigned)grid), dim3(kBatchThreads), lds, st,
                       phases.as<BatchPhase>(), items.as<BatchItem>(), queues,
                       ctr.as<uint32_t>(), spin_limit);
    hip_check(hipGetLastError(), "launch ntt_batch_kernel");
  }

  // After a run: every list drained and no wait ran out of spins.
  void check() {
    if (mode == PCGSQ_NTT_BATCH_GROUPED) return;  // plain launches: nothing to drain
    std::vector<uint32_t> h((size_t)kCtrDone);
    hip_check(hipMemcpy(h.data(), ctr.ptr, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
              "ntt batch status");
    require(h[kCtrErr] == 0, "ntt batch: a dependency wait exceeded its spin limit", kErrDevice);
    for (uint32_t qi = 0; qi < queues.n; ++qi)
      require(h[qi * kCtrLine] >= queues.end[qi] - queues.beg[qi],
              "ntt batch: a work list was not drained (no workgroup ran on its XCD)",
              kErrDevice);
  }

  ~pcgsq_ntt_batch() { drop_graph(); }
};

extern "C" {

int pcgsq_ntt_batch_create(pcgsq_ctx* ctx, const int* logN, int count, int ops,
                           pcgsq_ntt_batch** out) {
  return guarded([&] {
    require(ctx && logN && out, "ntt_batch_create: null argument");
    require(count >= 1 && count <= (1 << 16), "ntt_batch_create: count must be 1..65536");
    require(ops >= 1 && ops <= 3, "ntt_batch_create: ops must be 1 (forward), 2 (inverse) or 3");
    CtxGuard g(ctx);
    std::unique_ptr<pcgsq_ntt_batch> b(new pcgsq_ntt_batch);
    b->ctx = ctx;
    b->count = count;
    b->ops = ops;
    for (int i = 0; i < count; ++i) {
      require(batch_supported(logN[i]), "ntt_batch_create: lengths must be 2^10..2^22",
              kErrUnsupported);
      b->logN.push_back(logN[i]);
      b->off.push_back(b->total);
      b->total += (size_t)1 << logN[i];
      b->split[logN[i]] = default_batch_split(logN[i]);
    }
    const size_t bytes = b->total * sizeof(uint64_t);
    b->a.ensure(bytes, "hipMalloc ntt batch input");
    b->f.ensure(bytes, "hipMalloc ntt batch forward");
    b->r.ensure(bytes, "hipMalloc ntt batch inverse");
    b->tmp_f.ensure(bytes, "hipMalloc ntt batch scratch");
    b->tmp_i.ensure(bytes, "hipMalloc ntt batch scratch");
    b->build();
    *out = b.release();
  });
}

int pcgsq_ntt_batch_destroy(pcgsq_ntt_batch* b) {
  if (!b) return PCGSQ_OK;
  return guarded([&] {
    CtxGuard g(b->ctx);
    hip_check(hipStreamSynchronize(b->ctx->stream), "ntt batch sync");
    delete b;
  });
}

int pcgsq_ntt_batch_tune(pcgsq_ntt_batch* b, int logN, int n1_log, int wgs_per_cu,
                         uint32_t spin_limit) {
  return guarded([&] {
    require(b, "ntt_batch_tune: null batch");
    CtxGuard g(b->ctx);
    if (logN > 0) {
      require(b->split.count(logN), "ntt_batch_tune: no transform of that length in the batch");
      const int s = n1_log > 0 ? n1_log : default_batch_split(logN);
      require(batch_split_ok(logN, s),
              "ntt_batch_tune: one pass needs N <= 2^12; a split needs a column pass of 2^7..2^11 "
              "and a row pass of 2^6..2^11");
      b->split[logN] = s;
    }
    require(wgs_per_cu >= 0 && wgs_per_cu <= 16, "ntt_batch_tune: wgs_per_cu must be 0..16");
    b->wgs_per_cu = wgs_per_cu;
    b->spin_limit = spin_limit ? spin_limit : (1u << 22);
    hip_check(hipStreamSynchronize(b->ctx->stream), "ntt batch sync");
    b->build();
  });
}

int pcgsq_ntt_batch_set_mode(pcgsq_ntt_batch* b, int mode) {
  return guarded([&] {
    require(b, "ntt_batch_set_mode: null batch");
    require(mode == PCGSQ_NTT_BATCH_PERSISTENT || mode == PCGSQ_NTT_BATCH_GROUPED,
            "ntt_batch_set_mode: mode must be PCGSQ_NTT_BATCH_PERSISTENT or _GROUPED");
    CtxGuard g(b->ctx);
    hip_check(hipStreamSynchronize(b->ctx->stream), "ntt batch sync");
    b->mode = mode;
    b->build();
  });
}

int pcgsq_ntt_batch_dev(pcgsq_ntt_batch* b, uint64_t** a, uint64_t** f, uint64_t** r,
                        size_t* total_words) {
  return guarded([&] {
    require(b, "ntt_batch_dev: null batch");
    if (a) *a = b->a.as<uint64_t>();
    if (f) *f = b->f.as<uint64_t>();
    if (r) *r = b->r.as<uint64_t>();
    if (total_words) *total_words = b->total;
  });
}

int pcgsq_ntt_batch_load(pcgsq_ntt_batch* b, const uint64_t* host) {
  return guarded([&] {
    require(b && host, "ntt_batch_load: null argument");
    CtxGuard g(b->ctx);
    hip_check(hipMemcpyAsync(b->a.ptr, host, b->total * sizeof(uint64_t), hipMemcpyHostToDevice,
                             b->ctx->stream), "ntt batch upload");
    hip_check(hipStreamSynchronize(b->ctx->stream), "ntt batch upload");
  });
}

int pcgsq_ntt_batch_run(pcgsq_ntt_batch* b, int use_graph) {
  return guarded([&] {
    require(b, "ntt_batch_run: null batch");
    CtxGuard g(b->ctx);
    hipStream_t st = b->ctx->stream;
    if (!use_graph) {
      ProfZone pz(&b->ctx->prof, st, "ntt.batch");
      b->enqueue(st);
      return;
    }
    if (!b->exec) {
      hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "ntt batch capture");
      b->enqueue(st);
      hip_check(hipStreamEndCapture(st, &b->graph), "ntt batch capture");
      hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0),
                "ntt batch graph instantiate");
    }
    ProfZone pz(&b->ctx->prof, st, "ntt.batch");
    hip_check(hipGraphLaunch(b->exec, st), "ntt batch graph launch");
  });
}

int pcgsq_ntt_batch_finish(pcgsq_ntt_batch* b) {
  return guarded([&] {
    require(b, "ntt_batch_finish: null batch");
    CtxGuard g(b->ctx);
    hip_check(hipStreamSynchronize(b->ctx->stream), "ntt batch sync");
    b->check();
  });
}

int pcgsq_ntt_batch_read(pcgsq_ntt_batch* b, uint64_t* f_host, uint64_t* r_host) {
  return guarded([&] {
    require(b, "ntt_batch_read: null batch");
    CtxGuard g(b->ctx);
    hip_check(hipStreamSynchronize(b->ctx->stream), "ntt batch sync");
    const size_t bytes = b->total * sizeof(uint64_t);
    if (f_host) hip_check(hipMemcpy(f_host, b->f.ptr, bytes, hipMemcpyDeviceToHost), "ntt batch read");
    if (r_host) hip_check(hipMemcpy(r_host, b->r.ptr, bytes, hipMemcpyDeviceToHost), "ntt batch read");
  });
}

int pcgsq_ntt_batch_info(pcgsq_ntt_batch* b, int* nitems, int* grid, size_t* lds_bytes,
                         double* alg_bytes) {
  return guarded([&] {
    require(b, "ntt_batch_info: null batch");
    if (nitems) *nitems = (int)b->nitems;
    if (grid) *grid = b->grid;
    if (lds_bytes) *lds_bytes = b->lds;
    if (alg_bytes) *alg_bytes = b->alg_bytes;
  });
}

}  // extern "C"
