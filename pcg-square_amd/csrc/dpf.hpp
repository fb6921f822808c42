// Half-tree DPF full-domain evaluation engine (host side).
//
// The expansion is the reference's fused protocol
// (half_tree_dpf/dpf.cpp:485-578, GPU_baseline/common/dpf_gpu.cu:339-501):
// level by level, every party hashes its whole frontier, XOR-reduces each
// column into its correction-word share (seeded with DltSft), exchanges the
// B*(m-1) shares, and expands with CW = mine ^ peer.  On one MI355X both
// parties' trees run in the same launches and the exchange is the other
// party's reduced share already in HBM; with a host callback the shares
// cross PCIe once per level exactly as in the reference plugin.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "device_common.hpp"
#include "internal.hpp"
#include "profiler.hpp"

namespace pcgsq {

enum LeafMode : int {
  kLeafStore = 0,    // write the 2^n leaves of every tree
  kLeafConvert = 1,  // fuse the ChaCha8 out-hash: write C|tau<<63 and partial sums
  kLeafGather = 2,   // fuse the out-hash and the output layer's gather: per output
                     // coefficient, sum +-C and one tau bit per landing leaf
};

// Host exchange hook: fill peer[0..count) from mine[0..count) for `level`.
// Returns nonzero on failure.
typedef int (*HostExchangeFn)(void* user, int level, const Blk* mine, Blk* peer,
                              size_t count);

struct DpfRun {
  int ntrees = 0;      // trees in the launch (both parties: 2 * instances)
  int n = 0, w = 1;
  int prg = 1;         // 0 = AES, 1 = ChaCha8
  // inputs (device)
  const Blk* roots = nullptr;     // [ntrees] (or the first half, see roots1)
  const Blk* dltsft = nullptr;    // [ntrees][levels][nk] (reference layout)
  // When set, trees T >= ntrees/2 read roots1[T - ntrees/2] and dltsft1
  // instead (the two parties' inputs need not be adjacent).
  const Blk* roots1 = nullptr;
  const Blk* dltsft1 = nullptr;
  const uint32_t* aes_t0 = nullptr;  // AES T-table (device) when prg == AES
  const uint8_t* aes_rks = nullptr;  // [nk][176] round keys (device)
  // exchange: in-device pairing of tree T with T + ntrees/2 when host_ex is null
  HostExchangeFn host_ex = nullptr;
  void* host_user = nullptr;
  // outputs (device)
  LeafMode mode = kLeafConvert;
  Blk* leaves = nullptr;          // kLeafStore: [ntrees][2^n]
  uint64_t* ct = nullptr;         // kLeafConvert: [ntrees][2^n]
  uint64_t* partials = nullptr;   // kLeafConvert: [ntrees][nparts][2]; kLeafGather: nparts =
                                  // 2^n / kGatherTile
  // kLeafGather: the trees are [parties][sets][B] (sets blocks of B
  // instances per party, t^2 >= B; parties = 2 for both parties in one run,
  // 1 for one party with a host exchange); block bb of the run's party p
  // sums into g_acc + p * g_party_stride + bb * N (N = 2^N_log) and writes
  // its tau bits to g_mask [mask_words][parties][sets][N].  Needs
  // 2^(n-1) >= kGatherTile, t <= 64 and at least two levels.
  uint64_t* g_acc = nullptr;
  size_t g_party_stride = 0;
  uint64_t* g_mask = nullptr;
  int mask_words = 0, t = 0, B = 0, N_log = 0, sets = 0, parties = 2;
  Blk* cw_out = nullptr;          // optional [levels][ntrees][nk] copy of the shares
  int iters_max = 8;              // tiles of 256 parents per workgroup
  Profiler* prof = nullptr;       // optional per-launch timing
};

// Scratch owned by a context; grows on demand.
struct DpfScratch {
  DevBuf fh[2];    // ping-pong frontier (nodes + hashes)
  DevBuf cw;       // [levels][ntrees][nk]
  DevBuf peer;     // [levels][ntrees][nk] (host-exchange mode)
  DevBuf t0, rks;  // AES tables when the caller passes host tables
};

// Number of partial-sum slots per tree written by kLeafConvert.
int dpf_leaf_nparts(int n, int w, int iters_max);
// ... and by kLeafGather, whose workgroups own 256 * dpf_gather_lpt(n)
// output coefficients (2, 4 or 8 per thread, as many as one bin of 2^(n-1)
// coefficients holds).
int dpf_gather_lpt(int n);
int dpf_gather_nparts(int n);

// Whether a run expands its last three tree levels inside the leaf level
// (both parties in one run, w = 1, ChaCha8, fused gather with 8 coefficients
// per thread): the correction words of those levels come from the alpha
// path, and only levels up to levels - 4 are stored.
bool dpf_deep_leaf(const DpfRun& run);
bool dpf_deep_leaf(int n, int w, int prg, int mode, bool host_exchange, int parties);
// Nodes per tree of the widest stored frontier level.
size_t dpf_frontier_nodes(int n, int w, bool deep);

// Allocates everything dpf_expand needs for `ntrees` trees of domain 2^n and
// arity 2^w, so that a later dpf_expand of that shape allocates nothing.
void dpf_reserve(DpfScratch& scratch, int n, int w, size_t ntrees, bool host_exchange,
                 bool deep = false);

// Runs the whole level loop on `stream`.  Synchronizes with the host only in
// host-exchange mode (once per level, like the reference).
void dpf_expand(hipStream_t stream, const DpfRun& run, DpfScratch& scratch);

// Host AES helpers (FIPS-197): key expansion and the T-table used on device.
void aes128_expand_key_host(const uint8_t key[16], uint8_t rk[176]);
void aes_ttable_host(uint32_t t0[256]);
// sub_k = AES_master(k as LE u64 || 0^64) ^ that block, expanded to round keys
// (half_tree_dpf/tree_prg.h:45-56, dpf.cpp:599-619).
void aes_tree_round_keys_host(const uint8_t master[16], int nk, uint8_t* rks);
void aes128_encrypt_host(const uint8_t rk[176], const uint8_t in[16], uint8_t out[16]);

}  // namespace pcgsq
