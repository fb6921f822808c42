// One party's share of the 2PC preprocessing of PCG_OLE<P>::gen_and_expand
// (CPU_baseline/pcg_ole_2pc/pcg_ole_impl.h:62-199) on the device:
//   * its own draws from its own seed: noise (positions, payloads), its Delta
//     share, its F_p Beaver randomness r, s;
//   * random OTs in both directions by IKNP extension (ot.hpp);
//   * AND triples from the random OTs (one per cross term), used by GMW
//     (bool_circuit/circuit.cpp:245-380) for the Kogge-Stone position adder
//     (bool_circuit/adder.h:21-63) and the F_DeltaShiftShare circuit
//     (half_tree_dpf/dpf.cpp:172-240);
//   * beta = pay0 * pay1 and the Beaver triple's rs by Gilboa OT products
//     (TripleGen::zp_multiply / zp_triple, circuit.cpp:91-200).
// Every message between the parties is returned by pre_step as a PreMsg, in
// protocol order; the caller moves the bytes (a host callback for one party
// per process, a device copy for both parties in one process).
#pragma once

#include <stdint.h>

#include "device_common.hpp"
#include "internal.hpp"
#include "ot.hpp"

namespace pcgsq {

// Shape of the preprocessing of one expansion.
struct PreShape {
  int c = 0, t = 0, t_log = 0, B = 0, nblk = 0, ninst = 0;
  int levels = 0, nk = 0, w = 0, D_log = 0;
  int nb = 0;        // adder width: log2(bin) bits
  int S = 0;         // prefix steps of the adder
  int ops = 0;       // bit-sliced AND ops of the adder, 1 + 2S
  uint64_t bin = 0;
  // OT index layout (identical in both directions up to sec_gil; direction 0
  // adds the Gilboa OTs): adder gates (g, op, bit) at sec_add + (g*ops + op)*nb
  // + bit; DltSft AND-tree gates (g, e, j) at sec_tree + (g*L*nk + e)*(w-1) + j;
  // Delta ANDs (g, e) at sec_delta + g*L*nk + e (e = i*nk + k); Gilboa OT j of
  // product q of g at sec_gil + (g*3 + q)*64 + j.
  size_t sec_add = 0, sec_tree = 0, sec_delta = 0, sec_gil = 0;
  size_t m[2] = {0, 0};   // OTs of direction 0 and 1 (multiples of 512)
};
PreShape make_pre_shape(int N, int c, int t, int w);

// A message of the protocol: send `send_bytes` from device memory d_send and
// receive `recv_bytes` into d_recv (either may be 0).  Party 0 sends first.
// When nblocks > 1 the message is nblocks equal per-DPF-block pieces, block
// after block (the reference's per-block rounds); batch = AND gates per
// block (circuit_and), OTs or products.
struct PreMsg {
  const char* phase = "";
  int level = 0;
  int batch = 0;
  const void* d_send = nullptr;
  size_t send_bytes = 0;
  void* d_recv = nullptr;
  size_t recv_bytes = 0;
  int nblocks = 1;
};

// Where one party's preprocessing writes its results: the plan's arrays
// with this party's slots (instance g of party p at p*ninst + g).
struct PreOut {
  uint64_t* noise = nullptr;   // [pos | pay][party][c][t]
  Blk* roots = nullptr;        // [party][ninst]
  Blk* dlt = nullptr;          // [party][ninst][levels][nk]
  uint64_t* bt = nullptr;      // [beta | r | s | rs][party][ninst]
};

// One party's device state and progress.
struct PreRun {
  int party = 0;
  PreShape sh;
  uint64_t seed = 0, sid = 0;
  PreOut out;
  hipStream_t st = nullptr;
  int stage = 0;
  DevBuf base;                 // k0[128], k1[128], ks[128] on the device
  Blk base_s{0, 0};
  DevBuf choice, u_mine, u_peer, mc, m0, m1;
  DevBuf tri_add, tri_tree, tri_dlt, state, fac, open_mine[2], open_peer, gil, gshare;
};

// Device buffers for this shape; base OTs uploaded (once per party pair).
void pre_prepare(PreRun& r, const PreShape& sh, int party, const BaseOts& base);
// Start an expansion's preprocessing from the party's own seed with IKNP
// session sid (both parties use the same sid; never reuse one).
void pre_begin(PreRun& r, uint64_t seed, uint64_t sid, const PreOut& out, hipStream_t st);
// Enqueue the work up to the next message and describe it; false when the
// preprocessing is complete (its results are in `out` once the stream
// reaches this point).
bool pre_step(PreRun& r, PreMsg* msg);

}  // namespace pcgsq
