// Oblivious transfer for the two-party preprocessing (the reference runs
// emp-ot's FerretCOT: CPU_baseline/bool_circuit/circuit.cpp:21-25).
//
//   * Base OTs: 128 per direction, Chou-Orlandi "simplest OT" over P-256
//     (libcrypto's EC group arithmetic, SHA-256 as the key hash), two
//     message rounds for both directions at once.  Host code.
//   * OT extension: IKNP (Ishai-Kilian-Nissim-Petrank 2003) on the device,
//     producing random OTs: the sender gets (M0_j, M1_j), the receiver
//     M_{c_j} for its chosen bits c_j.  The column PRG and the
//     correlation-robust hash are ChaCha8.
//
// Directions: in direction d, party d is the extension sender and party
// 1 - d the extension receiver.  The receiver of direction d was the base-OT
// sender of that direction (it holds both seeds k0, k1 of every column); the
// sender was the base-OT receiver (choice bits s, the chosen seeds ks).
#pragma once

#include <stdint.h>

#include <functional>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"

namespace pcgsq {

constexpr int kKappa = 128;       // base OTs per direction (columns of the IKNP matrix)
constexpr size_t kOtChunk = 512;  // OTs per ChaCha8 block of every column

// One party's base-OT results for both directions.
struct BaseOts {
  Blk k0[kKappa], k1[kKappa];  // direction 1 - party: both seeds of every column
  Blk s;                       // direction party: the choice bits
  Blk ks[kKappa];              // direction party: the seeds chosen by s
};

// The base-OT protocol of one party, in steps (both directions at once):
//   msg1 = round1()                    its A = a*G (33 bytes), sent to the peer
//   msg2 = round2(peer's msg1)         its 128 points B_i (128 * 33 bytes)
//   finish(peer's msg2, out)
// Scalars and choice bits come from the party's own seed.
class BaseOtParty {
 public:
  static constexpr size_t kPointBytes = 33;
  static constexpr size_t kMsg1Bytes = kPointBytes;
  static constexpr size_t kMsg2Bytes = kKappa * kPointBytes;
  BaseOtParty(int party, uint64_t seed);
  ~BaseOtParty();
  BaseOtParty(const BaseOtParty&) = delete;
  BaseOtParty& operator=(const BaseOtParty&) = delete;
  std::vector<uint8_t> round1();
  std::vector<uint8_t> round2(const uint8_t* peer_msg1, size_t bytes);
  void finish(const uint8_t* peer_msg2, size_t bytes, BaseOts* out);

 private:
  struct Impl;
  Impl* impl_;
};

// ChaCha8 nonce of an IKNP use: column PRG (use 0) or message hash (use 1)
// of session `sid` and direction `dir`.
PCGSQ_HD uint64_t iknp_nonce(uint64_t sid, int dir, int use) {
  return 0x6f74000000000000ULL ^ (sid << 8) ^ ((uint64_t)dir << 4) ^ (uint64_t)use;
}

// Receiver of direction `dir`: m OTs (a multiple of kOtChunk) with choice
// bits choice[m/64]; column seeds k0, k1 [128] on the device.  Writes the
// matrix U for the sender, [m/512][128][8] words (chunk-major), and the
// received messages out[m] = H(j, t_j).
void iknp_receive(hipStream_t st, const Blk* k0, const Blk* k1, uint64_t sid, int dir,
                  const uint64_t* choice, size_t m, uint64_t* U, Blk* out);
// Sender of direction `dir`: column seeds ks [128] chosen by s, the
// receiver's U; writes out0[m] = H(j, q_j), out1[m] = H(j, q_j ^ s).
void iknp_send(hipStream_t st, const Blk* ks, Blk s, uint64_t sid, int dir, const uint64_t* U,
               size_t m, Blk* out0, Blk* out1);

}  // namespace pcgsq
