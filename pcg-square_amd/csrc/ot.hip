// Oblivious transfer for the two-party preprocessing: Chou-Orlandi base OTs
// on the host (libcrypto's P-256) and IKNP extension on the device.  The
// reference gets its OTs from emp-ot's FerretCOT
// (CPU_baseline/bool_circuit/circuit.cpp:21-25), an emp dependency that is
// absent here; this is a different, published OT extension with the same
// interface to its users: random OTs that TripleGen-style triples and
// Gilboa products are built from.
#include <string.h>

#include <openssl/bn.h>
#include <openssl/ec.h>
#include <openssl/evp.h>
#include <openssl/obj_mac.h>

#include <string>

#include "ot.hpp"

namespace pcgsq {

// ---- base OTs (Chou and Orlandi, "The Simplest Protocol for Oblivious
// Transfer", LATINCRYPT 2015), semi-honest, over P-256 ---------------------
//   sender:   a random, A = aG
//   receiver: b_i random, B_i = b_i G + s_i A
//   sender:   k0_i = H(A, B_i, a B_i), k1_i = H(A, B_i, a (B_i - A))
//   receiver: k_{s_i} = H(A, B_i, b_i A)
// H = SHA-256 over a tag, the direction, i and the three point encodings,
// truncated to 128 bits.

namespace {

void ossl(bool ok, const char* what) {
  if (!ok) throw Error(kErrInternal, std::string("base OT: ") + what + " failed");
}

// A scalar in [1, order) from the party's own seed.
BIGNUM* seeded_scalar(uint64_t seed, uint64_t kind, uint64_t i, const BIGNUM* order, BN_CTX* bn) {
  uint8_t raw[40];
  for (int w = 0; w < 5; ++w) {
    const uint64_t v = dealer_hd(seed, 17, kind, i * 8 + (uint64_t)w);
    for (int k = 0; k < 8; ++k) raw[8 * w + k] = (uint8_t)(v >> (8 * k));
  }
  BIGNUM* x = BN_bin2bn(raw, sizeof(raw), nullptr);
  ossl(x != nullptr, "BN_bin2bn");
  ossl(BN_nnmod(x, x, order, bn) == 1, "BN_nnmod");
  if (BN_is_zero(x)) ossl(BN_one(x) == 1, "BN_one");
  return x;
}

void encode(const EC_GROUP* g, const EC_POINT* p, uint8_t* out, BN_CTX* bn) {
  memset(out, 0, BaseOtParty::kPointBytes);
  const size_t n =
      EC_POINT_point2oct(g, p, POINT_CONVERSION_COMPRESSED, out, BaseOtParty::kPointBytes, bn);
  ossl(n != 0, "EC_POINT_point2oct");
}

Blk key_hash(int dir, int i, const uint8_t* A, const uint8_t* B, const uint8_t* K) {
  uint8_t buf[16 + 1 + 4 + 3 * BaseOtParty::kPointBytes];
  memcpy(buf, "pcgsq base ot v1", 16);
  buf[16] = (uint8_t)dir;
  for (int k = 0; k < 4; ++k) buf[17 + k] = (uint8_t)(i >> (8 * k));
  memcpy(buf + 21, A, BaseOtParty::kPointBytes);
  memcpy(buf + 21 + BaseOtParty::kPointBytes, B, BaseOtParty::kPointBytes);
  memcpy(buf + 21 + 2 * BaseOtParty::kPointBytes, K, BaseOtParty::kPointBytes);
  uint8_t md[EVP_MAX_MD_SIZE];
  unsigned int mdlen = 0;
  ossl(EVP_Digest(buf, sizeof(buf), md, &mdlen, EVP_sha256(), nullptr) == 1 && mdlen >= 16,
       "EVP_Digest");
  Blk k{0, 0};
  for (int j = 0; j < 8; ++j) {
    k.lo |= (uint64_t)md[j] << (8 * j);
    k.hi |= (uint64_t)md[8 + j] << (8 * j);
  }
  return k;
}

}  // namespace

struct BaseOtParty::Impl {
  int party = 0;
  uint64_t seed = 0;
  EC_GROUP* grp = nullptr;
  BN_CTX* bn = nullptr;
  BIGNUM* a = nullptr;          // sender scalar (direction 1 - party)
  uint8_t Aenc[kPointBytes] = {};
  BIGNUM* b[kKappa] = {};       // receiver scalars (direction party)
  Blk s{0, 0};                  // receiver choice bits
  uint8_t Apeer[kPointBytes] = {};
  std::vector<uint8_t> Benc;    // the receiver's B_i encodings
  int step = 0;

  ~Impl() {
    BN_free(a);
    for (BIGNUM* x : b) BN_free(x);
    BN_CTX_free(bn);
    EC_GROUP_free(grp);
  }
};

BaseOtParty::BaseOtParty(int party, uint64_t seed) : impl_(new Impl) {
  require(party == 0 || party == 1, "base OT: party must be 0 or 1");
  Impl& m = *impl_;
  m.party = party;
  m.seed = seed;
  m.grp = EC_GROUP_new_by_curve_name(NID_X9_62_prime256v1);
  m.bn = BN_CTX_new();
  if (!m.grp || !m.bn) {
    delete impl_;
    throw Error(kErrInternal, "base OT: P-256 group unavailable");
  }
}

BaseOtParty::~BaseOtParty() { delete impl_; }

std::vector<uint8_t> BaseOtParty::round1() {
  Impl& m = *impl_;
  require(m.step == 0, "base OT: round1 out of order", kErrInternal);
  const BIGNUM* order = EC_GROUP_get0_order(m.grp);
  m.a = seeded_scalar(m.seed, 0, 0, order, m.bn);
  EC_POINT* A = EC_POINT_new(m.grp);
  ossl(A != nullptr, "EC_POINT_new");
  const bool ok = EC_POINT_mul(m.grp, A, m.a, nullptr, nullptr, m.bn) == 1;
  if (ok) encode(m.grp, A, m.Aenc, m.bn);
  EC_POINT_free(A);
  ossl(ok, "EC_POINT_mul");
  m.step = 1;
  return std::vector<uint8_t>(m.Aenc, m.Aenc + kPointBytes);
}

std::vector<uint8_t> BaseOtParty::round2(const uint8_t* peer_msg1, size_t bytes) {
  Impl& m = *impl_;
  require(m.step == 1, "base OT: round2 out of order", kErrInternal);
  require(bytes == kMsg1Bytes, "base OT: first message has the wrong size", kErrCallback);
  memcpy(m.Apeer, peer_msg1, kPointBytes);
  const BIGNUM* order = EC_GROUP_get0_order(m.grp);
  m.s = Blk{dealer_hd(m.seed, 17, 2, 0), dealer_hd(m.seed, 17, 2, 1)};
  EC_POINT* Ap = EC_POINT_new(m.grp);
  EC_POINT* B = EC_POINT_new(m.grp);
  ossl(Ap && B, "EC_POINT_new");
  bool ok = EC_POINT_oct2point(m.grp, Ap, m.Apeer, kPointBytes, m.bn) == 1;
  m.Benc.assign((size_t)kKappa * kPointBytes, 0);
  for (int i = 0; ok && i < kKappa; ++i) {
    m.b[i] = seeded_scalar(m.seed, 1, (uint64_t)i, order, m.bn);
    const bool si = ((i < 64 ? m.s.lo >> i : m.s.hi >> (i - 64)) & 1) != 0;
    ok = EC_POINT_mul(m.grp, B, m.b[i], nullptr, nullptr, m.bn) == 1 &&
         (!si || EC_POINT_add(m.grp, B, B, Ap, m.bn) == 1);
    if (ok) encode(m.grp, B, m.Benc.data() + (size_t)i * kPointBytes, m.bn);
  }
  EC_POINT_free(Ap);
  EC_POINT_free(B);
  ossl(ok, "receiver points");
  m.step = 2;
  return m.Benc;
}

void BaseOtParty::finish(const uint8_t* peer_msg2, size_t bytes, BaseOts* out) {
  Impl& m = *impl_;
  require(m.step == 2, "base OT: finish out of order", kErrInternal);
  require(bytes == kMsg2Bytes, "base OT: second message has the wrong size", kErrCallback);
  const int dsend = 1 - m.party;   // direction where this party is the base-OT sender
  const int drecv = m.party;
  EC_POINT* A = EC_POINT_new(m.grp);
  EC_POINT* Ap = EC_POINT_new(m.grp);
  EC_POINT* B = EC_POINT_new(m.grp);
  EC_POINT* K = EC_POINT_new(m.grp);
  ossl(A && Ap && B && K, "EC_POINT_new");
  bool ok = EC_POINT_oct2point(m.grp, A, m.Aenc, kPointBytes, m.bn) == 1 &&
            EC_POINT_oct2point(m.grp, Ap, m.Apeer, kPointBytes, m.bn) == 1 &&
            EC_POINT_invert(m.grp, A, m.bn) == 1;   // A now holds -A
  uint8_t Ke[kPointBytes];
  for (int i = 0; ok && i < kKappa; ++i) {
    // sender role: the peer's B_i
    const uint8_t* Bi = peer_msg2 + (size_t)i * kPointBytes;
    ok = EC_POINT_oct2point(m.grp, B, Bi, kPointBytes, m.bn) == 1 &&
         EC_POINT_mul(m.grp, K, nullptr, B, m.a, m.bn) == 1;
    if (!ok) break;
    encode(m.grp, K, Ke, m.bn);
    out->k0[i] = key_hash(dsend, i, m.Aenc, Bi, Ke);
    ok = EC_POINT_add(m.grp, B, B, A, m.bn) == 1 &&
         EC_POINT_mul(m.grp, K, nullptr, B, m.a, m.bn) == 1;
    if (!ok) break;
    encode(m.grp, K, Ke, m.bn);
    out->k1[i] = key_hash(dsend, i, m.Aenc, Bi, Ke);
    // receiver role: b_i A_peer
    ok = EC_POINT_mul(m.grp, K, nullptr, Ap, m.b[i], m.bn) == 1;
    if (!ok) break;
    encode(m.grp, K, Ke, m.bn);
    out->ks[i] = key_hash(drecv, i, m.Apeer, m.Benc.data() + (size_t)i * kPointBytes, Ke);
  }
  EC_POINT_free(A);
  EC_POINT_free(Ap);
  EC_POINT_free(B);
  EC_POINT_free(K);
  ossl(ok, "key derivation");
  out->s = m.s;
  m.step = 3;
}

// ---- IKNP extension on the device --------------------------------------------
//
// Receiver (choice bits c, both seeds of every column r < 128):
//   T_r = G(k0_r), U_r = T_r ^ G(k1_r) ^ c           (U goes to the sender)
// Sender (choice bits s, seeds ks_r = k_{s_r, r}):
//   Q_r = G(ks_r) ^ s_r U_r = T_r ^ s_r c
// Row j of the m x 128 matrices: q_j = t_j ^ c_j s, so
//   sender: M0_j = H(j, q_j), M1_j = H(j, q_j ^ s);  receiver: H(j, t_j) = M_{c_j}.
// G(k) is the ChaCha8 keystream keyed by k (nonce: session, direction, use
// 0; block b covers OTs 512 b .. 512 b + 511 of the column), and
// H(j, x) = the first 16 bytes of ChaCha8 keyed by x at counter j (use 1).
//
// One wave per chunk of 512 OTs: lane i computes column block b of rows i
// and 64 + i, and the 128 x 64 bit tiles are transposed with wave ballots
// (ballot j over the lanes' bit j is row j of the tile's low or high half).

namespace {

__device__ __forceinline__ void column_block(Blk k, uint64_t b, uint64_t nonce, uint64_t* w) {
  uint32_t k0, k1, k2, k3, o[16];
  chacha8_words_of(k, k0, k1, k2, k3);
  chacha8<16>(k0, k1, k2, k3, b, nonce, o);
#pragma unroll
  for (int q = 0; q < 8; ++q) w[q] = (uint64_t)o[2 * q] | ((uint64_t)o[2 * q + 1] << 32);
}

__device__ __forceinline__ Blk ot_hash(uint64_t j, uint64_t lo, uint64_t hi, uint64_t nonce) {
  uint32_t o[4];
  chacha8<4>((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32), j, nonce, o);
  return Blk{(uint64_t)o[0] | ((uint64_t)o[1] << 32), (uint64_t)o[2] | ((uint64_t)o[3] << 32)};
}

// Row (lane) of the 64 x 64 tile whose columns are the lanes' words:
// bit r of the result = bit `lane` of lane r's word.
__device__ __forceinline__ uint64_t tile_row(uint64_t word) {
  const int lane = (int)threadIdx.x & 63;
  uint64_t mine = 0;
#pragma unroll 8
  for (int j = 0; j < 64; ++j) {
    const uint64_t m = __ballot((int)((word >> j) & 1ULL));
    if (lane == j) mine = m;
  }
  return mine;
}

__global__ __launch_bounds__(64) void iknp_receive_kernel(const Blk* __restrict__ k0,
                                                          const Blk* __restrict__ k1,
                                                          uint64_t ng, uint64_t nh,
                                                          const uint64_t* __restrict__ choice,
                                                          uint64_t* __restrict__ U,
                                                          Blk* __restrict__ out) {
  const uint64_t b = blockIdx.x;
  const int lane = (int)threadIdx.x;
  uint64_t t[2][8];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r = lane + 64 * h;
    uint64_t g1[8];
    column_block(k0[r], b, ng, t[h]);
    column_block(k1[r], b, ng, g1);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      U[((size_t)b * kKappa + (size_t)r) * 8 + (size_t)q] = t[h][q] ^ g1[q] ^ choice[8 * b + q];
  }
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    const uint64_t lo = tile_row(t[0][q]), hi = tile_row(t[1][q]);
    const uint64_t j = b * kOtChunk + (uint64_t)q * 64 + (uint64_t)lane;
    out[j] = ot_hash(j, lo, hi, nh);
  }
}

__global__ __launch_bounds__(64) void iknp_send_kernel(const Blk* __restrict__ ks, Blk s,
                                                       uint64_t ng, uint64_t nh,
                                                       const uint64_t* __restrict__ U,
                                                       Blk* __restrict__ out0,
                                                       Blk* __restrict__ out1) {
  const uint64_t b = blockIdx.x;
  const int lane = (int)threadIdx.x;
  uint64_t qv[2][8];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r = lane + 64 * h;
    column_block(ks[r], b, ng, qv[h]);
    const bool sr = ((h == 0 ? s.lo : s.hi) >> lane) & 1ULL;
    if (sr) {
#pragma unroll
      for (int q = 0; q < 8; ++q) qv[h][q] ^= U[((size_t)b * kKappa + (size_t)r) * 8 + (size_t)q];
    }
  }
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    const uint64_t lo = tile_row(qv[0][q]), hi = tile_row(qv[1][q]);
    const uint64_t j = b * kOtChunk + (uint64_t)q * 64 + (uint64_t)lane;
    out0[j] = ot_hash(j, lo, hi, nh);
    out1[j] = ot_hash(j, lo ^ s.lo, hi ^ s.hi, nh);
  }
}

}  // namespace

void iknp_receive(hipStream_t st, const Blk* k0, const Blk* k1, uint64_t sid, int dir,
                  const uint64_t* choice, size_t m, uint64_t* U, Blk* out) {
  require(m % kOtChunk == 0, "iknp: OT count must be a multiple of 512", kErrInternal);
  if (m == 0) return;
  require(m / kOtChunk < (1ULL << 31), "iknp: too many OTs", kErrUnsupported);
  hipLaunchKernelGGL(iknp_receive_kernel, dim3((unsigned)(m / kOtChunk)), dim3(64), 0, st, k0, k1,
                     iknp_nonce(sid, dir, 0), iknp_nonce(sid, dir, 1), choice, U, out);
  hip_check(hipGetLastError(), "launch iknp_receive_kernel");
}

void iknp_send(hipStream_t st, const Blk* ks, Blk s, uint64_t sid, int dir, const uint64_t* U,
               size_t m, Blk* out0, Blk* out1) {
  require(m % kOtChunk == 0, "iknp: OT count must be a multiple of 512", kErrInternal);
  if (m == 0) return;
  require(m / kOtChunk < (1ULL << 31), "iknp: too many OTs", kErrUnsupported);
  hipLaunchKernelGGL(iknp_send_kernel, dim3((unsigned)(m / kOtChunk)), dim3(64), 0, st, ks, s,
                     iknp_nonce(sid, dir, 0), iknp_nonce(sid, dir, 1), U, out0, out1);
  hip_check(hipGetLastError(), "launch iknp_send_kernel");
}

}  // namespace pcgsq
