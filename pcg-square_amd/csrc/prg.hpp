// Tree PRGs for the half-tree DPF on gfx950.
//
//   ChaCha8 (half_tree_dpf/tree_prg.h:80-111): H_k(x) = 16-byte chunk k of the
//     ChaCha8 keystream keyed by x || 0^128, nonce 0, block counter k/4.  All
//     state lives in VGPRs; no table, no LDS.
//   AES (half_tree_dpf/tree_prg.h:42-68, mkey_prp.h:24-79): CCR hash
//     H_k(x) = AES_{sub_k}(x) ^ x with FIPS-197 AES-128.  One T-table
//     (column word of MixColumns(SubBytes)) lives in LDS and the other three
//     tables are byte rotations of it; round keys for the nk sub-keys are
//     staged in LDS as 32-bit words.  The table is stored 32 times
//     interleaved (entry e, copy r at word 32e + r) and lane l reads copy
//     l mod 32, so the random byte lookups of a wave never collide in an LDS
//     bank; the address still costs one instruction after the byte extract.
#pragma once

#include "device_common.hpp"

namespace pcgsq {

enum PrgKind : int { kPrgAes = 0, kPrgChaCha8 = 1 };

constexpr int kAesCopies = 32;

// LDS staging for the AES tables (only used by AES instantiations).
struct AesLds {
  uint32_t t0[256 * kAesCopies];  // entry e, copy r at 32e + r
  uint32_t rk[15 * 44];           // nk <= 15 sub-keys x 44 round-key words
};

// T-table lookups: one table in any memory ...
struct AesTable {
  const uint32_t* t0;
  __device__ __forceinline__ uint32_t operator()(uint32_t byte) const { return t0[byte]; }
};
// ... or this lane's copy of the interleaved LDS table.
struct AesTableLds {
  const uint32_t* t0;  // AesLds::t0 + (lane mod 32)
  __device__ __forceinline__ uint32_t operator()(uint32_t byte) const {
    return t0[byte * kAesCopies];
  }
};
__device__ __forceinline__ AesTableLds aes_table(const AesLds* lds) {
  return AesTableLds{lds->t0 + (threadIdx.x & (kAesCopies - 1))};
}

template <class Tab>
__device__ __forceinline__ uint32_t aes_te(const Tab& t0, uint32_t x, int rot) {
  const uint32_t v = t0(x & 0xff);
  return rot ? __builtin_amdgcn_alignbit(v, v, 32 - 8 * rot) : v;
}

// AES-128 encryption of one block held as 4 little-endian column words.
template <class Tab>
__device__ __forceinline__ void aes128_encrypt_words(const Tab& t0,
                                                     const uint32_t* __restrict__ rk,
                                                     uint32_t& s0, uint32_t& s1,
                                                     uint32_t& s2, uint32_t& s3) {
  s0 ^= rk[0]; s1 ^= rk[1]; s2 ^= rk[2]; s3 ^= rk[3];
#pragma unroll
  for (int r = 1; r < 10; ++r) {
    const uint32_t a0 = aes_te(t0, s0, 0) ^ aes_te(t0, s1 >> 8, 1) ^ aes_te(t0, s2 >> 16, 2) ^
                        aes_te(t0, s3 >> 24, 3) ^ rk[4 * r + 0];
    const uint32_t a1 = aes_te(t0, s1, 0) ^ aes_te(t0, s2 >> 8, 1) ^ aes_te(t0, s3 >> 16, 2) ^
                        aes_te(t0, s0 >> 24, 3) ^ rk[4 * r + 1];
    const uint32_t a2 = aes_te(t0, s2, 0) ^ aes_te(t0, s3 >> 8, 1) ^ aes_te(t0, s0 >> 16, 2) ^
                        aes_te(t0, s1 >> 24, 3) ^ rk[4 * r + 2];
    const uint32_t a3 = aes_te(t0, s3, 0) ^ aes_te(t0, s0 >> 8, 1) ^ aes_te(t0, s1 >> 16, 2) ^
                        aes_te(t0, s2 >> 24, 3) ^ rk[4 * r + 3];
    s0 = a0; s1 = a1; s2 = a2; s3 = a3;
  }
  // final round: SubBytes + ShiftRows; S[x] is byte 1 of T0[x]
#define PCGSQ_SB(x) ((t0((x) & 0xff) >> 8) & 0xffu)
  const uint32_t b0 = PCGSQ_SB(s0) | (PCGSQ_SB(s1 >> 8) << 8) | (PCGSQ_SB(s2 >> 16) << 16) |
                      (PCGSQ_SB(s3 >> 24) << 24);
  const uint32_t b1 = PCGSQ_SB(s1) | (PCGSQ_SB(s2 >> 8) << 8) | (PCGSQ_SB(s3 >> 16) << 16) |
                      (PCGSQ_SB(s0 >> 24) << 24);
  const uint32_t b2 = PCGSQ_SB(s2) | (PCGSQ_SB(s3 >> 8) << 8) | (PCGSQ_SB(s0 >> 16) << 16) |
                      (PCGSQ_SB(s1 >> 24) << 24);
  const uint32_t b3 = PCGSQ_SB(s3) | (PCGSQ_SB(s0 >> 8) << 8) | (PCGSQ_SB(s1 >> 16) << 16) |
                      (PCGSQ_SB(s2 >> 24) << 24);
#undef PCGSQ_SB
  s0 = b0 ^ rk[40]; s1 = b1 ^ rk[41]; s2 = b2 ^ rk[42]; s3 = b3 ^ rk[43];
}

// H_k(x) for k < NK into h[0..NK).
template <int PRG, int NK>
__device__ __forceinline__ void tree_hash(Blk x, Blk* h, const AesLds* lds) {
  if constexpr (PRG == kPrgChaCha8) {
    uint32_t k0, k1, k2, k3;
    chacha8_words_of(x, k0, k1, k2, k3);
#pragma unroll
    for (int b = 0; b < (NK + 3) / 4; ++b) {
      uint32_t o[16];
      chacha8<16>(k0, k1, k2, k3, (uint64_t)b, 0, o);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = 4 * b + c;
        if (k < NK) {
          h[k].lo = (uint64_t)o[4 * c] | ((uint64_t)o[4 * c + 1] << 32);
          h[k].hi = (uint64_t)o[4 * c + 2] | ((uint64_t)o[4 * c + 3] << 32);
        }
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      uint32_t s0 = (uint32_t)x.lo, s1 = (uint32_t)(x.lo >> 32);
      uint32_t s2 = (uint32_t)x.hi, s3 = (uint32_t)(x.hi >> 32);
      aes128_encrypt_words(aes_table(lds), lds->rk + 44 * k, s0, s1, s2, s3);
      h[k].lo = ((uint64_t)s0 | ((uint64_t)s1 << 32)) ^ x.lo;
      h[k].hi = ((uint64_t)s2 | ((uint64_t)s3 << 32)) ^ x.hi;
    }
  }
}

// Stage the AES T-table and nk round-key schedules (176 B each) into LDS.
__device__ __forceinline__ void aes_stage(AesLds* lds, const uint32_t* __restrict__ t0_g,
                                          const uint8_t* __restrict__ rks, int nk) {
  for (int i = threadIdx.x; i < 256 * kAesCopies; i += blockDim.x) lds->t0[i] = t0_g[i / kAesCopies];
  for (int i = threadIdx.x; i < nk * 44; i += blockDim.x) {
    const uint8_t* p = rks + 4 * i;
    lds->rk[i] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                 ((uint32_t)p[3] << 24);
  }
  __syncthreads();
}

}  // namespace pcgsq
