/*
 * Utilities for the oracle (test infrastructure only; see oracle.h):
 *  - FNV-1a-64 over 64-bit words, as used for the naive NTT checksum
 *    (GPU_baseline/naive_ntt_bench.cu:92-96: h ^= x; h *= 1099511628211).
 *  - mt19937_64 (Matsumoto & Nishimura 2004; the engine behind
 *    std::mt19937_64 used by the reference recipes naive_ntt_bench.cu:127 and
 *    test/out_hash_cuda_bitexact.cu:57).  Pinned by the C++ standard's
 *    requirement that the 10000th output of a default-seeded engine is
 *    9981545732273789042.
 */
#include "oracle.h"

uint64_t orc_fnv1a64(const uint64_t* v, size_t n, uint64_t h) {
    for (size_t i = 0; i < n; ++i) {
        h ^= v[i];
        h *= 1099511628211ull;
    }
    return h;
}

#define MT_NN 312
#define MT_MM 156
#define MT_A 0xB5026F5AA96619E9ULL
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x7FFFFFFFULL

void orc_mt64_seed(orc_mt64* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < MT_NN; ++i)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) +
                   (uint64_t)i;
    s->idx = MT_NN;
}

uint64_t orc_mt64_next(orc_mt64* s) {
    if (s->idx >= MT_NN) {
        int i;
        uint64_t x;
        for (i = 0; i < MT_NN - MT_MM; ++i) {
            x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
            s->mt[i] = s->mt[i + MT_MM] ^ (x >> 1) ^ ((x & 1) ? MT_A : 0);
        }
        for (; i < MT_NN - 1; ++i) {
            x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
            s->mt[i] = s->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ ((x & 1) ? MT_A : 0);
        }
        x = (s->mt[MT_NN - 1] & MT_UM) | (s->mt[0] & MT_LM);
        s->mt[MT_NN - 1] = s->mt[MT_MM - 1] ^ (x >> 1) ^ ((x & 1) ? MT_A : 0);
        s->idx = 0;
    }
    uint64_t x = s->mt[s->idx++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

void orc_mt64_fill(orc_mt64* s, uint64_t* out, size_t n, uint64_t mod) {
    for (size_t i = 0; i < n; ++i) {
        uint64_t v = orc_mt64_next(s);
        out[i] = mod ? v % mod : v;
    }
}
