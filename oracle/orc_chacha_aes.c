/*
 * ChaCha8 PRG, the ChaCha8 tree PRG / out-hash, and AES-128 with the CCR
 * tree PRG: CPU restatement (test infrastructure only; see oracle.h).
 *
 * ChaCha8: CPU_baseline/chacha8.h:23-150 -- "expand 32-byte k" constants,
 * 8 key words little endian, word 12/13 = 64-bit block counter, word 14/15 =
 * 64-bit nonce, 4 double rounds, final addition of the input state, output
 * words little endian.
 * Tree PRG: half_tree_dpf/tree_prg.h:80-111 (key = parent || 0^128, nonce 0,
 * child k = 16-byte chunk k of blocks ctr = 0..ceil(nk/4)-1).
 * Out-hash: half_tree_dpf/tree_prg.h:128-148 (key = leaf || 0^128,
 * nonce = OUT_DOMAIN, first 16 bytes).
 * AES-128: FIPS-197 (the emp-tool AES the reference calls is standard
 * AES-128 on the block's memory bytes); the S-box is derived here from the
 * GF(2^8) inverse and the affine map rather than copied as a table.
 * CCR tree PRG: half_tree_dpf/mkey_prp.h:24-79 and tree_prg.h:42-68.
 */
#include <string.h>

#include "oracle.h"

static uint32_t ld32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}

static void st32(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}

static uint32_t rotl32(uint32_t v, int n) { return (v << n) | (v >> (32 - n)); }

#define QR(a, b, c, d)                      \
    a += b; d ^= a; d = rotl32(d, 16);      \
    c += d; b ^= c; b = rotl32(b, 12);      \
    a += b; d ^= a; d = rotl32(d, 8);       \
    c += d; b ^= c; b = rotl32(b, 7);

void orc_chacha8_block(const uint8_t key[32], uint64_t nonce, uint64_t ctr,
                       uint8_t out[64]) {
    uint32_t s[16];
    s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
    for (int i = 0; i < 8; ++i) s[4 + i] = ld32(key + 4 * i);
    s[12] = (uint32_t)ctr; s[13] = (uint32_t)(ctr >> 32);
    s[14] = (uint32_t)nonce; s[15] = (uint32_t)(nonce >> 32);
    uint32_t x[16];
    memcpy(x, s, sizeof(x));
    for (int r = 0; r < 4; ++r) {
        QR(x[0], x[4], x[8], x[12])
        QR(x[1], x[5], x[9], x[13])
        QR(x[2], x[6], x[10], x[14])
        QR(x[3], x[7], x[11], x[15])
        QR(x[0], x[5], x[10], x[15])
        QR(x[1], x[6], x[11], x[12])
        QR(x[2], x[7], x[8], x[13])
        QR(x[3], x[4], x[9], x[14])
    }
    for (int i = 0; i < 16; ++i) st32(out + 4 * i, x[i] + s[i]);
}

static void blk_bytes(orc_blk b, uint8_t* p) {
    for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(b.lo >> (8 * i));
    for (int i = 0; i < 8; ++i) p[8 + i] = (uint8_t)(b.hi >> (8 * i));
}

static orc_blk bytes_blk(const uint8_t* p) {
    orc_blk b = {0, 0};
    for (int i = 0; i < 8; ++i) b.lo |= (uint64_t)p[i] << (8 * i);
    for (int i = 0; i < 8; ++i) b.hi |= (uint64_t)p[8 + i] << (8 * i);
    return b;
}

void orc_tree_hash_chacha8(orc_blk* out, const orc_blk* parents, size_t n,
                           int nk) {
    const int nblocks = (nk + 3) / 4;
    for (size_t j = 0; j < n; ++j) {
        uint8_t key[32] = {0};
        blk_bytes(parents[j], key);
        uint8_t buf[256];
        for (int b = 0; b < nblocks; ++b)
            orc_chacha8_block(key, 0, (uint64_t)b, buf + 64 * b);
        for (int k = 0; k < nk; ++k) out[j * nk + k] = bytes_blk(buf + 16 * k);
    }
}

void orc_out_hash(orc_blk* out, const orc_blk* leaves, size_t n) {
    for (size_t j = 0; j < n; ++j) {
        uint8_t key[32] = {0}, buf[64];
        blk_bytes(leaves[j], key);
        orc_chacha8_block(key, ORC_OUT_DOMAIN, 0, buf);
        out[j] = bytes_blk(buf);
    }
}

uint64_t orc_out_hash_lo64(orc_blk leaf) {
    orc_blk h;
    orc_out_hash(&h, &leaf, 1);
    return h.lo;
}

/* ---- AES-128 ------------------------------------------------------------ */

static uint8_t g_sbox[256];
static int g_sbox_ready = 0;

static uint8_t gf_mul(uint8_t a, uint8_t b) {
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}

static uint8_t rotl8(uint8_t v, int n) { return (uint8_t)((v << n) | (v >> (8 - n))); }

/* FIPS-197 Sec. 5.1.1: multiplicative inverse in GF(2^8) followed by the
 * affine transformation b ^ rotl(b,1..4) ^ 0x63. */
static void build_sbox(void) {
    for (int x = 0; x < 256; ++x) {
        uint8_t inv = 0;
        if (x) {
            /* x^254 = x^-1 */
            uint8_t r = 1, base = (uint8_t)x;
            int e = 254;
            while (e) {
                if (e & 1) r = gf_mul(r, base);
                base = gf_mul(base, base);
                e >>= 1;
            }
            inv = r;
        }
        g_sbox[x] = (uint8_t)(inv ^ rotl8(inv, 1) ^ rotl8(inv, 2) ^
                              rotl8(inv, 3) ^ rotl8(inv, 4) ^ 0x63);
    }
    g_sbox_ready = 1;
}

__attribute__((constructor)) static void orc_aes_init(void) { build_sbox(); }

void orc_aes128_expand_key(const uint8_t key[16], uint8_t rk[176]) {
    if (!g_sbox_ready) build_sbox();
    static const uint8_t rcon[11] = {0x00, 0x01, 0x02, 0x04, 0x08, 0x10,
                                     0x20, 0x40, 0x80, 0x1b, 0x36};
    memcpy(rk, key, 16);
    for (int i = 4; i < 44; ++i) {
        uint8_t t[4];
        memcpy(t, rk + 4 * (i - 1), 4);
        if (i % 4 == 0) {
            uint8_t t0 = t[0];
            t[0] = g_sbox[t[1]] ^ rcon[i / 4];
            t[1] = g_sbox[t[2]];
            t[2] = g_sbox[t[3]];
            t[3] = g_sbox[t0];
        }
        for (int k = 0; k < 4; ++k) rk[4 * i + k] = rk[4 * (i - 4) + k] ^ t[k];
    }
}

void orc_aes128_encrypt(const uint8_t rk[176], const uint8_t in[16],
                        uint8_t out[16]) {
    if (!g_sbox_ready) build_sbox();
    uint8_t s[16], t[16];
    for (int i = 0; i < 16; ++i) s[i] = in[i] ^ rk[i];
    for (int round = 1; round <= 10; ++round) {
        for (int i = 0; i < 16; ++i) s[i] = g_sbox[s[i]];
        /* ShiftRows: row r of column c takes row r of column c+r. */
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) t[4 * c + r] = s[4 * ((c + r) & 3) + r];
        if (round < 10) {
            for (int c = 0; c < 4; ++c) {
                uint8_t a0 = t[4 * c], a1 = t[4 * c + 1], a2 = t[4 * c + 2],
                        a3 = t[4 * c + 3];
                s[4 * c + 0] = gf_mul(a0, 2) ^ gf_mul(a1, 3) ^ a2 ^ a3;
                s[4 * c + 1] = a0 ^ gf_mul(a1, 2) ^ gf_mul(a2, 3) ^ a3;
                s[4 * c + 2] = a0 ^ a1 ^ gf_mul(a2, 2) ^ gf_mul(a3, 3);
                s[4 * c + 3] = gf_mul(a0, 3) ^ a1 ^ a2 ^ gf_mul(a3, 2);
            }
        } else {
            memcpy(s, t, 16);
        }
        for (int i = 0; i < 16; ++i) s[i] ^= rk[16 * round + i];
    }
    memcpy(out, s, 16);
}

void orc_aes_tree_round_keys(const uint8_t master[16], int nk, uint8_t* rks) {
    uint8_t mrk[176];
    orc_aes128_expand_key(master, mrk);
    for (int k = 0; k < nk; ++k) {
        orc_blk tweak = {(uint64_t)k, 0}; /* makeBlock(0, k) */
        uint8_t tb[16], cb[16], sk[16];
        blk_bytes(tweak, tb);
        orc_aes128_encrypt(mrk, tb, cb);
        for (int i = 0; i < 16; ++i) sk[i] = cb[i] ^ tb[i];
        orc_aes128_expand_key(sk, rks + (size_t)k * 176);
    }
}

void orc_tree_hash_aes(const uint8_t* rks, orc_blk* out,
                       const orc_blk* parents, size_t n, int nk) {
    for (size_t j = 0; j < n; ++j) {
        uint8_t pb[16], cb[16];
        blk_bytes(parents[j], pb);
        for (int k = 0; k < nk; ++k) {
            orc_aes128_encrypt(rks + (size_t)k * 176, pb, cb);
            orc_blk c = bytes_blk(cb);
            c.lo ^= parents[j].lo;
            c.hi ^= parents[j].hi;
            out[j * nk + k] = c;
        }
    }
}
