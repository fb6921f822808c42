/*
 * Field FFp<P> and negacyclic NTT<P>: CPU restatement (test infrastructure
 * only; see oracle.h).
 *
 * Follows CPU_baseline/ffp.h:15-59 (u128 '%' reduction, add/sub with a
 * conditional subtract, extended-GCD inverse) and CPU_baseline/ntt.h:23-158
 * (psi = g^((P-1)/2N) with g the smallest primitive root, pre-twist psi^i,
 * bit-reverse + iterative DIT with roots omega^(j*N/len), inverse scaled by
 * N^-1 psi^-i).  Forward output is in natural order: out[k] = a(psi^(2k+1)).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef unsigned __int128 u128;

uint64_t orc_fmod(uint64_t x) { return x % ORC_P; }

uint64_t orc_fadd(uint64_t a, uint64_t b) {
    uint64_t s = a + b; /* a, b < P < 2^62 */
    return s >= ORC_P ? s - ORC_P : s;
}

uint64_t orc_fsub(uint64_t a, uint64_t b) {
    return a >= b ? a - b : a + ORC_P - b;
}

uint64_t orc_fneg(uint64_t a) { return a == 0 ? 0 : ORC_P - a; }

uint64_t orc_fmul(uint64_t a, uint64_t b) {
    return (uint64_t)(((u128)a * (u128)b) % ORC_P);
}

uint64_t orc_fpow(uint64_t b, uint64_t e) {
    uint64_t r = 1;
    b %= ORC_P;
    while (e) {
        if (e & 1) r = orc_fmul(r, b);
        b = orc_fmul(b, b);
        e >>= 1;
    }
    return r;
}

/* ffp.h:48-59: extended Euclid on signed 64-bit, result reduced into [0,P). */
uint64_t orc_finv(uint64_t a) {
    int64_t t0 = 0, t1 = 1;
    int64_t r0 = (int64_t)ORC_P, r1 = (int64_t)a;
    while (r1 != 0) {
        int64_t q = r0 / r1, tmp;
        tmp = t0 - q * t1; t0 = t1; t1 = tmp;
        tmp = r0 - q * r1; r0 = r1; r1 = tmp;
    }
    int64_t m = t0 % (int64_t)ORC_P;
    if (m < 0) m += (int64_t)ORC_P;
    return (uint64_t)m;
}

/* ntt.h:125-147: smallest primitive root via the distinct prime factors of
 * P-1 (trial division). */
uint64_t orc_primitive_root(void) {
    uint64_t phi = ORC_P - 1, x = phi;
    uint64_t f[64];
    int nf = 0;
    for (uint64_t d = 2; d * d <= x; ++d) {
        if (x % d) continue;
        f[nf++] = d;
        while (x % d == 0) x /= d;
    }
    if (x > 1) f[nf++] = x;
    for (uint64_t g = 2; g < ORC_P; ++g) {
        int ok = 1;
        for (int i = 0; i < nf; ++i)
            if (orc_fpow(g, phi / f[i]) == 1) { ok = 0; break; }
        if (ok) return g;
    }
    return 0;
}

struct orc_ntt {
    int N;
    uint64_t psi, inv_N;
    uint64_t *psi_pow, *psi_inv_pow, *root_fwd, *root_inv;
};

orc_ntt* orc_ntt_new(int N) {
    if (N <= 0 || (N & (N - 1)) != 0) return NULL;
    if ((ORC_P - 1) % (2 * (uint64_t)N) != 0) return NULL;
    orc_ntt* t = (orc_ntt*)calloc(1, sizeof(orc_ntt));
    t->N = N;
    t->inv_N = orc_finv((uint64_t)N);
    uint64_t g = orc_primitive_root();
    uint64_t psi = orc_fpow(g, (ORC_P - 1) / (2 * (uint64_t)N));
    uint64_t psi_inv = orc_finv(psi);
    t->psi = psi;
    t->psi_pow = (uint64_t*)malloc(sizeof(uint64_t) * N);
    t->psi_inv_pow = (uint64_t*)malloc(sizeof(uint64_t) * N);
    t->psi_pow[0] = t->psi_inv_pow[0] = 1;
    for (int i = 1; i < N; ++i) {
        t->psi_pow[i] = orc_fmul(t->psi_pow[i - 1], psi);
        t->psi_inv_pow[i] = orc_fmul(t->psi_inv_pow[i - 1], psi_inv);
    }
    uint64_t omega = orc_fmul(psi, psi), omega_inv = orc_fmul(psi_inv, psi_inv);
    int half = N / 2 > 0 ? N / 2 : 1;
    t->root_fwd = (uint64_t*)malloc(sizeof(uint64_t) * half);
    t->root_inv = (uint64_t*)malloc(sizeof(uint64_t) * half);
    t->root_fwd[0] = t->root_inv[0] = 1;
    for (int i = 1; i < N / 2; ++i) {
        t->root_fwd[i] = orc_fmul(t->root_fwd[i - 1], omega);
        t->root_inv[i] = orc_fmul(t->root_inv[i - 1], omega_inv);
    }
    return t;
}

void orc_ntt_free(orc_ntt* t) {
    if (!t) return;
    free(t->psi_pow); free(t->psi_inv_pow); free(t->root_fwd); free(t->root_inv);
    free(t);
}

int orc_ntt_size(const orc_ntt* t) { return t->N; }
uint64_t orc_ntt_psi(const orc_ntt* t) { return t->psi; }

/* ntt.h:175-197: bit-reverse permutation then radix-2 DIT stages. */
static void ntt_core(const orc_ntt* t, uint64_t* a, const uint64_t* roots) {
    const int N = t->N;
    for (int i = 1, j = 0; i < N; ++i) {
        int bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { uint64_t x = a[i]; a[i] = a[j]; a[j] = x; }
    }
    for (int len = 2; len <= N; len <<= 1) {
        int half = len >> 1, step = N / len;
        for (int i = 0; i < N; i += len)
            for (int j = 0; j < half; ++j) {
                uint64_t u = a[i + j];
                uint64_t v = orc_fmul(a[i + j + half], roots[j * step]);
                a[i + j] = orc_fadd(u, v);
                a[i + j + half] = orc_fsub(u, v);
            }
    }
}

void orc_ntt_forward(const orc_ntt* t, uint64_t* a) {
    for (int i = 0; i < t->N; ++i) a[i] = orc_fmul(a[i], t->psi_pow[i]);
    ntt_core(t, a, t->root_fwd);
}

void orc_ntt_inverse(const orc_ntt* t, uint64_t* a) {
    ntt_core(t, a, t->root_inv);
    for (int i = 0; i < t->N; ++i)
        a[i] = orc_fmul(a[i], orc_fmul(t->inv_N, t->psi_inv_pow[i]));
}

/* ntt.h:81-88 (out may alias a or b) */
void orc_ntt_poly_mul(const orc_ntt* t, uint64_t* c, const uint64_t* a,
                      const uint64_t* b) {
    const int N = t->N;
    uint64_t* fa = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* fb = (uint64_t*)malloc(sizeof(uint64_t) * N);
    memcpy(fa, a, sizeof(uint64_t) * N);
    memcpy(fb, b, sizeof(uint64_t) * N);
    orc_ntt_forward(t, fa);
    orc_ntt_forward(t, fb);
    for (int i = 0; i < N; ++i) fa[i] = orc_fmul(fa[i], fb[i]);
    orc_ntt_inverse(t, fa);
    memcpy(c, fa, sizeof(uint64_t) * N);
    free(fa);
    free(fb);
}

/* pcg_ole.h:92-106: O(N^2) negacyclic product (X^N = -1). */
void orc_poly_mul_schoolbook(uint64_t* c, const uint64_t* a,
                             const uint64_t* b, int N) {
    uint64_t* r = (uint64_t*)calloc(N, sizeof(uint64_t));
    for (int i = 0; i < N; ++i) {
        if (a[i] == 0) continue;
        for (int j = 0; j < N; ++j) {
            if (b[j] == 0) continue;
            int k = i + j;
            uint64_t v = orc_fmul(a[i], b[j]);
            if (k < N) r[k] = orc_fadd(r[k], v);
            else r[k - N] = orc_fsub(r[k - N], v);
        }
    }
    memcpy(c, r, sizeof(uint64_t) * N);
    free(r);
}

/* pcg_ole.h:119-132, with a per-call plan instead of the thread-local
 * cache (same arithmetic). */
void orc_poly_mul(uint64_t* c, const uint64_t* a, const uint64_t* b, int N) {
    if ((ORC_P - 1) % (2 * (uint64_t)N) == 0) {
        orc_ntt* t = orc_ntt_new(N);
        orc_ntt_poly_mul(t, c, a, b);
        orc_ntt_free(t);
    } else {
        orc_poly_mul_schoolbook(c, a, b, N);
    }
}

void orc_poly_mul_batch(uint64_t* c, const uint64_t* a, const uint64_t* b,
                        int batch, int N, int threads) {
    orc_ntt* t = orc_ntt_new(N);
    if (!t) {
        for (int k = 0; k < batch; ++k)
            orc_poly_mul_schoolbook(c + (size_t)k * N, a + (size_t)k * N,
                                    b + (size_t)k * N, N);
        return;
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int k = 0; k < batch; ++k)
        orc_ntt_poly_mul(t, c + (size_t)k * N, a + (size_t)k * N,
                         b + (size_t)k * N);
    orc_ntt_free(t);
}
