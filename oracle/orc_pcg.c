/*
 * PCG-OLE expansion for both parties, the Beaver-corrected output layer and
 * the seeded dealer: CPU restatement (test infrastructure only; see oracle.h).
 *
 * gen_and_expand follows CPU_baseline/pcg_ole_2pc/pcg_ole_impl.h:46-490:
 *   bin = N/t, D = 2*bin, dpf_n = log2 D;
 *   noise: pos[i][k] = k*bin + r % bin, payload = r % (P-1) + 1   (:62-77)
 *   per block (ii, jj), instance idx = kk*t + ll:
 *     alpha = pos0[ii][kk] % bin + pos1[jj][ll] % bin            (:118-157)
 *     beta  = payload0[ii][kk] * payload1[jj][ll], additively shared (:162-171)
 *     DPF batch full eval with DltSft_{i,k} = [digit_i(alpha)=k]*Delta (:180-199)
 *     output layer C = lo64(H'(leaf)) % P, tau = lsb(leaf)        (:330-365)
 *       A = (-1)^s sum tau, B = (-1)^(1-s) sum C + beta_s,
 *       Beaver open d = A-r, e = B-s; CW_s = d*s_s + e*r_s + rs_s (+d*e, s=0)
 *       y = (-1)^s (C + tau*CW) scattered into g[(kk+ll)*bin + d]  (:367-446)
 *   step 4: a = (1, public_a...), x = sum_i a_i e_i,
 *           z = sum_ij (a_i a_j) * fold(g_ij), fold(g)[d] = g[d] - g[d+N]
 *                                                               (:454-484)
 * The reference draws the noise, Delta, W and public_a from emp/OpenSSL
 * randomness and gets alpha/beta/DltSft/Beaver shares from 2PC
 * preprocessing (out of scope).  They come from the seeded dealer below; the
 * arithmetic that consumes them is unchanged.
 *
 * Dealer spec (shared with the product's host dealer, which must agree):
 *   U(seed, dom, a, b) = sm(sm(sm(seed ^ dom*K) ^ a) ^ b), sm = splitmix64
 *   finalizer, K = 0xD6E8FEB86659FD93.
 *   pos[s][i][k]  : r = U(seed, 1, s, i*t+k), pos = k*bin + r % bin
 *   pay[s][i][k]  : r = U(seed, 2, s, i*t+k), pay = r % (P-1) + 1
 *   global instance g = blk*B + idx, blk = ii*c + jj
 *   delta_s       : lo = U(seed,3,s,2g), hi = U(seed,3,s,2g+1);
 *                   party 0 sets lsb, party 1 clears it (dpf.cpp:78-82)
 *   root_s        : delta_s ^ W
 *   dltsft0[i][k] : lo = U(seed,4,g,2(i*nk+k)), hi = U(seed,4,g,2(i*nk+k)+1)
 *   dltsft1       : dltsft0 ^ ([digit_i(alpha)=k] ? Delta : 0)
 *   beta0 = U(seed,5,g,0) % P, beta1 = beta - beta0
 *   r_s = U(seed,6,g,s) % P, s_s = U(seed,7,g,s) % P,
 *   rs0 = U(seed,8,g,0) % P, rs1 = (r0+r1)(s0+s1) - rs0
 *   public_a[i][j] = U(pub_seed, 9, i, j) % P
 *   W = the reference's F_RAND_SEED block (lo 0x0123456789abcdef,
 *       hi 0xdeadbeefcafe0001); any public W cancels between the parties.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

static uint64_t sm64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t orc_dealer_u64(uint64_t seed, uint32_t domain, uint64_t a,
                        uint64_t b) {
    return sm64(sm64(sm64(seed ^ ((uint64_t)domain * 0xD6E8FEB86659FD93ULL)) ^ a) ^ b);
}

orc_blk orc_public_W(void) {
    orc_blk w = {0x0123456789abcdefULL, 0xdeadbeefcafe0001ULL};
    return w;
}

static int ilog2(uint64_t x) {
    int r = 0;
    while ((1ULL << r) < x) ++r;
    return r;
}

size_t orc_dealer_sizes(const orc_params* p, size_t* n_pos, size_t* n_inst,
                        size_t* n_dlt) {
    const int B = p->t * p->t;
    const int D = 2 * (p->N / p->t);
    const int n = ilog2((uint64_t)D);
    const int m = 1 << p->w, nk = m - 1, levels = n / p->w;
    *n_pos = (size_t)p->c * p->t;
    *n_inst = (size_t)p->c * p->c * B;
    *n_dlt = *n_inst * levels * nk;
    return *n_inst;
}

static int check_params(const orc_params* p) {
    if (p->N <= 0 || (p->N & (p->N - 1))) return -1;
    if (p->t <= 0 || p->N % p->t) return -1;
    if (p->c < 1) return -1;
    int D = 2 * (p->N / p->t);
    if (D & (D - 1)) return -1;
    int n = ilog2((uint64_t)D);
    if (p->w < 1 || p->w > 4 || n % p->w) return -1;
    return 0;
}

int orc_dealer(const orc_params* p, uint64_t seed, orc_party_inputs* in0,
               orc_party_inputs* in1, uint64_t* alpha_out) {
    if (check_params(p)) return -1;
    const int N = p->N, c = p->c, t = p->t, B = t * t;
    const uint64_t bin = (uint64_t)(N / t);
    const int D = 2 * (int)bin, n = ilog2((uint64_t)D);
    const int m = 1 << p->w, nk = m - 1, levels = n / p->w;
    orc_party_inputs* in[2] = {in0, in1};
    const orc_blk W = orc_public_W();

    for (int s = 0; s < 2; ++s)
        for (int i = 0; i < c; ++i)
            for (int k = 0; k < t; ++k) {
                uint64_t r = orc_dealer_u64(seed, 1, (uint64_t)s, (uint64_t)(i * t + k));
                in[s]->pos[i * t + k] = (uint64_t)k * bin + r % bin;
                r = orc_dealer_u64(seed, 2, (uint64_t)s, (uint64_t)(i * t + k));
                in[s]->pay[i * t + k] = r % (ORC_P - 1) + 1;
            }

    for (int ii = 0; ii < c; ++ii)
        for (int jj = 0; jj < c; ++jj) {
            const int blk = ii * c + jj;
            for (int idx = 0; idx < B; ++idx) {
                const int kk = idx / t, ll = idx % t;
                const uint64_t g = (uint64_t)blk * B + idx;
                const uint64_t alpha = in0->pos[ii * t + kk] % bin +
                                       in1->pos[jj * t + ll] % bin;
                if (alpha_out) alpha_out[g] = alpha;
                orc_blk d0 = {orc_dealer_u64(seed, 3, 0, 2 * g),
                              orc_dealer_u64(seed, 3, 0, 2 * g + 1)};
                orc_blk d1 = {orc_dealer_u64(seed, 3, 1, 2 * g),
                              orc_dealer_u64(seed, 3, 1, 2 * g + 1)};
                d0.lo |= 1ULL;
                d1.lo &= ~1ULL;
                const orc_blk Delta = {d0.lo ^ d1.lo, d0.hi ^ d1.hi};
                in0->root[g].lo = d0.lo ^ W.lo; in0->root[g].hi = d0.hi ^ W.hi;
                in1->root[g].lo = d1.lo ^ W.lo; in1->root[g].hi = d1.hi ^ W.hi;
                for (int i = 0; i < levels; ++i) {
                    const int digit = orc_dpf_digit(n, p->w, i, alpha);
                    for (int k = 0; k < nk; ++k) {
                        const uint64_t e = (uint64_t)(i * nk + k);
                        const size_t off = (size_t)g * levels * nk + e;
                        orc_blk s0 = {orc_dealer_u64(seed, 4, g, 2 * e),
                                      orc_dealer_u64(seed, 4, g, 2 * e + 1)};
                        in0->dltsft[off] = s0;
                        if (digit == k) { s0.lo ^= Delta.lo; s0.hi ^= Delta.hi; }
                        in1->dltsft[off] = s0;
                    }
                }
                const uint64_t beta = orc_fmul(in0->pay[ii * t + kk], in1->pay[jj * t + ll]);
                in0->beta[g] = orc_dealer_u64(seed, 5, g, 0) % ORC_P;
                in1->beta[g] = orc_fsub(beta, in0->beta[g]);
                in0->r[g] = orc_dealer_u64(seed, 6, g, 0) % ORC_P;
                in1->r[g] = orc_dealer_u64(seed, 6, g, 1) % ORC_P;
                in0->s[g] = orc_dealer_u64(seed, 7, g, 0) % ORC_P;
                in1->s[g] = orc_dealer_u64(seed, 7, g, 1) % ORC_P;
                const uint64_t rs = orc_fmul(orc_fadd(in0->r[g], in1->r[g]),
                                             orc_fadd(in0->s[g], in1->s[g]));
                in0->rs[g] = orc_dealer_u64(seed, 8, g, 0) % ORC_P;
                in1->rs[g] = orc_fsub(rs, in0->rs[g]);
            }
        }
    return 0;
}

void orc_public_a(uint64_t pub_seed, int c, int N, uint64_t* public_a) {
    for (int i = 0; i < c - 1; ++i)
        for (int j = 0; j < N; ++j)
            public_a[(size_t)i * N + j] =
                orc_dealer_u64(pub_seed, 9, (uint64_t)i, (uint64_t)j) % ORC_P;
}

/* ---- output layer -------------------------------------------------------- */

void orc_out_sums(const orc_blk* leaves, int B, size_t D, uint64_t* sumC,
                  uint64_t* sumT) {
    for (int b = 0; b < B; ++b) {
        uint64_t c = 0, t = 0;
        for (size_t d = 0; d < D; ++d) {
            const orc_blk lf = leaves[(size_t)b * D + d];
            c = orc_fadd(c, orc_out_hash_lo64(lf) % ORC_P);
            t += lf.lo & 1;
        }
        sumC[b] = c;
        sumT[b] = t;
    }
}

void orc_out_scatter_g(const orc_blk* leaves, int B, size_t D, int t,
                       int party, const uint64_t* CW, uint64_t* g, int N) {
    const size_t bin = D / 2;
    memset(g, 0, sizeof(uint64_t) * (size_t)N);
    for (int b = 0; b < B; ++b) {
        const int kk = b / t, ll = b % t;
        for (size_t d = 0; d < D; ++d) {
            const orc_blk lf = leaves[(size_t)b * D + d];
            uint64_t y = orc_out_hash_lo64(lf) % ORC_P;
            if (lf.lo & 1) y = orc_fadd(y, CW[b]);
            if (party != 0) y = orc_fneg(y);
            size_t pos = (size_t)(kk + ll) * bin + d;
            if (pos >= (size_t)N) { pos -= (size_t)N; y = orc_fneg(y); }
            g[pos] = orc_fadd(g[pos], y);
        }
    }
}

void orc_output_layer_both(const orc_blk* leaves0, const orc_blk* leaves1,
                           int B, size_t D, int t, const uint64_t* beta0,
                           const uint64_t* beta1, const uint64_t* r0,
                           const uint64_t* r1, const uint64_t* s0,
                           const uint64_t* s1, const uint64_t* rs0,
                           const uint64_t* rs1, uint64_t* g_red0,
                           uint64_t* g_red1, int N, uint64_t* cw_out) {
    const orc_blk* lv[2] = {leaves0, leaves1};
    const uint64_t* beta[2] = {beta0, beta1};
    const uint64_t* r[2] = {r0, r1};
    const uint64_t* s[2] = {s0, s1};
    const uint64_t* rs[2] = {rs0, rs1};
    uint64_t* gr[2] = {g_red0, g_red1};
    uint64_t* sumC = (uint64_t*)malloc(sizeof(uint64_t) * B * 2);
    uint64_t* sumT = (uint64_t*)malloc(sizeof(uint64_t) * B * 2);
    uint64_t* de = (uint64_t*)malloc(sizeof(uint64_t) * B * 4);
    uint64_t* cwp = (uint64_t*)malloc(sizeof(uint64_t) * B * 2);
    uint64_t* CW = (uint64_t*)malloc(sizeof(uint64_t) * B);
    /* all_C / all_tau (pcg_ole_impl.h:338-357): C = lo64(H'(leaf)) % P and
     * tau = lsb(leaf) per leaf, hashed once, with the per-instance sums. */
    uint64_t* allC[2];
    uint8_t* allT[2];
    for (int p = 0; p < 2; ++p) {
        allC[p] = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)B * D);
        allT[p] = (uint8_t*)malloc((size_t)B * D);
#pragma omp parallel for schedule(static)
        for (int b = 0; b < B; ++b) {
            uint64_t c = 0, tt = 0;
            for (size_t d = 0; d < D; ++d) {
                const orc_blk lf = lv[p][(size_t)b * D + d];
                const uint64_t C = orc_out_hash_lo64(lf) % ORC_P;
                const uint8_t tau = (uint8_t)(lf.lo & 1);
                allC[p][(size_t)b * D + d] = C;
                allT[p][(size_t)b * D + d] = tau;
                c = orc_fadd(c, C);
                tt += tau;
            }
            sumC[p * B + b] = c;
            sumT[p * B + b] = tt;
        }
    }
    /* A/B shares and the first Beaver open (pcg_ole_impl.h:358-371) */
    for (int p = 0; p < 2; ++p)
        for (int b = 0; b < B; ++b) {
            uint64_t A = orc_fmod(sumT[p * B + b]);
            if (p == 1) A = orc_fneg(A);
            uint64_t Bs = sumC[p * B + b];
            if (p == 0) Bs = orc_fneg(Bs);
            Bs = orc_fadd(Bs, beta[p][b]);
            de[(p * 2 + 0) * B + b] = orc_fsub(A, r[p][b]);
            de[(p * 2 + 1) * B + b] = orc_fsub(Bs, s[p][b]);
        }
    /* CW shares and the second open (:400-407, :434) */
    for (int p = 0; p < 2; ++p)
        for (int b = 0; b < B; ++b) {
            uint64_t dv = orc_fadd(de[0 * B + b], de[2 * B + b]);
            uint64_t ev = orc_fadd(de[1 * B + b], de[3 * B + b]);
            uint64_t cw = orc_fadd(orc_fadd(orc_fmul(dv, s[p][b]), orc_fmul(ev, r[p][b])),
                                   rs[p][b]);
            if (p == 0) cw = orc_fadd(cw, orc_fmul(dv, ev));
            cwp[p * B + b] = cw;
        }
    for (int b = 0; b < B; ++b) CW[b] = orc_fadd(cwp[b], cwp[B + b]);
    if (cw_out) memcpy(cw_out, CW, sizeof(uint64_t) * B);
    /* scatter into 2N then fold (:432-446, :474-478) */
    const size_t bin = D / 2;
    uint64_t* g2 = (uint64_t*)malloc(sizeof(uint64_t) * 2 * (size_t)N);
    for (int p = 0; p < 2; ++p) {
        memset(g2, 0, sizeof(uint64_t) * 2 * (size_t)N);
        for (int b = 0; b < B; ++b) {
            const int kk = b / t, ll = b % t;
            const size_t base = (size_t)(kk + ll) * bin;
            for (size_t d = 0; d < D; ++d) {
                uint64_t y = allC[p][(size_t)b * D + d];
                if (allT[p][(size_t)b * D + d]) y = orc_fadd(y, CW[b]);
                if (p == 1) y = orc_fneg(y);
                const size_t pos = base + d;
                if (pos < 2 * (size_t)N) g2[pos] = orc_fadd(g2[pos], y);
            }
        }
        for (int d = 0; d < N; ++d) gr[p][d] = orc_fsub(g2[d], g2[(size_t)d + N]);
    }
    for (int p = 0; p < 2; ++p) { free(allC[p]); free(allT[p]); }
    free(g2); free(sumC); free(sumT); free(de); free(cwp); free(CW);
}

/* ---- full expansion -------------------------------------------------------- */

typedef struct {
    orc_party_inputs in[2];
    uint64_t* alpha;
    void* mem;
} dealt;

static int deal(const orc_params* p, uint64_t seed, dealt* dl) {
    size_t npos, ninst, ndlt;
    orc_dealer_sizes(p, &npos, &ninst, &ndlt);
    const size_t per = npos * 2 * 8 + ninst * (16 + 4 * 8) + ndlt * 16;
    char* m = (char*)malloc(per * 2 + ninst * 8 + 64);
    if (!m) return -1;
    dl->mem = m;
    for (int s = 0; s < 2; ++s) {
        orc_party_inputs* q = &dl->in[s];
        q->pos = (uint64_t*)m; m += npos * 8;
        q->pay = (uint64_t*)m; m += npos * 8;
        q->root = (orc_blk*)m; m += ninst * 16;
        q->dltsft = (orc_blk*)m; m += ndlt * 16;
        q->beta = (uint64_t*)m; m += ninst * 8;
        q->r = (uint64_t*)m; m += ninst * 8;
        q->s = (uint64_t*)m; m += ninst * 8;
        q->rs = (uint64_t*)m; m += ninst * 8;
    }
    dl->alpha = (uint64_t*)m;
    return orc_dealer(p, seed, &dl->in[0], &dl->in[1], dl->alpha);
}

/* One DPF block (ii,jj) for both parties: instances are independent (the
 * batched exchange is a concatenation), so they run one per OpenMP task. */
static int run_block(const orc_params* p, const dealt* dl, int blk,
                     uint64_t* g_red0, uint64_t* g_red1, int threads) {
    const int N = p->N, t = p->t, B = t * t;
    const size_t D = (size_t)2 * (N / t);
    const int n = ilog2(D), m = 1 << p->w, nk = m - 1, levels = n / p->w;
    uint8_t* rks = NULL;
    if (p->prg == ORC_PRG_AES) {
        static const uint8_t S[16] = {0x2b, 0x7e, 0x15, 0x16, 0x28, 0xae, 0xd2, 0xa6,
                                      0xab, 0xf7, 0x15, 0x88, 0x09, 0xcf, 0x4f, 0x3c};
        rks = (uint8_t*)malloc((size_t)nk * 176);
        orc_aes_tree_round_keys(S, nk, rks); /* pcg_ole_impl.h:89-92 */
    }
    orc_blk* lv0 = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * D);
    orc_blk* lv1 = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * D);
    if (!lv0 || !lv1) { free(lv0); free(lv1); free(rks); return -1; }
    const size_t g0 = (size_t)blk * B;
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) reduction(|: err)
    for (int b = 0; b < B; ++b) {
        err |= orc_dpf_batch_full_eval_both(
            1, n, p->w, p->prg, rks, &dl->in[0].root[g0 + b], &dl->in[1].root[g0 + b],
            &dl->in[0].dltsft[(g0 + b) * levels * nk],
            &dl->in[1].dltsft[(g0 + b) * levels * nk], lv0 + (size_t)b * D,
            lv1 + (size_t)b * D, NULL);
    }
    if (!err && g_red0 && g_red1)
        orc_output_layer_both(lv0, lv1, B, D, t, dl->in[0].beta + g0, dl->in[1].beta + g0,
                              dl->in[0].r + g0, dl->in[1].r + g0, dl->in[0].s + g0,
                              dl->in[1].s + g0, dl->in[0].rs + g0, dl->in[1].rs + g0,
                              g_red0, g_red1, N, NULL);
    free(lv0); free(lv1); free(rks);
    return err ? -1 : 0;
}

int orc_pcg_ole_blocks(const orc_params* p, uint64_t seed, int blk_begin,
                       int blk_end, uint64_t* g_red0, uint64_t* g_red1,
                       int threads) {
    dealt dl;
    if (deal(p, seed, &dl)) { free(dl.mem); return -1; }
    int err = 0;
    for (int blk = blk_begin; blk < blk_end && !err; ++blk)
        err = run_block(p, &dl, blk,
                        g_red0 ? g_red0 + (size_t)(blk - blk_begin) * p->N : NULL,
                        g_red1 ? g_red1 + (size_t)(blk - blk_begin) * p->N : NULL,
                        threads);
    free(dl.mem);
    return err;
}

int orc_pcg_ole_expand(const orc_params* p, uint64_t seed, uint64_t pub_seed,
                       uint64_t* x0, uint64_t* z0, uint64_t* x1, uint64_t* z1,
                       int threads) {
    dealt dl;
    if (deal(p, seed, &dl)) { free(dl.mem); return -1; }
    const int N = p->N, c = p->c, t = p->t;
    uint64_t* g0 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)c * c * N);
    uint64_t* g1 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)c * c * N);
    int err = 0;
    for (int blk = 0; blk < c * c && !err; ++blk)
        err = run_block(p, &dl, blk, g0 + (size_t)blk * N, g1 + (size_t)blk * N, threads);
    if (err) { free(g0); free(g1); free(dl.mem); return -1; }

    /* step 4 (pcg_ole_impl.h:454-484) */
    uint64_t* a = (uint64_t*)calloc((size_t)c * N, sizeof(uint64_t));
    a[0] = 1;
    if (c > 1) orc_public_a(pub_seed, c, N, a + N);
    uint64_t* e = (uint64_t*)calloc((size_t)N, sizeof(uint64_t));
    uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)N);
    uint64_t* aij = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)c * c * N);
    const orc_party_inputs* in[2] = {&dl.in[0], &dl.in[1]};
    uint64_t* xs[2] = {x0, x1};
    uint64_t* zs[2] = {z0, z1};
    uint64_t* gs[2] = {g0, g1};
    for (int i = 0; i < c; ++i)
        for (int j = 0; j < c; ++j)
            orc_poly_mul(aij + (size_t)(i * c + j) * N, a + (size_t)i * N,
                         a + (size_t)j * N, N);
    for (int s = 0; s < 2; ++s) {
        memset(xs[s], 0, sizeof(uint64_t) * N);
        memset(zs[s], 0, sizeof(uint64_t) * N);
        for (int i = 0; i < c; ++i) {
            memset(e, 0, sizeof(uint64_t) * N);
            for (int k = 0; k < t; ++k) e[in[s]->pos[i * t + k]] = in[s]->pay[i * t + k];
            orc_poly_mul(tmp, a + (size_t)i * N, e, N);
            for (int d = 0; d < N; ++d) xs[s][d] = orc_fadd(xs[s][d], tmp[d]);
        }
        for (int ij = 0; ij < c * c; ++ij) {
            orc_poly_mul(tmp, aij + (size_t)ij * N, gs[s] + (size_t)ij * N, N);
            for (int d = 0; d < N; ++d) zs[s][d] = orc_fadd(zs[s][d], tmp[d]);
        }
    }
    free(a); free(e); free(tmp); free(aij); free(g0); free(g1); free(dl.mem);
    return 0;
}

/* ---- GPU_baseline/dpf_real_bench.cu:48-110 recipe -------------------------- */

static void zero_exchange(void* ctx, int level, const orc_blk* mine,
                          orc_blk* peer, size_t count) {
    (void)ctx; (void)level; (void)mine;
    memset(peer, 0, sizeof(orc_blk) * count);
}

uint64_t orc_dpf_real_bench_checksum(int n, int B, int t, int w, int threads) {
    if (t == 0) { t = 1; while ((t + 1) * (t + 1) <= B) ++t; }
    const size_t D = (size_t)1 << n;
    const int gN = (int)(t * (D / 2));
    const int levels = n / w, nk = (1 << w) - 1;
    orc_blk* roots = (orc_blk*)malloc(sizeof(orc_blk) * B);
    orc_blk* dlt = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * levels * nk);
    for (int b = 0; b < B; ++b) {
        roots[b].lo = 0x9e3779b97f4a7c15ULL * (uint64_t)(b + 1);
        roots[b].hi = (uint64_t)(b + 1);
    }
    for (size_t i = 0; i < (size_t)B * levels * nk; ++i) {
        dlt[i].lo = 0x1234567ULL * (uint64_t)(i + 1);
        dlt[i].hi = (uint64_t)i;
    }
    uint64_t* sumC = (uint64_t*)calloc(B, 8);
    uint64_t* sumT = (uint64_t*)calloc(B, 8);
    int nth = threads > 0 ? threads : 1;
    uint64_t* gpart = (uint64_t*)calloc((size_t)nth * gN, 8);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
    for (int b = 0; b < B; ++b) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        orc_blk* lv = (orc_blk*)malloc(sizeof(orc_blk) * D);
        orc_dpf_batch_full_eval_party(1, n, w, ORC_PRG_CHACHA8, NULL, &roots[b],
                                      &dlt[(size_t)b * levels * nk], zero_exchange,
                                      NULL, lv);
        const uint64_t cw = (0xABCDEFULL * (uint64_t)(b + 1)) % ORC_P;
        const int kk = b / t, ll = b % t;
        uint64_t sc = 0, st = 0;
        uint64_t* g = gpart + (size_t)tid * gN;
        for (size_t d = 0; d < D; ++d) {
            const uint64_t C = orc_out_hash_lo64(lv[d]) % ORC_P;
            const uint64_t tau = lv[d].lo & 1;
            sc = orc_fadd(sc, C);
            st += tau;
            uint64_t y = C;
            if (tau) y = orc_fadd(y, cw);
            size_t pos = (size_t)(kk + ll) * (D / 2) + d;
            if (pos >= (size_t)gN) { pos -= (size_t)gN; y = orc_fneg(y); }
            g[pos] = orc_fadd(g[pos], y);
        }
        sumC[b] = sc;
        sumT[b] = st;
        free(lv);
    }
    uint64_t cks = 0;
    for (int b = 0; b < B; ++b) cks ^= sumC[b] ^ sumT[b];
    for (int i = 0; i < gN; i += 97) {
        uint64_t v = 0;
        for (int k = 0; k < nth; ++k) v = orc_fadd(v, gpart[(size_t)k * gN + i]);
        cks ^= v;
    }
    free(roots); free(dlt); free(sumC); free(sumT); free(gpart);
    return cks;
}

/* generate_triples (CPU_baseline/pcg_ole_2pc/triple.h:45-110): two OLEs from
 * independent seeds, ring triple shares
 *   a_0 = x_{0,1}, b_0 = x_{0,2}, c_0 = a_0 b_0 + z_{0,1} + z_{0,2},
 *   a_1 = x_{1,2}, b_1 = x_{1,1}, c_1 = a_1 b_1 + z_{1,1} + z_{1,2},
 * then the forward NTT of a, b and c of each party: slot i is one Z_p triple
 * share.  Outputs are [N] each. */
int orc_generate_triples(const orc_params* p, uint64_t seed1, uint64_t seed2, uint64_t pub_seed,
                         uint64_t* a0, uint64_t* b0, uint64_t* c0, uint64_t* a1,
                         uint64_t* b1, uint64_t* c1, int threads) {
    const int N = p->N;
    const size_t n = (size_t)N;
    uint64_t* m = (uint64_t*)malloc(8 * n * sizeof(uint64_t));
    if (!m) return -1;
    uint64_t *x01 = m, *z01 = m + n, *x11 = m + 2 * n, *z11 = m + 3 * n;
    uint64_t *x02 = m + 4 * n, *z02 = m + 5 * n, *x12 = m + 6 * n, *z12 = m + 7 * n;
    if (orc_pcg_ole_expand(p, seed1, pub_seed, x01, z01, x11, z11, threads) ||
        orc_pcg_ole_expand(p, seed2, pub_seed, x02, z02, x12, z12, threads)) {
        free(m);
        return -1;
    }
    memcpy(a0, x01, n * 8);
    memcpy(b0, x02, n * 8);
    memcpy(a1, x12, n * 8);
    memcpy(b1, x11, n * 8);
    orc_poly_mul(c0, a0, b0, N);
    orc_poly_mul(c1, a1, b1, N);
    for (size_t i = 0; i < n; ++i) {
        c0[i] = orc_fadd(orc_fadd(c0[i], z01[i]), z02[i]);
        c1[i] = orc_fadd(orc_fadd(c1[i], z11[i]), z12[i]);
    }
    orc_ntt* t = orc_ntt_new(N);
    if (!t) { free(m); return -1; }
    uint64_t* v[6] = {a0, b0, c0, a1, b1, c1};
    for (int k = 0; k < 6; ++k) orc_ntt_forward(t, v[k]);
    orc_ntt_free(t);
    free(m);
    return 0;
}
