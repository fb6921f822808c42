/*
 * GMW preprocessing of the PCG-OLE for both parties: CPU restatement, gate
 * by gate (test infrastructure only; see oracle.h).
 *
 * Follows CPU_baseline/pcg_ole_2pc/pcg_ole_impl.h:114-199:
 *   3a. alpha = off0 + off1 by the Kogge-Stone adder circuit
 *       (bool_circuit/adder.h:21-63) evaluated under GMW, giving each party
 *       an XOR share of the nb+1 sum bits, nb = log2(bin);
 *   3c. DltSft by the F_DeltaShiftShare circuit (half_tree_dpf/dpf.cpp:
 *       61-255): per level i and child k the indicator
 *       prod_c (alpha_{n-w(i+1)+c} xor not k_c) over an AND tree, then 128
 *       AND gates with Delta; outputs stay XOR-shared.
 *   3b. beta = pay0 * pay1 and the F_p Beaver triples (r, s, rs) by
 *       Gilboa's OT product (bool_circuit/circuit.cpp:91-200).
 * GMW rules (bool_circuit/circuit.cpp:245-380): an input owned by party o
 * is shared as (x ^ m, m) with o holding x ^ m; XOR is local; NOT flips
 * party 0's share; AND with the triple (a, b, c), c = a & b, opens
 * d = x ^ a, e = y ^ b and gives z_0 = d&e ^ d&b_0 ^ e&a_0 ^ c_0,
 * z_1 = d&b_1 ^ e&a_1 ^ c_1.
 * Every AND triple is made from two OTs as TripleGen::generate makes it
 * (circuit.cpp:30-85): each party draws its a, b; party 0, the OT sender,
 * draws r, s and offers (r, r ^ a_0) and (s, s ^ b_0); party 1 chooses with
 * b_1 and a_1.  So c_0 = a_0&b_0 ^ r ^ s and
 * c_1 = a_1&b_1 ^ (r ^ a_0&b_1) ^ (s ^ b_0&a_1).
 * The OTs are ideal (the receiver gets the chosen message) where the
 * reference runs FerretCOT.
 *
 * Randomness (the reference draws from emp::PRG): dealer domain 14, word
 * M(g, j) = U(seed, 14, g, j) for global instance g.
 *   j = 0 / 1: masks of party 0's / party 1's adder input (bit i = input i)
 *   j = 2 / 3: masks of party 0's / party 1's alpha-share input of DltSft
 *   j = 4,5 / 6,7: masks of party 0's / party 1's Delta-share input
 *                  (lo, hi words; bit j of the block = input j)
 *   AND op o: a0, a1, b0, b1, r, s = M(8+6o+0..5); the gate with slot
 *             (o, bit) uses bit `bit` of each word.
 * Gilboa (circuit.cpp:91-153): party 0 holds x, party 1 holds y < P; for
 * bit j < 62 of y party 0 offers (r_j, r_j + x 2^j mod P) with
 * r_j = U(seed, 15, g, 64q + j) % P for product q of instance g; party 0's
 * share is P - sum r_j, party 1's the sum of the messages it chose.
 * Products: q = 0 beta (x = pay0, y = pay1); zp_triple (circuit.cpp:
 * 156-200) with the dealer's r_s, s_s: q = 1 r0 * s1, q = 2 s0 * r1, and
 * rs_s = r_s s_s + (share s of q = 1) + (share s of q = 2).
 *   Ops: adder op 0 = the generate ANDs G_i = a_i & b_i (bit i); prefix
 *        step s = 1.. (distance d = 2^(s-1) < nb): op 2s-1 = P_i & G_{i-d},
 *        op 2s = P_i & P_{i-d} (bit i, i >= d).  DltSft (level i, child k,
 *        e = i*nk + k): base op = adder ops + e*(w+1); its AND-tree gates in
 *        creation order take ops base + j (bit 0), j < w-1; the Delta ANDs
 *        take op base + w-1 (bits 0..63) and base + w (bits 64..127).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

enum { G_IN, G_XOR, G_NOT, G_AND };

typedef struct {
    int type, in0, in1, owner, op, bit;
} gate;

typedef struct {
    gate* g;
    int n, cap;
    int* outs;
    int nout, outcap;
} circuit;

static int add(circuit* c, int type, int in0, int in1, int owner, int op, int bit) {
    if (c->n == c->cap) {
        c->cap = c->cap ? 2 * c->cap : 256;
        c->g = (gate*)realloc(c->g, (size_t)c->cap * sizeof(gate));
    }
    gate* x = &c->g[c->n];
    x->type = type; x->in0 = in0; x->in1 = in1; x->owner = owner; x->op = op; x->bit = bit;
    return c->n++;
}
static int c_in(circuit* c, int owner) { return add(c, G_IN, -1, -1, owner, 0, 0); }
static int c_xor(circuit* c, int a, int b) { return add(c, G_XOR, a, b, -1, 0, 0); }
static int c_not(circuit* c, int a) { return add(c, G_NOT, a, -1, -1, 0, 0); }
static int c_and(circuit* c, int a, int b, int op, int bit) { return add(c, G_AND, a, b, -1, op, bit); }
static void c_out(circuit* c, int w) {
    if (c->nout == c->outcap) {
        c->outcap = c->outcap ? 2 * c->outcap : 256;
        c->outs = (int*)realloc(c->outs, (size_t)c->outcap * sizeof(int));
    }
    c->outs[c->nout++] = w;
}
static void c_free(circuit* c) { free(c->g); free(c->outs); memset(c, 0, sizeof(*c)); }

static uint64_t M(uint64_t seed, uint64_t g, uint64_t j) {
    return orc_dealer_u64(seed, 14, g, j);
}

/* Evaluate both parties' shares gate by gate.  in_bits[o] lists owner o's
 * input bits in add order, in_masks[o] the matching mask bits. */
static void evaluate(const circuit* c, uint64_t seed, uint64_t g, const uint8_t* const in_bits[2],
                     const uint8_t* const in_masks[2], uint8_t* s0, uint8_t* s1) {
    int next[2] = {0, 0};
    for (int i = 0; i < c->n; ++i) {
        const gate* x = &c->g[i];
        switch (x->type) {
        case G_IN: {
            const int o = x->owner, k = next[o]++;
            const uint8_t v = in_bits[o][k], m = in_masks[o][k];
            s0[i] = o == 0 ? (uint8_t)(v ^ m) : m;
            s1[i] = o == 1 ? (uint8_t)(v ^ m) : m;
            break;
        }
        case G_XOR:
            s0[i] = s0[x->in0] ^ s0[x->in1];
            s1[i] = s1[x->in0] ^ s1[x->in1];
            break;
        case G_NOT:
            s0[i] = 1 ^ s0[x->in0];
            s1[i] = s1[x->in0];
            break;
        default: {
            const uint64_t base = 8 + 6 * (uint64_t)x->op;
            const int b = x->bit;
            const uint8_t a0 = (M(seed, g, base) >> b) & 1, a1 = (M(seed, g, base + 1) >> b) & 1;
            const uint8_t b0 = (M(seed, g, base + 2) >> b) & 1, b1 = (M(seed, g, base + 3) >> b) & 1;
            const uint8_t r = (M(seed, g, base + 4) >> b) & 1, sr = (M(seed, g, base + 5) >> b) & 1;
            /* the two OTs: party 1 receives r ^ a0&b1 (choice b1), sr ^ b0&a1 (choice a1) */
            const uint8_t ot1 = (uint8_t)(b1 ? r ^ a0 : r), ot2 = (uint8_t)(a1 ? sr ^ b0 : sr);
            const uint8_t c0 = (uint8_t)((a0 & b0) ^ r ^ sr), c1 = (uint8_t)((a1 & b1) ^ ot1 ^ ot2);
            const uint8_t d = s0[x->in0] ^ a0 ^ s1[x->in0] ^ a1;
            const uint8_t e = s0[x->in1] ^ b0 ^ s1[x->in1] ^ b1;
            s0[i] = (uint8_t)((d & e) ^ (d & b0) ^ (e & a0) ^ c0);
            s1[i] = (uint8_t)((d & b1) ^ (e & a1) ^ c1);
            break;
        }
        }
    }
}

static int steps_of(int nb) {
    int s = 0;
    for (int d = 1; d < nb; d <<= 1) ++s;
    return s;
}

/* The Kogge-Stone adder (adder.h:21-63) for one instance; fills the sum
 * wires (nb + 1, LSB first). */
static void build_adder(circuit* c, int nb, int* sum) {
    int* a = (int*)malloc(sizeof(int) * (size_t)nb);
    int* b = (int*)malloc(sizeof(int) * (size_t)nb);
    int* G = (int*)malloc(sizeof(int) * (size_t)nb);
    int* P = (int*)malloc(sizeof(int) * (size_t)nb);
    int* P0 = (int*)malloc(sizeof(int) * (size_t)nb);
    int* Gn = (int*)malloc(sizeof(int) * (size_t)nb);
    int* Pn = (int*)malloc(sizeof(int) * (size_t)nb);
    for (int i = 0; i < nb; ++i) a[i] = c_in(c, 0);
    for (int i = 0; i < nb; ++i) b[i] = c_in(c, 1);
    for (int i = 0; i < nb; ++i) {
        G[i] = c_and(c, a[i], b[i], 0, i);
        P[i] = c_xor(c, a[i], b[i]);
        P0[i] = P[i];
    }
    int s = 1;
    for (int d = 1; d < nb; d <<= 1, ++s) {
        memcpy(Gn, G, sizeof(int) * (size_t)nb);
        memcpy(Pn, P, sizeof(int) * (size_t)nb);
        for (int i = d; i < nb; ++i) {
            const int pg = c_and(c, P[i], G[i - d], 2 * s - 1, i);
            Gn[i] = c_xor(c, G[i], pg);
            Pn[i] = c_and(c, P[i], P[i - d], 2 * s, i);
        }
        memcpy(G, Gn, sizeof(int) * (size_t)nb);
        memcpy(P, Pn, sizeof(int) * (size_t)nb);
    }
    sum[0] = P0[0];
    for (int i = 1; i < nb; ++i) sum[i] = c_xor(c, P0[i], G[i - 1]);
    sum[nb] = G[nb - 1];
    free(a); free(b); free(G); free(P); free(P0); free(Gn); free(Pn);
}

/* The DltSft circuit (dpf.cpp:172-240) for one instance. */
static void build_dltsft(circuit* c, int n, int w, int op0) {
    const int m = 1 << w, nk = m - 1, levels = n / w;
    int* aa = (int*)malloc(sizeof(int) * (size_t)n);
    int* ba = (int*)malloc(sizeof(int) * (size_t)n);
    int ad[128], bd[128], dx[128];
    int* ax = (int*)malloc(sizeof(int) * (size_t)n);
    for (int j = 0; j < n; ++j) aa[j] = c_in(c, 0);
    for (int j = 0; j < n; ++j) ba[j] = c_in(c, 1);
    for (int j = 0; j < 128; ++j) ad[j] = c_in(c, 0);
    for (int j = 0; j < 128; ++j) bd[j] = c_in(c, 1);
    for (int j = 0; j < n; ++j) ax[j] = c_xor(c, aa[j], ba[j]);
    for (int j = 0; j < 128; ++j) dx[j] = c_xor(c, ad[j], bd[j]);
    int f[8], nf;
    for (int i = 0; i < levels; ++i)
        for (int k = 0; k < nk; ++k) {
            const int base = op0 + (i * nk + k) * (w + 1);
            int tree = 0;
            for (int cc = 0; cc < w; ++cc) {
                const int pos = n - w * (i + 1) + cc;
                f[cc] = ((k >> cc) & 1) ? ax[pos] : c_not(c, ax[pos]);
            }
            nf = w;
            while (nf > 1) {
                int nn = 0;
                for (int j = 0; j + 1 < nf; j += 2) f[nn++] = c_and(c, f[j], f[j + 1], base + tree++, 0);
                if (nf & 1) f[nn++] = f[nf - 1];
                nf = nn;
            }
            for (int j = 0; j < 128; ++j)
                c_out(c, c_and(c, f[0], dx[j], base + w - 1 + (j >> 6), j & 63));
        }
    free(aa); free(ba); free(ax);
}

int orc_gmw_adder_ops(int nb) { return 1 + 2 * steps_of(nb); }

/* Gilboa's product for instance g, product q: both parties' shares. */
void orc_gilboa(uint64_t x, uint64_t y, uint64_t seed, uint64_t g, uint32_t q, uint64_t* sh0,
                uint64_t* sh1) {
    int nbits = 0;
    for (uint64_t v = ORC_P - 1; v; v >>= 1) ++nbits;
    uint64_t rsum = 0, got = 0, xs = x % ORC_P;
    for (int j = 0; j < nbits; ++j) {
        const uint64_t r = orc_dealer_u64(seed, 15, g, 64 * (uint64_t)q + (uint64_t)j) % ORC_P;
        const uint64_t m0 = r, m1 = orc_fadd(r, xs);   /* the sender's two messages */
        rsum = orc_fadd(rsum, r);
        got = orc_fadd(got, ((y % ORC_P) >> j) & 1 ? m1 : m0);
        xs = orc_fadd(xs, xs);
    }
    *sh0 = orc_fsub(0, rsum);
    *sh1 = got;
}

int orc_gmw_preprocess(const orc_params* p, uint64_t seed, orc_party_inputs* in0,
                       orc_party_inputs* in1, uint64_t* alpha_sh0, uint64_t* alpha_sh1) {
    if (orc_dealer(p, seed, in0, in1, NULL)) return -1;
    const int N = p->N, cc = p->c, t = p->t, B = t * t, w = p->w;
    const uint64_t bin = (uint64_t)(N / t);
    int nb = 0;
    while ((1ULL << nb) < bin) ++nb;
    const int n = nb + 1, m = 1 << w, nk = m - 1, levels = n / w;
    const int aops = orc_gmw_adder_ops(nb);
    const orc_blk W = orc_public_W();

    circuit add_c = {0}, dlt_c = {0};
    int* sum = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
    build_adder(&add_c, nb, sum);
    build_dltsft(&dlt_c, n, w, aops);
    uint8_t* s0 = (uint8_t*)malloc((size_t)(add_c.n > dlt_c.n ? add_c.n : dlt_c.n));
    uint8_t* s1 = (uint8_t*)malloc((size_t)(add_c.n > dlt_c.n ? add_c.n : dlt_c.n));
    uint8_t bits0[64 + 128], bits1[64 + 128], msk0[64 + 128], msk1[64 + 128];

    for (int ii = 0; ii < cc; ++ii)
        for (int jj = 0; jj < cc; ++jj)
            for (int idx = 0; idx < B; ++idx) {
                const int kk = idx / t, ll = idx % t;
                const uint64_t g = (uint64_t)(ii * cc + jj) * B + idx;
                /* 3a: the position adder */
                const uint64_t off0 = in0->pos[ii * t + kk] % bin, off1 = in1->pos[jj * t + ll] % bin;
                const uint64_t m0 = M(seed, g, 0), m1 = M(seed, g, 1);
                for (int i = 0; i < nb; ++i) {
                    bits0[i] = (off0 >> i) & 1; msk0[i] = (m0 >> i) & 1;
                    bits1[i] = (off1 >> i) & 1; msk1[i] = (m1 >> i) & 1;
                }
                const uint8_t* ib[2] = {bits0, bits1};
                const uint8_t* im[2] = {msk0, msk1};
                evaluate(&add_c, seed, g, ib, im, s0, s1);
                uint64_t a0 = 0, a1 = 0;
                for (int i = 0; i <= nb; ++i) {
                    a0 |= (uint64_t)s0[sum[i]] << i;
                    a1 |= (uint64_t)s1[sum[i]] << i;
                }
                if (alpha_sh0) alpha_sh0[g] = a0;
                if (alpha_sh1) alpha_sh1[g] = a1;
                /* 3c: DltSft from the alpha shares and the Delta shares */
                orc_blk d0 = {in0->root[g].lo ^ W.lo, in0->root[g].hi ^ W.hi};
                orc_blk d1 = {in1->root[g].lo ^ W.lo, in1->root[g].hi ^ W.hi};
                const uint64_t m2 = M(seed, g, 2), m3 = M(seed, g, 3);
                const uint64_t md0[2] = {M(seed, g, 4), M(seed, g, 5)};
                const uint64_t md1[2] = {M(seed, g, 6), M(seed, g, 7)};
                for (int j = 0; j < n; ++j) {
                    bits0[j] = (a0 >> j) & 1; msk0[j] = (m2 >> j) & 1;
                    bits1[j] = (a1 >> j) & 1; msk1[j] = (m3 >> j) & 1;
                }
                for (int j = 0; j < 128; ++j) {
                    const uint64_t v0 = j < 64 ? d0.lo : d0.hi, v1 = j < 64 ? d1.lo : d1.hi;
                    bits0[n + j] = (v0 >> (j & 63)) & 1; msk0[n + j] = (md0[j >> 6] >> (j & 63)) & 1;
                    bits1[n + j] = (v1 >> (j & 63)) & 1; msk1[n + j] = (md1[j >> 6] >> (j & 63)) & 1;
                }
                evaluate(&dlt_c, seed, g, ib, im, s0, s1);
                int o = 0;
                for (int i = 0; i < levels; ++i)
                    for (int k = 0; k < nk; ++k) {
                        orc_blk x0 = {0, 0}, x1 = {0, 0};
                        for (int j = 0; j < 128; ++j, ++o) {
                            const int wi = dlt_c.outs[o];
                            if (j < 64) {
                                x0.lo |= (uint64_t)s0[wi] << j; x1.lo |= (uint64_t)s1[wi] << j;
                            } else {
                                x0.hi |= (uint64_t)s0[wi] << (j - 64); x1.hi |= (uint64_t)s1[wi] << (j - 64);
                            }
                        }
                        const size_t off = ((size_t)g * levels + i) * nk + k;
                        in0->dltsft[off] = x0;
                        in1->dltsft[off] = x1;
                    }
                /* 3b: beta and the Beaver triple by Gilboa products */
                orc_gilboa(in0->pay[ii * t + kk], in1->pay[jj * t + ll], seed, g, 0, &in0->beta[g],
                           &in1->beta[g]);
                uint64_t c10, c11, c20, c21;
                orc_gilboa(in0->r[g], in1->s[g], seed, g, 1, &c10, &c11);
                orc_gilboa(in0->s[g], in1->r[g], seed, g, 2, &c20, &c21);
                in0->rs[g] = orc_fadd(orc_fadd(orc_fmul(in0->r[g], in0->s[g]), c10), c20);
                in1->rs[g] = orc_fadd(orc_fadd(orc_fmul(in1->r[g], in1->s[g]), c11), c21);
            }
    free(s0); free(s1); free(sum);
    c_free(&add_c); c_free(&dlt_c);
    return 0;
}

int orc_gmw_and_gates(int nb, int w, int n) {
    circuit a = {0}, d = {0};
    int* sum = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
    build_adder(&a, nb, sum);
    build_dltsft(&d, n, w, orc_gmw_adder_ops(nb));
    int cnt = 0;
    for (int i = 0; i < a.n; ++i) cnt += a.g[i].type == G_AND;
    for (int i = 0; i < d.n; ++i) cnt += d.g[i].type == G_AND;
    free(sum);
    c_free(&a);
    c_free(&d);
    return cnt;
}
