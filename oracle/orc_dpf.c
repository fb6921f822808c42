/*
 * Half-tree m-ary DPF: CPU restatement (test infrastructure only; see
 * oracle.h).
 *
 * batch_gen_full_eval follows half_tree_dpf/dpf.cpp:485-578:
 *   root = delta_share ^ W;  per level i (frontier size sz = m^i):
 *     h[j][k]  = H_k(parent_j)                       (tree PRG)
 *     hx[j]    = XOR_k h[j][k]
 *     my_k     = XOR_j h[j][k] ^ DltSft[i][k]        (this party's CW share)
 *     exchange my <-> peer (one round for all B*(m-1) shares)
 *     CW_k     = my_k ^ peer_k,  CW_{m-1} = XOR_k CW_k
 *     child[j*m+k]   = h[j][k] ^ (t ? CW_k : 0),        k < m-1
 *     child[j*m+m-1] = hx[j] ^ parent_j ^ (t ? CW_{m-1} : 0),
 *   with t = lsb(parent_j) = parent.lo & 1.
 * eval follows dpf.cpp:302-327 (digits MSB first).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void tree_hash(int prg, const uint8_t* rks, orc_blk* out,
                      const orc_blk* parents, size_t n, int nk) {
    if (prg == ORC_PRG_AES)
        orc_tree_hash_aes(rks, out, parents, n, nk);
    else
        orc_tree_hash_chacha8(out, parents, n, nk);
}

static orc_blk bx(orc_blk a, orc_blk b) {
    orc_blk r = {a.lo ^ b.lo, a.hi ^ b.hi};
    return r;
}

typedef struct {
    int B, n, w, m, nk, levels, prg;
    size_t D;
    const uint8_t* rks;
    const orc_blk* dlt;
    orc_blk *cur, *nxt, *h, *hx, *my;
} party_state;

static int party_init(party_state* s, int B, int n, int w, int prg,
                      const uint8_t* rks, const orc_blk* roots,
                      const orc_blk* dlt) {
    if (w < 1 || w > 4 || n % w != 0 || B < 0) return -1;
    if (prg == ORC_PRG_AES && rks == NULL) return -1;
    memset(s, 0, sizeof(*s));
    s->B = B; s->n = n; s->w = w; s->m = 1 << w; s->nk = s->m - 1;
    s->levels = n / w; s->prg = prg; s->rks = rks; s->dlt = dlt;
    s->D = (size_t)1 << n;
    size_t maxsz = s->D / (size_t)s->m;
    s->cur = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * s->D + 16);
    s->nxt = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * s->D + 16);
    s->h = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * maxsz * s->nk + 16);
    s->hx = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * maxsz + 16);
    s->my = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * s->nk + 16);
    for (int b = 0; b < B; ++b) s->cur[(size_t)b * s->D] = roots[b];
    return 0;
}

static void party_free(party_state* s) {
    free(s->cur); free(s->nxt); free(s->h); free(s->hx); free(s->my);
}

/* dpf.cpp:507-532: hash all parents, XOR columns into this party's shares. */
static void party_hash_level(party_state* s, int i) {
    const size_t sz = (size_t)1 << (s->w * i);
    const size_t maxsz = s->D / (size_t)s->m;
    for (int b = 0; b < s->B; ++b) {
        orc_blk* h = s->h + (size_t)b * maxsz * s->nk;
        orc_blk* hx = s->hx + (size_t)b * maxsz;
        const orc_blk* cur = s->cur + (size_t)b * s->D;
        tree_hash(s->prg, s->rks, h, cur, sz, s->nk);
        for (size_t j = 0; j < sz; ++j) {
            orc_blk x = {0, 0};
            for (int k = 0; k < s->nk; ++k) x = bx(x, h[j * s->nk + k]);
            hx[j] = x;
        }
        for (int k = 0; k < s->nk; ++k) {
            orc_blk cw = {0, 0};
            for (size_t j = 0; j < sz; ++j) cw = bx(cw, h[j * s->nk + k]);
            cw = bx(cw, s->dlt[((size_t)b * s->levels + i) * s->nk + k]);
            s->my[(size_t)b * s->nk + k] = cw;
        }
    }
}

/* dpf.cpp:553-574: rebuild CW from both shares and expand every parent. */
static void party_expand_level(party_state* s, int i, const orc_blk* peer,
                               orc_blk* cw_out) {
    const size_t sz = (size_t)1 << (s->w * i);
    const size_t maxsz = s->D / (size_t)s->m;
    for (int b = 0; b < s->B; ++b) {
        orc_blk CW[16];
        orc_blk fold = {0, 0};
        for (int k = 0; k < s->nk; ++k) {
            CW[k] = bx(s->my[(size_t)b * s->nk + k], peer[(size_t)b * s->nk + k]);
            fold = bx(fold, CW[k]);
        }
        CW[s->m - 1] = fold;
        if (cw_out)
            for (int k = 0; k < s->m; ++k)
                cw_out[((size_t)b * s->levels + i) * s->m + k] = CW[k];
        const orc_blk* h = s->h + (size_t)b * maxsz * s->nk;
        const orc_blk* hx = s->hx + (size_t)b * maxsz;
        const orc_blk* cur = s->cur + (size_t)b * s->D;
        orc_blk* nxt = s->nxt + (size_t)b * s->D;
        for (size_t j = 0; j < sz; ++j) {
            int t = (int)(cur[j].lo & 1);
            for (int k = 0; k < s->nk; ++k) {
                orc_blk c = h[j * s->nk + k];
                if (t) c = bx(c, CW[k]);
                nxt[j * s->m + k] = c;
            }
            orc_blk last = bx(hx[j], cur[j]);
            if (t) last = bx(last, CW[s->m - 1]);
            nxt[j * s->m + s->m - 1] = last;
        }
    }
    orc_blk* tmp = s->cur; s->cur = s->nxt; s->nxt = tmp;
}

int orc_dpf_batch_full_eval_party(int B, int n, int w, int prg,
                                  const uint8_t* rks, const orc_blk* roots,
                                  const orc_blk* dltsft, orc_exchange_fn ex,
                                  void* ctx, orc_blk* leaves) {
    party_state s;
    if (party_init(&s, B, n, w, prg, rks, roots, dltsft)) return -1;
    orc_blk* peer = (orc_blk*)malloc(sizeof(orc_blk) * (size_t)B * s.nk + 16);
    for (int i = 0; i < s.levels; ++i) {
        party_hash_level(&s, i);
        ex(ctx, i, s.my, peer, (size_t)B * s.nk);
        party_expand_level(&s, i, peer, NULL);
    }
    memcpy(leaves, s.cur, sizeof(orc_blk) * (size_t)B * s.D);
    free(peer);
    party_free(&s);
    return 0;
}

int orc_dpf_batch_full_eval_both(int B, int n, int w, int prg,
                                 const uint8_t* rks, const orc_blk* roots0,
                                 const orc_blk* roots1, const orc_blk* dlt0,
                                 const orc_blk* dlt1, orc_blk* leaves0,
                                 orc_blk* leaves1, orc_blk* cw_out) {
    party_state s0, s1;
    if (party_init(&s0, B, n, w, prg, rks, roots0, dlt0)) return -1;
    if (party_init(&s1, B, n, w, prg, rks, roots1, dlt1)) { party_free(&s0); return -1; }
    for (int i = 0; i < s0.levels; ++i) {
        party_hash_level(&s0, i);
        party_hash_level(&s1, i);
        /* in-process exchange: each party receives the other's shares */
        party_expand_level(&s0, i, s1.my, cw_out);
        party_expand_level(&s1, i, s0.my, NULL);
    }
    memcpy(leaves0, s0.cur, sizeof(orc_blk) * (size_t)B * s0.D);
    memcpy(leaves1, s1.cur, sizeof(orc_blk) * (size_t)B * s1.D);
    party_free(&s0);
    party_free(&s1);
    return 0;
}

/* dpf.cpp:302-327 */
orc_blk orc_dpf_eval(int n, int w, int prg, const uint8_t* rks, orc_blk root,
                     const orc_blk* cw, uint64_t j) {
    const int m = 1 << w, levels = n / w, nk = m - 1;
    orc_blk seed = root, h[16];
    for (int i = 0; i < levels; ++i) {
        int t = (int)(seed.lo & 1);
        uint64_t digit = (j >> (n - w * (i + 1))) & (uint64_t)(m - 1);
        tree_hash(prg, rks, h, &seed, 1, nk);
        orc_blk hx = {0, 0};
        for (int k = 0; k < nk; ++k) hx = bx(hx, h[k]);
        if ((int)digit < m - 1) {
            seed = h[digit];
            if (t) seed = bx(seed, cw[(size_t)i * m + digit]);
        } else {
            seed = bx(hx, seed);
            if (t) seed = bx(seed, cw[(size_t)i * m + m - 1]);
        }
    }
    return seed;
}

int orc_dpf_digit(int n, int w, int level, uint64_t alpha) {
    const int m = 1 << w;
    return (int)((alpha >> (n - w * (level + 1))) & (uint64_t)(m - 1));
}

void orc_dpf_path_cw(int n, int w, int prg, const uint8_t* rks, orc_blk root0,
                     orc_blk root1, const orc_blk* dlt0, const orc_blk* dlt1,
                     uint64_t alpha, orc_blk* cw_out) {
    const int m = 1 << w, levels = n / w, nk = m - 1;
    orc_blk s[2] = {root0, root1}, h[2][16];
    for (int i = 0; i < levels; ++i) {
        tree_hash(prg, rks, h[0], &s[0], 1, nk);
        tree_hash(prg, rks, h[1], &s[1], 1, nk);
        orc_blk* CW = cw_out + (size_t)i * m;
        orc_blk fold = {0, 0};
        for (int k = 0; k < nk; ++k) {
            CW[k] = bx(bx(h[0][k], h[1][k]),
                       bx(dlt0[(size_t)i * nk + k], dlt1[(size_t)i * nk + k]));
            fold = bx(fold, CW[k]);
        }
        CW[m - 1] = fold;
        int digit = orc_dpf_digit(n, w, i, alpha);
        for (int p = 0; p < 2; ++p) {
            int t = (int)(s[p].lo & 1);
            orc_blk nx;
            if (digit < m - 1) {
                nx = h[p][digit];
            } else {
                orc_blk hx = {0, 0};
                for (int k = 0; k < nk; ++k) hx = bx(hx, h[p][k]);
                nx = bx(hx, s[p]);
            }
            if (t) nx = bx(nx, CW[digit]);
            s[p] = nx;
        }
    }
}

/* ---- dealer Gen + sync-free full-domain eval (batched DPF keys) --------- */
/* Gen from a dealer who knows alpha and both seeds (w = 1).  Roots follow the
 * fused protocol's setup, root_s = Delta_s ^ W (dpf.cpp:338): Delta_0 = d0,
 * Delta_1 = d0 ^ delta.  Each level's CW is the path-only value
 * H(s0) ^ H(s1) ^ [bit_i(alpha) == 0] * delta -- the value the fused exchange
 * of dpf.cpp:358-385 reconstructs (off-path nodes cancel) -- and the path
 * advances with the children of dpf.cpp:391-397.  The output word makes the
 * converted leaves y_s(j) = (-1)^s (C_s(j) + tau_s(j) * cw_out) (the
 * conversion of pcg_ole_impl.h:432-446) sum to beta * [j == alpha]:
 * cw_out = (beta - C_0(alpha) + C_1(alpha)) * (tau_0(alpha) - tau_1(alpha)). */
void orc_dpf_key_gen(int n, int prg, const uint8_t* rks, uint64_t alpha, uint64_t beta,
                     orc_blk delta, orc_blk d0, orc_blk W, orc_blk* root0, orc_blk* root1,
                     orc_blk* cw, uint64_t* cw_out) {
    orc_blk s[2], h[2];
    s[0] = bx(d0, W);
    s[1] = bx(bx(d0, delta), W);
    *root0 = s[0];
    *root1 = s[1];
    for (int i = 0; i < n; ++i) {
        tree_hash(prg, rks, &h[0], &s[0], 1, 1);
        tree_hash(prg, rks, &h[1], &s[1], 1, 1);
        const int bit = (int)((alpha >> (n - 1 - i)) & 1);
        orc_blk c = bx(h[0], h[1]);
        if (bit == 0) c = bx(c, delta);
        cw[i] = c;
        for (int p = 0; p < 2; ++p) {
            const int t = (int)(s[p].lo & 1);
            orc_blk nx = bit ? bx(h[p], s[p]) : h[p];
            if (t) nx = bx(nx, c);
            s[p] = nx;
        }
    }
    const uint64_t C0 = orc_fmod(orc_out_hash_lo64(s[0]));
    const uint64_t C1 = orc_fmod(orc_out_hash_lo64(s[1]));
    const uint64_t v = orc_fsub(orc_fadd(orc_fmod(beta), C1), C0);
    *cw_out = (s[0].lo & 1) ? v : orc_fneg(v);
}

/* Party s's full domain from its key (w = 1): every leaf as DPF.Eval of
 * dpf.cpp:302-327 would compute it (level by level, CW[i] known), then
 * y(j) = (-1)^s (C(j) + tau(j) * cw_out) with C = lo64(H'(leaf)) mod P. */
void orc_dpf_key_eval_y(int n, int prg, const uint8_t* rks, orc_blk root, const orc_blk* cw,
                        uint64_t cw_out, int party, uint64_t* y) {
    const size_t D = (size_t)1 << n;
    orc_blk* cur = (orc_blk*)malloc(D * sizeof(orc_blk));
    orc_blk* nxt = (orc_blk*)malloc(D * sizeof(orc_blk));
    cur[0] = root;
    for (int i = 0; i < n; ++i) {
        const size_t sz = (size_t)1 << i;
        for (size_t j = 0; j < sz; ++j) {
            orc_blk h;
            tree_hash(prg, rks, &h, &cur[j], 1, 1);
            orc_blk c0 = h;
            if (cur[j].lo & 1) c0 = bx(c0, cw[i]);
            nxt[2 * j] = c0;
            nxt[2 * j + 1] = bx(c0, cur[j]);
        }
        orc_blk* t = cur;
        cur = nxt;
        nxt = t;
    }
    for (size_t j = 0; j < D; ++j) {
        uint64_t v = orc_fmod(orc_out_hash_lo64(cur[j]));
        if (cur[j].lo & 1) v = orc_fadd(v, cw_out);
        y[j] = party ? orc_fneg(v) : v;
    }
    free(cur);
    free(nxt);
}
