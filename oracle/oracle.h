/*
 * oracle.h -- CPU restatement of the reference PCG-OLE expansion path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code.  Only tests/, __graft_entry__.smoke() (as the checker) and
 * bench.py's cpu_baseline leg may use it.
 *
 * Every function restates a reference routine and cites it (paths relative
 * to the reference root).  Arithmetic is kept in the reference's own form
 * (u128 '%' for the field, per-node ChaCha8 for the tree PRG) so that the
 * oracle is a faithful check, not a fast path.
 *
 * Pinning (what independently checks each function):
 *   field / NTT / poly-mul : reference checksums of the naive and 4-step
 *                            poly-mul recipes at N=2^20 (8599dd17bf241164,
 *                            3969340b039f5d5d), schoolbook product, the
 *                            compiled reference ntt.h (oracle/_ref).
 *   ChaCha8                : published zero-key ChaCha8 keystream, compiled
 *                            reference chacha8.h.
 *   AES-128                : FIPS-197 Appendix B/C known answers.
 *   tree PRG / DPF / output: reference DPF+convert checksums (plateau recipe),
 *                            point-function invariant, brute force of eval()
 *                            on tiny n, fused-CW == path-only CW.
 *   PCG-OLE                : x0*x1 == z0+z1 over R_p.
 *   Dealer (seeded 2PC stand-in) and the public W constant: parity unpinned
 *   against emp (emp-tool is not vendored); they are a documented spec.
 */
#ifndef PCGSQ_ORACLE_H
#define PCGSQ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* P = 2^62 - 3*2^25 + 1 (GPU_baseline/test/out_hash_cuda_bitexact.cu:27,
 * CPU_baseline/test/pcg_ole_2pc.cpp:18). */
#define ORC_P 4611686018326724609ULL
/* ChaCha8OutHash::OUT_DOMAIN ("out_hash"), half_tree_dpf/tree_prg.h:132 */
#define ORC_OUT_DOMAIN 0x6f75745f68617368ULL

/* 128-bit block: lo = bytes 0..7, hi = bytes 8..15 (little endian), the
 * in-memory layout of emp::block (GPU_baseline/common/dpf_gpu.h:19-26). */
typedef struct {
    uint64_t lo, hi;
} orc_blk;

enum { ORC_PRG_AES = 0, ORC_PRG_CHACHA8 = 1 };

/* ---- field FFp<P> (CPU_baseline/ffp.h) ---------------------------------- */
uint64_t orc_fmod(uint64_t x);
uint64_t orc_fadd(uint64_t a, uint64_t b);
uint64_t orc_fsub(uint64_t a, uint64_t b);
uint64_t orc_fneg(uint64_t a);
uint64_t orc_fmul(uint64_t a, uint64_t b);
uint64_t orc_fpow(uint64_t b, uint64_t e);
uint64_t orc_finv(uint64_t a);

/* ---- NTT<P> (CPU_baseline/ntt.h) --------------------------------------- */
typedef struct orc_ntt orc_ntt;
orc_ntt* orc_ntt_new(int N);               /* NULL if N invalid for P */
void orc_ntt_free(orc_ntt* t);
int orc_ntt_size(const orc_ntt* t);
uint64_t orc_ntt_psi(const orc_ntt* t);
void orc_ntt_forward(const orc_ntt* t, uint64_t* a);
void orc_ntt_inverse(const orc_ntt* t, uint64_t* a);
void orc_ntt_poly_mul(const orc_ntt* t, uint64_t* c, const uint64_t* a,
                      const uint64_t* b);
uint64_t orc_primitive_root(void);
/* pcg_ole.h:92-106 */
void orc_poly_mul_schoolbook(uint64_t* c, const uint64_t* a,
                             const uint64_t* b, int N);
/* pcg_ole.h:119-132 (NTT when (P-1) % 2N == 0, else schoolbook) */
void orc_poly_mul(uint64_t* c, const uint64_t* a, const uint64_t* b, int N);
/* batch poly-mul, OpenMP over items when threads > 1 */
void orc_poly_mul_batch(uint64_t* c, const uint64_t* a, const uint64_t* b,
                        int batch, int N, int threads);

/* ---- ChaCha8 (CPU_baseline/chacha8.h) ---------------------------------- */
void orc_chacha8_block(const uint8_t key[32], uint64_t nonce, uint64_t ctr,
                       uint8_t out[64]);
/* ChaCha8TreePRG::hash (half_tree_dpf/tree_prg.h:86-104):
 * out[j*nk + k] = 16-byte chunk k of ChaCha8(key = parent_j || 0^128). */
void orc_tree_hash_chacha8(orc_blk* out, const orc_blk* parents, size_t n,
                           int nk);
/* ChaCha8OutHash::hash (half_tree_dpf/tree_prg.h:134-147) */
void orc_out_hash(orc_blk* out, const orc_blk* leaves, size_t n);
uint64_t orc_out_hash_lo64(orc_blk leaf);

/* ---- AES-128 (FIPS-197) + CCR tree PRG (mkey_prp.h, tree_prg.h:42-68) --- */
void orc_aes128_expand_key(const uint8_t key[16], uint8_t rk[176]);
void orc_aes128_encrypt(const uint8_t rk[176], const uint8_t in[16],
                        uint8_t out[16]);
/* sub_k = AES_master(makeBlock(0,k)) ^ makeBlock(0,k), expanded to round keys;
 * rks has nk*176 bytes (tree_prg.h:45-56, dpf.cpp:599-619). */
void orc_aes_tree_round_keys(const uint8_t master[16], int nk, uint8_t* rks);
/* H_k(x) = AES_{sub_k}(x) ^ x ; out[j*nk + k] */
void orc_tree_hash_aes(const uint8_t* rks, orc_blk* out,
                       const orc_blk* parents, size_t n, int nk);

/* ---- half-tree DPF (half_tree_dpf/dpf.cpp) ----------------------------- */
/* Per-level exchange: fill peer[0..count) given mine[0..count). */
typedef void (*orc_exchange_fn)(void* ctx, int level, const orc_blk* mine,
                                orc_blk* peer, size_t count);

/* One party's batch_gen_full_eval (dpf.cpp:485-578) with a host exchange.
 * roots[B]; dltsft[B*levels*nk] indexed [b][i][k]; leaves[B * 2^n]. */
int orc_dpf_batch_full_eval_party(int B, int n, int w, int prg,
                                  const uint8_t* rks, const orc_blk* roots,
                                  const orc_blk* dltsft, orc_exchange_fn ex,
                                  void* ctx, orc_blk* leaves);

/* Both parties in one process: the exchange is my_0 ^ my_1 per level.
 * Same arguments as above, for party 0 and party 1.  If cw_out is not NULL it
 * receives the reconstructed correction words [b][i][k], k < m (k = m-1 is the
 * fold word). */
int orc_dpf_batch_full_eval_both(int B, int n, int w, int prg,
                                 const uint8_t* rks, const orc_blk* roots0,
                                 const orc_blk* roots1, const orc_blk* dlt0,
                                 const orc_blk* dlt1, orc_blk* leaves0,
                                 orc_blk* leaves1, orc_blk* cw_out);

/* DPF.Eval of a single point (dpf.cpp:302-327); cw[levels*m]. */
orc_blk orc_dpf_eval(int n, int w, int prg, const uint8_t* rks, orc_blk root,
                     const orc_blk* cw, uint64_t j);

/* Path-only correction words: CW_{i,k} = H_k(s0) ^ H_k(s1) ^ dlt0 ^ dlt1
 * along the path alpha (no frontier).  Used to check that it equals the fused
 * protocol's CWs (off-path nodes are equal in both trees). */
void orc_dpf_path_cw(int n, int w, int prg, const uint8_t* rks, orc_blk root0,
                     orc_blk root1, const orc_blk* dlt0, const orc_blk* dlt1,
                     uint64_t alpha, orc_blk* cw_out);

/* Dealer Gen of a w = 1 key pair from alpha, beta, Delta (lsb 1), Delta_0
 * share and public W: roots, path-only CW[n], output word cw_out (see
 * orc_dpf.c).  Eval of party s's whole domain into y[2^n] (field shares,
 * y_0 + y_1 = beta * [j == alpha]). */
void orc_dpf_key_gen(int n, int prg, const uint8_t* rks, uint64_t alpha, uint64_t beta,
                     orc_blk delta, orc_blk d0, orc_blk W, orc_blk* root0, orc_blk* root1,
                     orc_blk* cw, uint64_t* cw_out);
void orc_dpf_key_eval_y(int n, int prg, const uint8_t* rks, orc_blk root, const orc_blk* cw,
                        uint64_t cw_out, int party, uint64_t* y);

/* DltSft for a dealer: value_{i,k} = [digit_i(alpha) == k] * Delta
 * (dpf.cpp:104-130, indicator over bits n - w(i+1) + c). */
int orc_dpf_digit(int n, int w, int level, uint64_t alpha);

/* ---- output layer (pcg_ole_impl.h:330-449, leaf_convert_cuda.cu) ------ */
/* Sign-free sums: sumC[b] = sum_d (lo64(H'(leaf)) % P) mod P,
 * sumT[b] = sum_d lsb(leaf) (plain). */
void orc_out_sums(const orc_blk* leaves, int B, size_t D, uint64_t* sumC,
                  uint64_t* sumT);
/* y = C + (tau ? CW : 0); party 1 negates; pos = (b/t + b%t)*(D/2) + d;
 * g_red[pos mod N] += pos < N ? y : -y (mod P); g_red has N entries and is
 * overwritten (leaf_convert_cuda.cu:207-234). */
void orc_out_scatter_g(const orc_blk* leaves, int B, size_t D, int t,
                       int party, const uint64_t* CW, uint64_t* g_red, int N);
/* Host output layer for both parties incl. Beaver opens
 * (pcg_ole_impl.h:330-449) followed by the negacyclic fold (:474-478).
 * beta0/1[B], r0/1, s0/1, rs0/1[B] are the two parties' shares. */
void orc_output_layer_both(const orc_blk* leaves0, const orc_blk* leaves1,
                           int B, size_t D, int t, const uint64_t* beta0,
                           const uint64_t* beta1, const uint64_t* r0,
                           const uint64_t* r1, const uint64_t* s0,
                           const uint64_t* s1, const uint64_t* rs0,
                           const uint64_t* rs1, uint64_t* g_red0,
                           uint64_t* g_red1, int N, uint64_t* cw_out);

/* ---- dealer (seeded stand-in for emp/OpenSSL randomness & 2PC) ---------- */
uint64_t orc_dealer_u64(uint64_t seed, uint32_t domain, uint64_t a,
                        uint64_t b);

typedef struct {
    int N, c, t, w, prg;
} orc_params;

/* Per-expansion dealer output, one party.  Sizes: c*t positions/payloads;
 * per block (c*c) and instance (B = t*t): delta, dltsft[levels*nk], beta,
 * r, s, rs. */
typedef struct {
    uint64_t* pos;      /* [c][t] */
    uint64_t* pay;      /* [c][t] */
    orc_blk* root;      /* [c*c][B]  delta_share ^ W */
    orc_blk* dltsft;    /* [c*c][B][levels][nk] */
    uint64_t* beta;     /* [c*c][B] */
    uint64_t* r;        /* [c*c][B] */
    uint64_t* s;        /* [c*c][B] */
    uint64_t* rs;       /* [c*c][B] */
} orc_party_inputs;

size_t orc_dealer_sizes(const orc_params* p, size_t* n_pos, size_t* n_inst,
                        size_t* n_dlt);
int orc_dealer(const orc_params* p, uint64_t seed, orc_party_inputs* in0,
               orc_party_inputs* in1, uint64_t* alpha /* [c*c][B] or NULL */);
/* The dealer's inputs with DltSft, beta and rs produced instead by the
 * preprocessing of pcg_ole_impl.h:114-199: the Kogge-Stone position adder
 * and the F_DeltaShiftShare circuit of dpf.cpp:172-240 evaluated gate by
 * gate under GMW for both parties, with AND triples made from OTs; beta and
 * the Beaver triples' rs by Gilboa OT products (orc_gmw.c gives the spec).
 * The DltSft, beta and rs shares differ from orc_dealer's; what they
 * reconstruct to, and everything else, is the same.  alpha_sh0/1[c*c][B]
 * (or NULL): the adder's XOR shares of alpha. */
int orc_gmw_preprocess(const orc_params* p, uint64_t seed, orc_party_inputs* in0,
                       orc_party_inputs* in1, uint64_t* alpha_sh0, uint64_t* alpha_sh1);
/* Gilboa's OT product (bool_circuit/circuit.cpp:91-153), product q of
 * instance g: party 0 holds x, party 1 holds y; *sh0 + *sh1 = x*y mod P. */
void orc_gilboa(uint64_t x, uint64_t y, uint64_t seed, uint64_t g, uint32_t q, uint64_t* sh0,
                uint64_t* sh1);
/* AND ops of the nb-bit Kogge-Stone adder (1 + 2*ceil(log2 nb)). */
int orc_gmw_adder_ops(int nb);
/* AND gates of one instance's two circuits (adder of nb bits, DltSft of
 * depth n at arity 2^w). */
int orc_gmw_and_gates(int nb, int w, int n);
/* public_a[c-1][N] = dealer_u64(pub_seed, PUB, i, j) % P */
void orc_public_a(uint64_t pub_seed, int c, int N, uint64_t* public_a);
/* Fixed public W (stands in for emp::PRG(&F_RAND_SEED), dpf.cpp:25-26,291). */
orc_blk orc_public_W(void);

/* Full PCG_OLE<P>::gen_and_expand for both parties (pcg_ole_impl.h:46-490),
 * in-process exchange.  x0,z0,x1,z1 have N entries.  threads>1 uses OpenMP
 * over instances.  Returns 0 on success. */
int orc_pcg_ole_expand(const orc_params* p, uint64_t seed, uint64_t pub_seed,
                       uint64_t* x0, uint64_t* z0, uint64_t* x1, uint64_t* z1,
                       int threads);
/* generate_triples (pcg_ole_2pc/triple.h:45-110): N Z_p triple shares per
 * party from the OLEs of seed1 and seed2: (a_s[i], b_s[i], c_s[i]) with
 * c_0 + c_1 = (a_0 + a_1)(b_0 + b_1) per slot. */
int orc_generate_triples(const orc_params* p, uint64_t seed1, uint64_t seed2, uint64_t pub_seed,
                         uint64_t* a0, uint64_t* b0, uint64_t* c0, uint64_t* a1,
                         uint64_t* b1, uint64_t* c1, int threads);
/* Same, but only blocks [blk_begin, blk_end) of the c*c DPF blocks and no
 * step 4 (for bounded CPU timing samples).  g_red0/1 may be NULL. */
int orc_pcg_ole_blocks(const orc_params* p, uint64_t seed, int blk_begin,
                       int blk_end, uint64_t* g_red0, uint64_t* g_red1,
                       int threads);

/* ---- utilities ---------------------------------------------------------- */
uint64_t orc_fnv1a64(const uint64_t* v, size_t n, uint64_t h);
typedef struct {
    uint64_t mt[312];
    int idx;
} orc_mt64;
void orc_mt64_seed(orc_mt64* s, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64* s);
/* n successive outputs, each taken % mod when mod != 0 */
void orc_mt64_fill(orc_mt64* s, uint64_t* out, size_t n, uint64_t mod);

/* GPU_baseline/dpf_real_bench.cu:48-110 recipe (ChaCha8, party 0, peer CW 0):
 * returns the printed checksum. */
uint64_t orc_dpf_real_bench_checksum(int n, int B, int t, int w, int threads);

#ifdef __cplusplus
}
#endif
#endif
