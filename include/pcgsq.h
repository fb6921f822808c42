/*
 * pcgsq.h -- C ABI of libpcgsq, the MI355X (gfx950) implementation of the
 * PCG-OLE expansion path of tkgong/pcg-square.
 *
 * Every entry point replaces a function of the reference's device plugin seam
 * (GPU_baseline/common/ headers, namespace pcg_cuda) or a local-compute entry of
 * CPU_baseline, cited next to it.  Conventions:
 *   - return value: PCGSQ_OK (0) or a PCGSQ_E* status; no C++ exception ever
 *     crosses this boundary; pcgsq_last_error() holds the message (per thread);
 *   - PCGSQ_EUNSUPPORTED is the reference's "exit 77" case (shape or prime
 *     the device path does not cover);
 *   - all buffers are plain pointers; "host" arguments are host memory, "dev"
 *     arguments are device memory of the context's GPU;
 *   - a context owns its device scratch (no hidden process-global pools);
 *     calls on one context are serialized by the context; different contexts
 *     are independent.
 * Field: P = 2^62 - 3*2^25 + 1 (the only prime accepted, as in
 * GPU_baseline/common/poly_mul_cuda.cu:389-393).
 */
#ifndef PCGSQ_H
#define PCGSQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCGSQ_PRIME 4611686018326724609ULL

enum {
  PCGSQ_OK = 0,
  PCGSQ_EINVAL = 1,
  PCGSQ_EUNSUPPORTED = 2,
  PCGSQ_EDEVICE = 3,
  PCGSQ_ENOMEM = 4,
  PCGSQ_ECALLBACK = 5,
  PCGSQ_EINTERNAL = 6
};

/* Tree PRG (half_tree_dpf::PRGType, dpf.h:95; pcg_cuda::DpfGpuPrg, dpf_gpu.h:38). */
enum { PCGSQ_PRG_AES = 0, PCGSQ_PRG_CHACHA8 = 1 };

/* Poly-mul / NTT backends (pcg_ole::PolyMulBackend, pcg_ole.h:33). */
enum { PCGSQ_NTT_FOURSTEP = 0, PCGSQ_NTT_NAIVE = 1 };

/* 128-bit block, lo = bytes 0..7, hi = bytes 8..15 (pcg_cuda::DpfBlk,
 * dpf_gpu.h:23-26; the in-memory layout of emp::block). */
typedef struct pcgsq_blk {
  uint64_t lo, hi;
} pcgsq_blk;

typedef struct pcgsq_ctx pcgsq_ctx;

/* ---- context / runtime -------------------------------------------------- */
/* Replaces pcg_cuda::is_cuda_available (poly_mul_cuda.h:9): 1 when a gfx950
 * device is visible. */
int pcgsq_device_available(void);
/* Device name/runtime strings (pcg_cuda::device_name / runtime_version,
 * poly_mul_cuda.h:11-12). */
int pcgsq_device_name(char* buf, size_t len);
const char* pcgsq_last_error(void);
int pcgsq_ctx_create(int device, pcgsq_ctx** out);
int pcgsq_ctx_destroy(pcgsq_ctx* ctx);
/* Stream the context launches on (hipStream_t, non-null, default priority). */
void* pcgsq_ctx_stream(pcgsq_ctx* ctx);
int pcgsq_ctx_synchronize(pcgsq_ctx* ctx);
/* Per-launch device timing on the context stream, by zone name (the
 * counterpart of the reference's PCG_DPF_LEVEL_MS / PCG_SQUARE_STAGE_MS
 * splits, dpf_gpu.cu:427-493).  enable resets the totals;  read writes
 * {"zone": [launches, total_ms], ...} as JSON text. */
int pcgsq_ctx_profile(pcgsq_ctx* ctx, int enable);
int pcgsq_ctx_profile_read(pcgsq_ctx* ctx, char* buf, size_t len);
/* Register-only ChaCha8 block throughput of the device (blocks/s): the
 * compute ceiling the DPF kernels are measured against. */
int pcgsq_probe_chacha8(pcgsq_ctx* ctx, int iters, double* blocks_per_s);
/* Device F_p arithmetic on host arrays of n values, for checking it against
 * exact arithmetic (no counterpart in the reference; a self-test of the
 * device field code): op 0 high 64 bits of a*b (the butterflies' form),
 * 1 Shoup a*b mod P with c = floor(b*2^64/P), result in [0, 2P) (the
 * butterflies' form), 2 the same in the general form, 3 a*b mod P for
 * a, b < 2^62.  c may be NULL for ops 0 and 3.  Synchronizes. */
int pcgsq_field_op(pcgsq_ctx* ctx, int op, size_t n, const uint64_t* a, const uint64_t* b,
                   const uint64_t* c, uint64_t* out);
/* Plain device copy of `bytes` (read + write counted): achievable HBM
 * bytes/s (the best of the shapes pcgsq_probe_hbm sweeps), and the
 * known-size kernel that HBM counters are scaled by. */
int pcgsq_probe_copy(pcgsq_ctx* ctx, size_t bytes, double* bytes_per_s);
/* HBM probes over `bytes`: copy (read + write counted; best over 1..8
 * 16-byte loads in flight per lane, plain or non-temporal, 2..32 workgroups
 * per CU; copy_shape[3] = loads per lane, non-temporal, workgroups per CU),
 * read-only and write-only bytes/s.  Any output pointer may be NULL. */
int pcgsq_probe_hbm(pcgsq_ctx* ctx, size_t bytes, double* copy_bps, double* read_bps,
                    double* write_bps, int* copy_shape);
/* Device memory for ctypes callers that hold their own device-resident
 * buffers (no counterpart in the reference, whose plugin only takes host
 * pointers). */
int pcgsq_dev_alloc(pcgsq_ctx* ctx, size_t bytes, void** dev_out);
int pcgsq_dev_free(pcgsq_ctx* ctx, void* dev);
/* Page-locked host memory (no counterpart in the reference): outputs copied
 * into it (pcgsq_ole_outputs and the like) move by DMA at PCIe speed instead
 * of through the runtime's staging of pageable memory.  Free with
 * pcgsq_host_free. */
int pcgsq_host_alloc(size_t bytes, void** host_out);
int pcgsq_host_free(void* host);
int pcgsq_memcpy_h2d(pcgsq_ctx* ctx, void* dev_dst, const void* host_src,
                     size_t bytes);
int pcgsq_memcpy_d2h(pcgsq_ctx* ctx, void* host_dst, const void* dev_src,
                     size_t bytes);

/* ---- half-tree DPF full-domain evaluation -------------------------------- */
/* Per-level correction-word exchange (pcg_cuda::DpfExchangeFn,
 * dpf_gpu.h:31-32, as a C function pointer + user argument): fill
 * peer[0..count) from mine[0..count) for tree level `level`; return 0, or
 * nonzero to abort the expansion with PCGSQ_ECALLBACK. */
typedef int (*pcgsq_exchange_fn)(void* user, int level, const pcgsq_blk* mine,
                                 pcgsq_blk* peer, size_t count);

/* pcg_cuda::dpf_gpu_batch_full_eval (dpf_gpu.h:55-62): one party's batched
 * half-tree expansion; host roots[B], dltsft[B*levels*nk] ([b][i][k]),
 * round_keys nk*176 bytes for AES (NULL for ChaCha8), host out_leaves[B*2^n]. */
int pcgsq_dpf_batch_full_eval(pcgsq_ctx* ctx, int B, int n, int w,
                              const pcgsq_blk* roots, const pcgsq_blk* dltsft,
                              const uint8_t* round_keys, int party,
                              pcgsq_exchange_fn exchange, void* user,
                              pcgsq_blk* out_leaves, int prg);

/* pcg_cuda::dpf_gpu_batch_full_eval_device (dpf_gpu.h:69-75): same, leaves
 * stay on the device; *d_leaves aliases context scratch and is valid until
 * the next DPF call on this context. */
int pcgsq_dpf_batch_full_eval_device(pcgsq_ctx* ctx, int B, int n, int w,
                                     const pcgsq_blk* roots,
                                     const pcgsq_blk* dltsft,
                                     const uint8_t* round_keys, int party,
                                     pcgsq_exchange_fn exchange, void* user,
                                     const pcgsq_blk** d_leaves, int prg);

/* Both parties in one process (no counterpart in the reference, which runs
 * one party per process: half_tree_dpf.cpp test drivers).  The exchange is
 * the other party's reduced share, already on the device.  Host arrays as
 * above per party; optional cw_out[levels*B*nk] ([i][b][k]) receives the
 * reconstructed CWs (k < m-1). */
int pcgsq_dpf_batch_full_eval_both(pcgsq_ctx* ctx, int B, int n, int w,
                                   const pcgsq_blk* roots0,
                                   const pcgsq_blk* roots1,
                                   const pcgsq_blk* dltsft0,
                                   const pcgsq_blk* dltsft1,
                                   const uint8_t* round_keys,
                                   pcgsq_blk* leaves0, pcgsq_blk* leaves1,
                                   pcgsq_blk* cw_out, int prg);

/* ---- output layer (leaf_convert_cuda.h) ----------------------------------- */
/* pcg_cuda::dpf_out_sums (leaf_convert_cuda.h:43-44): d_leaves device
 * [B][D]; host sumC[B] (mod P), sumT[B]. */
int pcgsq_dpf_out_sums(pcgsq_ctx* ctx, const pcgsq_blk* d_leaves, int B,
                       size_t D, uint64_t prime, uint64_t* sumC,
                       uint64_t* sumT);
/* No counterpart in the reference: pcgsq_dpf_out_sums for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops or
 * harnesses that forbid hipMalloc).  It allocates nothing and does not
 * synchronize: the sums of the device leaves d_leaves [B][D] go to the device
 * arrays d_sumC[B] (mod P) and d_sumT[B], enqueued on the context stream.
 * Its partial sums live in a fixed device array of
 * PCGSQ_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends):
 * 16 * B * max(1, D / min(D, 2048)) bytes must fit, else PCGSQ_EINVAL.  All
 * calls on one GPU share that array, so calls that overlap on one GPU (other
 * contexts or streams) race; calls on one context are ordered by its stream. */
#define PCGSQ_NOALLOC_SCRATCH_BYTES (4u << 20)
int pcgsq_dpf_out_sums_noalloc(pcgsq_ctx* ctx, const pcgsq_blk* d_leaves, int B, size_t D,
                               uint64_t prime, uint64_t* d_sumC, uint64_t* d_sumT);
/* pcg_cuda::dpf_out_scatter_g (leaf_convert_cuda.h:49-52): host CW[B],
 * host g_out[N] overwritten with the folded y = +-(C + tau*CW). */
int pcgsq_dpf_out_scatter_g(pcgsq_ctx* ctx, const pcgsq_blk* d_leaves, int B,
                            size_t D, int t, int party, uint64_t prime,
                            const uint64_t* CW, uint64_t* g_out, int N);

/* ---- polynomial multiplication over R_p = F_p[X]/(X^N+1) ------------------ */
/* pcg_cuda::poly_mul_supported / poly_mul_gpuntt_square_supported
 * (poly_mul_cuda.h:10,163): 1 if supported, else 0 with a reason. */
int pcgsq_poly_mul_supported(uint64_t prime, int N, int backend, char* reason,
                             size_t reason_len);
/* pcg_cuda::poly_mul_u64 / poly_mul_u64_gpuntt_square (poly_mul_cuda.h:
 * 154-166): host buffers of batch*N words; out may alias a or b.
 * device_ms (nullable) receives the device time of the call. */
int pcgsq_poly_mul_u64(pcgsq_ctx* ctx, uint64_t* out, const uint64_t* a,
                       const uint64_t* b, int batch, int N, uint64_t prime,
                       int backend, double* device_ms);
/* Device-buffer variant (no copies): dev_out, dev_a, dev_b. */
int pcgsq_poly_mul_u64_dev(pcgsq_ctx* ctx, uint64_t* dev_out,
                           const uint64_t* dev_a, const uint64_t* dev_b,
                           int batch, int N, int backend);
/* NTT<P>::forward / inverse (CPU_baseline/ntt.h:65-78), batched, host
 * buffers, in place; forward output in natural order a(psi^(2k+1)). */
int pcgsq_ntt_forward(pcgsq_ctx* ctx, uint64_t* a, int batch, int N,
                      int backend);
int pcgsq_ntt_inverse(pcgsq_ctx* ctx, uint64_t* a, int batch, int N,
                      int backend);
int pcgsq_ntt_forward_dev(pcgsq_ctx* ctx, uint64_t* dev_a, int batch, int N,
                          int backend);
int pcgsq_ntt_inverse_dev(pcgsq_ctx* ctx, uint64_t* dev_a, int batch, int N,
                          int backend);

/* Shape of the NTT passes for length N on this context (no counterpart in the
 * reference; set by the autotuner, pcgsq/autotune.py): column length
 * 2^n1_log (n1_log = log2 N: one pass, allowed up to N = 2^13; 0: default,
 * one pass up to 2^13 and ceil(log2 N / 2) above), variant bit 0 / bit 1 =
 * alternative register grouping of the length-n1 / length-n2 passes, cols1 /
 * cols2 = sub-transforms per workgroup tile of those passes (power of two,
 * 0 = automatic).  Results do not depend on the shape. */
int pcgsq_ntt_tune(pcgsq_ctx* ctx, int N, int n1_log, int variant, int cols1, int cols2);

/* ---- Batched NTT of uneven lengths ------------------------------------------
 * No counterpart in the reference (its GPU NTT tiers take one length per
 * call: poly_mul_cuda.h:9-31, ntt_batch_bench.cu).  For callers that hold
 * many transforms of different lengths at once: `count` transforms of
 * lengths 2^logN[i] (2^10..2^22) in one persistent launch that drains a
 * device work list (pass-major tiles with per-pass completion counters).
 * ops: PCGSQ_NTT_BATCH_FORWARD (a -> f), PCGSQ_NTT_BATCH_INVERSE (a -> r) or
 * both (a -> f -> r).  Transform i occupies words [off_i, off_i + 2^logN[i])
 * of the batch buffers, off_i = sum of the earlier lengths.  Results equal
 * pcgsq_ntt_forward / pcgsq_ntt_inverse per transform, bit for bit. */
typedef struct pcgsq_ntt_batch pcgsq_ntt_batch;
#define PCGSQ_NTT_BATCH_FORWARD 1
#define PCGSQ_NTT_BATCH_INVERSE 2
#define PCGSQ_NTT_BATCH_PERSISTENT 0 /* one persistent launch (default) */
#define PCGSQ_NTT_BATCH_GROUPED 1    /* per-length pass launches over offset tables */
int pcgsq_ntt_batch_create(pcgsq_ctx* ctx, const int* logN, int count, int ops,
                           pcgsq_ntt_batch** out);
int pcgsq_ntt_batch_destroy(pcgsq_ntt_batch* b);
/* Shape (set by the autotuner): split of length 2^logN (n1_log = logN: one
 * pass, allowed up to 2^12; else a column pass of 2^7..2^11 and a row pass of
 * 2^6..2^11; 0: default), resident
 * workgroups per CU (0 = as many as fit) and the dependency spin limit
 * (0 = default); logN = 0 changes only the last two. */
int pcgsq_ntt_batch_tune(pcgsq_ntt_batch* b, int logN, int n1_log, int wgs_per_cu,
                         uint32_t spin_limit);
/* How a run executes (results are the same): PCGSQ_NTT_BATCH_PERSISTENT, one
 * persistent launch draining per-XCD work lists; or PCGSQ_NTT_BATCH_GROUPED,
 * for each length the per-length pass launches (the context's tuned shape for
 * that length) over all transforms of that length through an offset table. */
int pcgsq_ntt_batch_set_mode(pcgsq_ntt_batch* b, int mode);
/* Device buffers (inputs a, forward f, inverse r) and their length in words. */
int pcgsq_ntt_batch_dev(pcgsq_ntt_batch* b, uint64_t** a, uint64_t** f, uint64_t** r,
                        size_t* total_words);
/* Copy the concatenated inputs from the host (synchronizes). */
int pcgsq_ntt_batch_load(pcgsq_ntt_batch* b, const uint64_t* host);
/* Enqueue the whole batch on the context stream (use_graph: replay a
 * captured HIP graph of the launch). */
int pcgsq_ntt_batch_run(pcgsq_ntt_batch* b, int use_graph);
/* Wait for the batch; PCGSQ_EDEVICE if a dependency wait hit its limit. */
int pcgsq_ntt_batch_finish(pcgsq_ntt_batch* b);
/* Copy f and/or r to the host (either may be NULL; synchronizes). */
int pcgsq_ntt_batch_read(pcgsq_ntt_batch* b, uint64_t* f_host, uint64_t* r_host);
/* Work items (grouped mode: launches), grid and LDS bytes per workgroup of
 * the persistent launch (0 in grouped mode), and the algorithmic HBM bytes of
 * one run: 16 * N * p(N) per transform and direction, p = 1 up to 2^13 (the
 * transform fits one workgroup's LDS), else 2. */
int pcgsq_ntt_batch_info(pcgsq_ntt_batch* b, int* nitems, int* grid, size_t* lds_bytes,
                         double* alg_bytes);

/* ---- Batched DPF keys of uneven domains -------------------------------------
 * No single counterpart in the reference: DPF.Gen / DPF.Eval of
 * half_tree_dpf/dpf.cpp:302-398 (the gen/eval primitives behind the fused
 * path) for `count` key pairs with domains 2^logn[i] (2^10..2^30), w = 1, and
 * the output conversion of pcg_ole_impl.h:432-446.  Gen is the dealer form:
 * roots Delta_s ^ W and the path-only correction words, which equal the
 * fused protocol's exchange result; then both parties' whole domains are
 * expanded sync-free in one persistent launch and converted to field shares
 * y_s(j) = (-1)^s (C_s(j) + tau_s(j) cw_out), y_0 + y_1 = beta [j == alpha].
 * Outputs: party s, key i at words [s * total + off_i, + 2^logn[i]),
 * off_i = sum of the earlier domains. */
typedef struct pcgsq_dpf_batch pcgsq_dpf_batch;
int pcgsq_dpf_batch_create(pcgsq_ctx* ctx, const int* logn, int count, int prg,
                           pcgsq_dpf_batch** out);
int pcgsq_dpf_batch_destroy(pcgsq_dpf_batch* b);
/* Gen from the dealer's seeded draws (alpha, beta, Delta, Delta_0 share). */
int pcgsq_dpf_batch_deal(pcgsq_dpf_batch* b, uint64_t seed);
/* Gen from given points: alpha[i] < 2^logn[i], beta[i] < P, Delta with lsb 1,
 * Delta_0 the first party's share of Delta. */
int pcgsq_dpf_batch_gen(pcgsq_dpf_batch* b, const uint64_t* alpha, const uint64_t* beta,
                        const pcgsq_blk* delta, const pcgsq_blk* delta0);
/* The dealt points (either pointer may be NULL). */
int pcgsq_dpf_batch_points(pcgsq_dpf_batch* b, uint64_t* alpha, uint64_t* beta);
/* The keys: roots [2][count], correction words (key i's n words after the
 * earlier keys'), output words [count]; any pointer may be NULL. */
int pcgsq_dpf_batch_keys(pcgsq_dpf_batch* b, pcgsq_blk* roots, pcgsq_blk* cw, uint64_t* cw_out);
/* Enqueue the full-domain eval of every key of both parties (use_graph:
 * replay a captured HIP graph of the launch). */
int pcgsq_dpf_batch_run(pcgsq_dpf_batch* b, int use_graph);
/* Copy the outputs of party 0 / party 1 (either may be NULL; synchronizes). */
int pcgsq_dpf_batch_read(pcgsq_dpf_batch* b, uint64_t* y0, uint64_t* y1);
/* Device pointer of the outputs [2][total] and total. */
int pcgsq_dpf_batch_dev(pcgsq_dpf_batch* b, uint64_t** y, size_t* total_words);
/* Work items, grid, and the ChaCha8/AES tree blocks and out-hash blocks of
 * one run (both parties). */
int pcgsq_dpf_batch_info(pcgsq_dpf_batch* b, int* nitems, int* grid, double* tree_blocks,
                         double* out_blocks);

/* ---- AES (aes_gpu.h) -------------------------------------------------------- */
/* pcg_cuda::aes128_expand_key_host (aes_gpu.h:13). */
int pcgsq_aes128_expand_key(const uint8_t key[16], uint8_t round_keys[176]);
/* pcg_cuda::aes128_ecb_encrypt_gpu (aes_gpu.h:18-19): count blocks, one
 * 16-byte key per block, host buffers. */
int pcgsq_aes128_ecb_encrypt(pcgsq_ctx* ctx, const uint8_t* keys,
                             const uint8_t* in, uint8_t* out, int count);
/* AES tree-PRG round keys for nk sub-keys derived from a master key
 * (AESTreePRG ctor, half_tree_dpf/tree_prg.h:45-56): rks[nk*176]. */
int pcgsq_aes_tree_round_keys(const uint8_t master[16], int nk, uint8_t* rks);

/* ---- PCG-OLE expansion (pcg_ole::PCG_OLE<P>::gen_and_expand) --------------- */
/* pcg_ole::PCGParams (pcg_ole.h:271-281) without the backends, which are
 * this library. */
typedef struct pcgsq_params {
  int N;    /* ring dimension, power of two, 2N | P-1 */
  int c;    /* sparse polynomials per party */
  int t;    /* nonzeros per sparse polynomial, t | N */
  int w;    /* DPF arity exponent, w | log2(2N/t) */
  int prg;  /* PCGSQ_PRG_* */
} pcgsq_params;

typedef struct pcgsq_ole_plan pcgsq_ole_plan;

/* Plan for one parameter set: device buffers, the public a_i (drawn from
 * pub_seed) and their NTTs stay resident.  hbm_budget_bytes bounds the
 * expansion's working set (0 = choose). */
int pcgsq_ole_plan_create(pcgsq_ctx* ctx, const pcgsq_params* p,
                          uint64_t pub_seed, size_t hbm_budget_bytes,
                          pcgsq_ole_plan** out);
int pcgsq_ole_plan_destroy(pcgsq_ole_plan* plan);
/* Device bytes the plan holds. */
size_t pcgsq_ole_plan_bytes(const pcgsq_ole_plan* plan);
/* Device bytes a plan of these parameters and budget would take, including
 * its context's NTT scratch (0 if the parameters are invalid). */
size_t pcgsq_ole_plan_estimate(const pcgsq_params* p, size_t hbm_budget_bytes);

/* ---- One party of the 2PC expansion ------------------------------------------
 * PCG_OLE<P>::gen_and_expand as one party's process runs it
 * (pcg_ole_impl.h:46-490, with emp::NetIO replaced by a callback): the party
 * expands only its own DPF trees, and every exchange of the protocol goes
 * through `fn`, in the order the reference makes them.  Rows (phase, level,
 * batch, send bytes, receive bytes are the columns of the reference's
 * comm_trace.h rows):
 *   "base_ot"     the first run of a plan dealt with pcgsq_ole_party_deal_own:
 *                 the Chou-Orlandi base OTs, level 0 (33 bytes) and 1
 *                 (128 * 33 bytes);
 *   "ot_ext"      per run of such a plan: the IKNP matrix of the OTs this
 *                 party receives (16 bytes per OT; the two parties' sizes
 *                 differ);
 *   "gilboa"      party 0's corrections of the Gilboa OTs (party 1 sends 0);
 *   "circuit_and" per AND depth of the position adder, then of the DltSft
 *                 circuit, and per DPF block: the opened d, e of the layer's
 *                 AND gates (bool_circuit/circuit.cpp:306-330); level = the
 *                 circuit's depth, batch = AND gates;
 *   "dpf_gen"     per DPF level and block: the level's correction-word shares,
 *                 B * (2^w - 1) blocks of 16 bytes (dpf.cpp:520-551);
 *   "beaver_de"   per block: d and e of the first Beaver open, 2B u64;
 *   "beaver_cw"   per block: the CW shares of the second open, B u64
 *                 (pcg_ole_impl.h:231-295); level = block index ii*c + jj.
 * fn sends send_bytes from `send` and receives recv_bytes from the other
 * party into `recv` (either may be 0; party 0 sends first, party 1 receives
 * first, as NetIO's send/flush/recv pairs run), and returns 0 (nonzero
 * aborts the run with PCGSQ_ECALLBACK).
 * Inputs:
 *   pcgsq_ole_party_deal_own(own_seed): the party's own noise, Delta share
 *     and Beaver randomness from its own seed; the run first computes alpha,
 *     DltSft, beta and rs with the other party (OT extension from base OTs,
 *     GMW, Gilboa products: the 2PC preprocessing of pcg_ole_impl.h:62-199).
 *     No party ever holds the other's secrets.
 *   pcgsq_ole_deal(seed): a trusted-dealer simulation: both parties' shares
 *     derived from one shared seed (each party process can recompute the
 *     other's secrets), the run starting at the DPF.
 * Two parties' plans together give exactly the outputs of one both-party
 * plan dealt alike (deal_own with seeds S, S equals pcgsq_ole_deal(S) and
 * pcgsq_ole_deal_gmw(S)). */
typedef int (*pcgsq_ole_exchange_fn)(void* user, const char* phase, int level, int batch,
                                     const void* send, size_t send_bytes, void* recv,
                                     size_t recv_bytes);
/* Exchange granularity: per DPF block, as the reference's serial arm
 * (default), or one exchange per DPF level, Beaver open and circuit depth
 * for a whole group of blocks, as its batched arm
 * (GPU_baseline/e2e_full_bench.cu:128-153).  Results are the same. */
enum { PCGSQ_EXCHANGE_PER_BLOCK = 0, PCGSQ_EXCHANGE_BATCHED = 1 };
int pcgsq_ole_party_create(pcgsq_ctx* ctx, const pcgsq_params* p, int party, uint64_t pub_seed,
                           size_t hbm_budget_bytes, pcgsq_ole_plan** out);
int pcgsq_ole_party_run(pcgsq_ole_plan* plan, pcgsq_ole_exchange_fn fn, void* user);
int pcgsq_ole_party_outputs(pcgsq_ole_plan* plan, uint64_t* x, uint64_t* z);
int pcgsq_ole_party_deal_own(pcgsq_ole_plan* plan, uint64_t own_seed);
int pcgsq_ole_party_set_exchange_mode(pcgsq_ole_plan* plan, int mode);
/* One party's base OTs (Chou-Orlandi over P-256, two "base_ot" exchanges)
 * without a device: for each direction d the party that receives extended
 * OTs there holds both 128-bit seeds of every column (k0, k1: direction
 * 1 - party), the other its choice bits s and the chosen seeds ks
 * (direction party), ks[i] = s_i ? k1[i] : k0[i] of the other party. */
typedef struct pcgsq_base_ots {
  pcgsq_blk k0[128], k1[128];
  pcgsq_blk s;
  pcgsq_blk ks[128];
} pcgsq_base_ots;
int pcgsq_base_ots_run(int party, uint64_t seed, pcgsq_ole_exchange_fn fn, void* user,
                       pcgsq_base_ots* out);
/* What one party's last preprocessing left on the device, for tests of the
 * protocol's invariants (no counterpart in the reference): the choice words
 * of the OTs it received, the 16-byte messages it received, the two it sent,
 * its adder AND triples [ninst][ops][a, b, c] (u64 bit-sliced words), its
 * Delta-AND triples [ninst][levels*nk][a, b.lo, b.hi, c.lo, c.hi] and its
 * AND-tree triples [ninst][levels*nk] (u32: a bits 0..3, b 4..7, c 8..11).
 * *needed receives the size; host may be NULL to query it.  Synchronizes. */
enum {
  PCGSQ_PRE_CHOICE = 0,
  PCGSQ_PRE_RECEIVED = 1,
  PCGSQ_PRE_SENT0 = 2,
  PCGSQ_PRE_SENT1 = 3,
  PCGSQ_PRE_ADDER_TRIPLES = 4,
  PCGSQ_PRE_DELTA_TRIPLES = 5,
  PCGSQ_PRE_TREE_TRIPLES = 6
};
int pcgsq_ole_preprocessing_read(pcgsq_ole_plan* plan, int party, int what, void* host,
                                 size_t bytes, size_t* needed);
/* Deal the per-expansion inputs from `seed` (the seeded stand-in for the
 * reference's noise sampling and 2PC preprocessing, pcg_ole_impl.h:62-199):
 * the dealer runs on the device, asynchronously on the context stream.  On a
 * one-party plan this is the trusted-dealer simulation described above. */
int pcgsq_ole_deal(pcgsq_ole_plan* plan, uint64_t seed);
/* Both-party plans: the same draws, with alpha, DltSft, beta and rs computed
 * by the two parties' 2PC preprocessing (pcg_ole_impl.h:114-199) run
 * in-process: base OTs (once per plan), IKNP random OTs both ways, the
 * Kogge-Stone adder (bool_circuit/adder.h:21-63) and the F_DeltaShiftShare
 * circuit (half_tree_dpf/dpf.cpp:172-240) under GMW
 * (bool_circuit/circuit.cpp:245-380) with AND triples from the OTs, and beta
 * and the Beaver triples' rs by Gilboa OT products (circuit.cpp:91-200).
 * Each party uses `seed` as its own seed; every message moves by a device
 * copy between the parties' buffers.  The DltSft, beta and rs shares differ
 * from pcgsq_ole_deal's but reconstruct to the same values, so the
 * expansion's outputs are identical. */
int pcgsq_ole_deal_gmw(pcgsq_ole_plan* plan, uint64_t seed);
/* Both parties' dealt DltSft shares, [2][c*c*t*t][levels][2^w - 1] blocks
 * (synchronizes). */
int pcgsq_ole_dealt_dltsft(pcgsq_ole_plan* plan, pcgsq_blk* dltsft);
/* Both parties' dealt F_p shares, out[8][c*c*t*t] = beta0, beta1, r0, r1,
 * s0, s1, rs0, rs1 per instance (synchronizes). */
int pcgsq_ole_dealt_products(pcgsq_ole_plan* plan, uint64_t* out);
/* Run the whole expansion for both parties on the device-resident inputs
 * (DPF batches, output layer with Beaver opens, step 4), asynchronously on
 * the context stream. */
int pcgsq_ole_run(pcgsq_ole_plan* plan);
/* Copy the outputs (x_s, z_s with z0 + z1 = x0 * x1) to the host; any
 * pointer may be NULL.  Synchronizes. */
int pcgsq_ole_outputs(pcgsq_ole_plan* plan, uint64_t* x0, uint64_t* z0,
                      uint64_t* x1, uint64_t* z1);
/* Device pointers of the outputs [4][N] (x0, z0, x1, z1). */
const uint64_t* pcgsq_ole_outputs_dev(pcgsq_ole_plan* plan);
/* Convenience: deal + run + outputs. */
int pcgsq_ole_expand(pcgsq_ctx* ctx, const pcgsq_params* p, uint64_t seed,
                     uint64_t pub_seed, uint64_t* x0, uint64_t* z0,
                     uint64_t* x1, uint64_t* z1);
/* Only the DPF blocks and the output layer (no step 4): afterwards
 * pcgsq_ole_g returns the folded g_ij of every block for both parties,
 * [2][c*c][N] (pcg_ole_impl.h:432-449 + the fold of :474-478).  After a full
 * pcgsq_ole_run that buffer holds NTT-domain values instead. */
int pcgsq_ole_run_blocks(pcgsq_ole_plan* plan);
int pcgsq_ole_g(pcgsq_ole_plan* plan, uint64_t* g_host);

/* ---- Serving: continuous batching of expansion requests -------------------
 * No counterpart in the reference (its benchmarks run one expansion per
 * process).  A server on `device` runs up to `slots` requests at once, each
 * on a pooled plan with its own stream.  When a slot is free it admits the
 * oldest waiting request that can start now: one whose shape has an idle
 * plan (requests of one shape keep their order).  A shape with no plan gets
 * one built within the device-memory budget (0 = 70% of free memory; idle
 * plans of other shapes are evicted least recently used first); further
 * plans of a shape, so that its requests overlap, come from
 * pcgsq_server_reserve.  Running requests overlap on the device.  Seeds are
 * dealt on the device and outputs stay there, in buffers pooled by size,
 * until copied or released.  A request's result does not depend on what else
 * runs. */
typedef struct pcgsq_server pcgsq_server;
enum { PCGSQ_REQ_QUEUED = 0, PCGSQ_REQ_RUNNING = 1, PCGSQ_REQ_DONE = 2, PCGSQ_REQ_FAILED = 3 };
int pcgsq_server_create(int device, size_t hbm_budget_bytes, int slots, pcgsq_server** out);
int pcgsq_server_destroy(pcgsq_server* s);
/* Queue one expansion (PCGSQ_ENOMEM if it can never fit the budget). */
int pcgsq_server_submit(pcgsq_server* s, const pcgsq_params* p, uint64_t seed,
                        uint64_t pub_seed, uint64_t* id);
int pcgsq_server_poll(pcgsq_server* s, uint64_t id, int* state);
/* Block until the request is done (timeout_s <= 0: no limit). */
int pcgsq_server_wait(pcgsq_server* s, uint64_t id, double timeout_s);
/* Copy x0, z0, x1, z1 of a finished request (any pointer may be NULL). */
int pcgsq_server_outputs(pcgsq_server* s, uint64_t id, uint64_t* x0, uint64_t* z0,
                         uint64_t* x1, uint64_t* z1);
/* Device pointer of a finished request's outputs [4][N]. */
int pcgsq_server_outputs_dev(pcgsq_server* s, uint64_t id, const uint64_t** out);
/* Drop a finished request; its output buffer returns to the server's pool. */
int pcgsq_server_release(pcgsq_server* s, uint64_t id);
/* Submit, admission and completion times (seconds of pcgsq_server_clock). */
int pcgsq_server_timing(pcgsq_server* s, uint64_t id, double* t_submit, double* t_start,
                        double* t_done);
/* Budget, bytes in use, peak bytes in use, completed and queued requests. */
int pcgsq_server_stats(pcgsq_server* s, size_t* budget, size_t* used, size_t* peak,
                       long* completed, int* queued);
/* Create plans of this shape ahead of requests until `count` exist (they
 * stay until evicted for memory). */
int pcgsq_server_reserve(pcgsq_server* s, const pcgsq_params* p, uint64_t pub_seed, int count);
/* Device bytes the server charges a request of these parameters (its plan
 * under the server's per-plan budget, plus the outputs). */
int pcgsq_server_charge(pcgsq_server* s, const pcgsq_params* p, size_t* bytes);
/* Work window: the most ring dimension (sum of N) that running requests may
 * hold at once (0, the default: no limit).  A request that would exceed it
 * waits unless nothing runs, and so does every later request at least as
 * large, so that large requests start in arrival order while smaller ones
 * fill the slots around them. */
int pcgsq_server_set_window(pcgsq_server* s, uint64_t max_running_N);
/* Admission policy (0 switches either part off, the default):
 *   exclusive_N: requests with N >= exclusive_N (expansions that fill the
 *     GPU on their own) run one at a time and start in arrival order among
 *     themselves; smaller requests fill the other slots around them;
 *   age_bound_s: a waiting request older than this that cannot start is
 *     passed by no later request (admission stops at it until it starts), so
 *     skipping ahead delays no request by more than the bound. */
int pcgsq_server_set_policy(pcgsq_server* s, uint64_t exclusive_N, double age_bound_s);
/* Plans in the pool, plans built and plans evicted since creation. */
int pcgsq_server_plans(pcgsq_server* s, int* plans, long* built, long* evicted);
/* The server's monotonic clock (seconds). */
double pcgsq_server_clock(void);

/* pcg_ole::generate_triples (CPU_baseline/pcg_ole_2pc/triple.h:45-110) for
 * both parties: the plan's expansion for seed1 and for seed2 (the reference's
 * two gen_and_expand calls), ring triple shares a_0 = x_{0,1}, b_0 = x_{0,2},
 * a_1 = x_{1,2}, b_1 = x_{1,1}, c_s = a_s b_s + z_{s,1} + z_{s,2}, and their
 * forward NTTs: N Z_p triple shares per party with
 * c0[i] + c1[i] = (a0[i] + a1[i]) (b0[i] + b1[i]).  Outputs [N] (any may be
 * NULL); synchronizes. */
int pcgsq_triples_generate(pcgsq_ole_plan* plan, uint64_t seed1, uint64_t seed2, uint64_t* a0,
                           uint64_t* b0, uint64_t* c0, uint64_t* a1, uint64_t* b1,
                           uint64_t* c1);

/* Dealer value U(seed, domain, a, b) of the documented dealer spec. */
uint64_t pcgsq_dealer_u64(uint64_t seed, uint32_t domain, uint64_t a,
                          uint64_t b);

#ifdef __cplusplus
}
#endif
#endif
