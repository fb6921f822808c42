// pcg_hip.hpp -- C++ drop-in for the reference's device plugin seam.
//
// Same names, argument meaning, ownership and error behaviour as the
// reference's pcg_cuda:: headers (GPU_baseline/common/dpf_gpu.h,
// leaf_convert_cuda.h, poly_mul_cuda.h, aes_gpu.h), implemented over the C
// ABI of libpcgsq (include/pcgsq.h).  A CPU_baseline build switches from the
// CUDA plugin to this one by including this header where it includes the
// pcg_cuda headers and aliasing the namespace:
//
//     #include "pcgsq/pcg_hip.hpp"
//     namespace pcg_cuda = pcg_hip;
//
// Like the reference plugin, the functions use one process-wide device
// context (device 0), are synchronous, serialize on an internal mutex, take
// host pointers, and throw std::runtime_error on failure.  Pointers returned
// by the *_device functions alias internal scratch and stay valid until the
// next DPF call.
#ifndef PCGSQ_PCG_HIP_HPP
#define PCGSQ_PCG_HIP_HPP

#include <stddef.h>
#include <stdint.h>

#include <exception>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "pcgsq.h"

namespace pcg_hip {

// ---- common/dpf_gpu.h -------------------------------------------------------
struct alignas(16) DpfBlk {   // dpf_gpu.h:23-26
  uint64_t lo;
  uint64_t hi;
};
using DpfExchangeFn =         // dpf_gpu.h:31-32
    std::function<void(const DpfBlk* my_cws, DpfBlk* peer_cws, size_t count)>;
enum class DpfGpuPrg { AES, CHACHA8 };   // dpf_gpu.h:38

// ---- common/leaf_convert_cuda.h ----------------------------------------------
constexpr uint64_t kOutHashDomain = 0x6f75745f68617368ULL;   // leaf_convert_cuda.h:38

// ---- common/poly_mul_cuda.h --------------------------------------------------
struct PolyMulStats {   // poly_mul_cuda.h:9-11
  double device_ms = 0.0;
};

namespace detail {

inline pcgsq_ctx* ctx() {
  static pcgsq_ctx* c = [] {
    pcgsq_ctx* p = nullptr;
    if (pcgsq_ctx_create(0, &p) != PCGSQ_OK)
      throw std::runtime_error(std::string("pcg_hip: ") + pcgsq_last_error());
    return p;
  }();
  return c;
}

inline std::mutex& mu() {
  static std::mutex m;
  return m;
}

inline void check(int rc) {
  if (rc != PCGSQ_OK) throw std::runtime_error(std::string("pcg_hip: ") + pcgsq_last_error());
}

// The caller's exchange and the exception it threw: the trampoline keeps the
// exception, the C call returns PCGSQ_ECALLBACK, and the wrapper rethrows the
// exchange's own exception unchanged (as the reference's NetIO errors reach
// its caller), not a generic runtime_error.
struct ExchangeCall {
  const DpfExchangeFn* fn;
  std::exception_ptr err;
};

inline int exchange_tramp(void* user, int /*level*/, const pcgsq_blk* mine, pcgsq_blk* peer,
                          size_t count) {
  ExchangeCall* x = static_cast<ExchangeCall*>(user);
  try {
    (*x->fn)(reinterpret_cast<const DpfBlk*>(mine), reinterpret_cast<DpfBlk*>(peer), count);
    return 0;
  } catch (...) {
    x->err = std::current_exception();
    return 1;
  }
}

// Status of a C call that took an exchange: the exchange's exception first.
inline void check_call(int rc, const ExchangeCall& x) {
  if (x.err) std::rethrow_exception(x.err);
  check(rc);
}

inline int prg_code(DpfGpuPrg p) { return p == DpfGpuPrg::AES ? PCGSQ_PRG_AES : PCGSQ_PRG_CHACHA8; }

}  // namespace detail

// dpf_gpu.h:55-62
inline void dpf_gpu_batch_full_eval(int B, int n, int w, const DpfBlk* roots,
                                    const DpfBlk* dltsft, const uint8_t* round_keys, int party,
                                    const DpfExchangeFn& exchange, DpfBlk* out_leaves,
                                    DpfGpuPrg prg = DpfGpuPrg::AES) {
  if (B <= 0) return;
  std::lock_guard<std::mutex> lk(detail::mu());
  detail::ExchangeCall call{&exchange, nullptr};
  const int rc = pcgsq_dpf_batch_full_eval(
      detail::ctx(), B, n, w, reinterpret_cast<const pcgsq_blk*>(roots),
      reinterpret_cast<const pcgsq_blk*>(dltsft), round_keys, party, detail::exchange_tramp,
      &call, reinterpret_cast<pcgsq_blk*>(out_leaves), detail::prg_code(prg));
  detail::check_call(rc, call);
}

// dpf_gpu.h:69-75
inline const DpfBlk* dpf_gpu_batch_full_eval_device(int B, int n, int w, const DpfBlk* roots,
                                                    const DpfBlk* dltsft,
                                                    const uint8_t* round_keys, int party,
                                                    const DpfExchangeFn& exchange,
                                                    DpfGpuPrg prg = DpfGpuPrg::AES) {
  if (B <= 0) return nullptr;
  std::lock_guard<std::mutex> lk(detail::mu());
  const pcgsq_blk* d = nullptr;
  detail::ExchangeCall call{&exchange, nullptr};
  const int rc = pcgsq_dpf_batch_full_eval_device(
      detail::ctx(), B, n, w, reinterpret_cast<const pcgsq_blk*>(roots),
      reinterpret_cast<const pcgsq_blk*>(dltsft), round_keys, party, detail::exchange_tramp,
      &call, &d, detail::prg_code(prg));
  detail::check_call(rc, call);
  return reinterpret_cast<const DpfBlk*>(d);
}

// leaf_convert_cuda.h:43-44
inline void dpf_out_sums(const DpfBlk* d_leaves, int B, size_t D, uint64_t prime,
                         uint64_t* sumC_host, uint64_t* sumT_host) {
  if (B <= 0) return;
  std::lock_guard<std::mutex> lk(detail::mu());
  detail::check(pcgsq_dpf_out_sums(detail::ctx(), reinterpret_cast<const pcgsq_blk*>(d_leaves), B,
                                   D, prime, sumC_host, sumT_host));
}

// leaf_convert_cuda.h:49-52
inline void dpf_out_scatter_g(const DpfBlk* d_leaves, int B, size_t D, int t, int party,
                              uint64_t prime, const uint64_t* CW_host, uint64_t* g_out_host,
                              int N) {
  if (B <= 0) return;
  std::lock_guard<std::mutex> lk(detail::mu());
  detail::check(pcgsq_dpf_out_scatter_g(detail::ctx(),
                                        reinterpret_cast<const pcgsq_blk*>(d_leaves), B, D, t,
                                        party, prime, CW_host, g_out_host, N));
}

// poly_mul_cuda.h:9
inline bool is_cuda_available() { return pcgsq_device_available() != 0; }

inline std::string device_name() {
  char buf[256] = {0};
  std::lock_guard<std::mutex> lk(detail::mu());
  (void)detail::ctx();
  detail::check(pcgsq_device_name(buf, sizeof buf));
  return buf;
}

inline std::string runtime_version() { return device_name(); }

namespace detail {
inline bool supported(uint64_t prime, int N, int backend, std::string* reason) {
  char buf[256] = {0};
  const bool ok = pcgsq_poly_mul_supported(prime, N, backend, buf, sizeof buf) != 0;
  if (!ok && reason) *reason = buf;
  return ok;
}
inline void poly_mul(uint64_t* out, const uint64_t* a, const uint64_t* b, int batch, int N,
                     uint64_t prime, int backend, PolyMulStats* stats) {
  std::lock_guard<std::mutex> lk(mu());
  double ms = 0;
  check(pcgsq_poly_mul_u64(ctx(), out, a, b, batch, N, prime, backend, &ms));
  if (stats) stats->device_ms = ms;
}
}  // namespace detail

// v1 seam (poly_mul_cuda.h:10,15-17): served by the naive per-stage NTT.
inline bool poly_mul_supported(uint64_t prime, int N, std::string* reason = nullptr) {
  return detail::supported(prime, N, PCGSQ_NTT_NAIVE, reason);
}
inline void poly_mul_u64(uint64_t* out, const uint64_t* a, const uint64_t* b, int batch, int N,
                         uint64_t prime, PolyMulStats* stats = nullptr) {
  detail::poly_mul(out, a, b, batch, N, prime, PCGSQ_NTT_NAIVE, stats);
}
// four-step ("square") seam (poly_mul_cuda.h:24-29)
inline bool poly_mul_gpuntt_square_supported(uint64_t prime, int N,
                                             std::string* reason = nullptr) {
  return detail::supported(prime, N, PCGSQ_NTT_FOURSTEP, reason);
}
inline void poly_mul_u64_gpuntt_square(uint64_t* out, const uint64_t* a, const uint64_t* b,
                                       int batch, int N, uint64_t prime,
                                       PolyMulStats* stats = nullptr) {
  detail::poly_mul(out, a, b, batch, N, prime, PCGSQ_NTT_FOURSTEP, stats);
}
// The names pcg_ole.h:147-157,226-230 calls for its "merge" and "4step"
// backends (declared by no reference header): both map to the four-step NTT.
inline bool poly_mul_gpuntt_supported(uint64_t prime, int N, std::string* reason = nullptr) {
  return poly_mul_gpuntt_square_supported(prime, N, reason);
}
inline void poly_mul_u64_gpuntt(uint64_t* out, const uint64_t* a, const uint64_t* b, int batch,
                                int N, uint64_t prime, PolyMulStats* stats = nullptr) {
  poly_mul_u64_gpuntt_square(out, a, b, batch, N, prime, stats);
}
inline bool poly_mul_gpuntt_4step_supported(uint64_t prime, int N,
                                            std::string* reason = nullptr) {
  return poly_mul_gpuntt_square_supported(prime, N, reason);
}
inline void poly_mul_u64_gpuntt_4step(uint64_t* out, const uint64_t* a, const uint64_t* b,
                                      int batch, int N, uint64_t prime,
                                      PolyMulStats* stats = nullptr) {
  poly_mul_u64_gpuntt_square(out, a, b, batch, N, prime, stats);
}

// ---- common/aes_gpu.h ---------------------------------------------------------
inline void aes128_expand_key_host(const uint8_t key[16], uint8_t round_keys[176]) {
  detail::check(pcgsq_aes128_expand_key(key, round_keys));
}
inline void aes128_ecb_encrypt_gpu(const uint8_t* keys, const uint8_t* in, uint8_t* out,
                                   int count) {
  std::lock_guard<std::mutex> lk(detail::mu());
  detail::check(pcgsq_aes128_ecb_encrypt(detail::ctx(), keys, in, out, count));
}

// ---- CPU_baseline/ntt.h: NTT<P> on the device ------------------------------------
// Same interface as the reference's NTT<P> (ntt.h:23-90) for the project prime,
// over raw residues (FFp<P> stores exactly one uint64 residue).
template <uint64_t P>
class NTT {
  static_assert(P == PCGSQ_PRIME, "pcg_hip::NTT supports P = 2^62 - 3*2^25 + 1 only");

 public:
  explicit NTT(int N) : N_(N) {
    std::string why;
    if (!detail::supported(P, N, PCGSQ_NTT_FOURSTEP, &why))
      throw std::runtime_error("pcg_hip::NTT: " + why);
  }
  void forward(uint64_t* a) const { run(a, false); }
  void inverse(uint64_t* a) const { run(a, true); }
  void poly_mul(uint64_t* c, const uint64_t* a, const uint64_t* b) const {
    detail::poly_mul(c, a, b, 1, N_, P, PCGSQ_NTT_FOURSTEP, nullptr);
  }
  int N() const { return N_; }

 private:
  void run(uint64_t* a, bool inv) const {
    std::lock_guard<std::mutex> lk(detail::mu());
    detail::check(inv ? pcgsq_ntt_inverse(detail::ctx(), a, 1, N_, PCGSQ_NTT_FOURSTEP)
                      : pcgsq_ntt_forward(detail::ctx(), a, 1, N_, PCGSQ_NTT_FOURSTEP));
  }
  int N_;
};

// ---- PCG-OLE expansion for both parties on one device ----------------------------
// The local compute of pcg_ole::PCG_OLE<P>::gen_and_expand
// (pcg_ole_impl.h:46-490) for both parties at once, with the 2PC
// preprocessing replaced by the seeded dealer.
class OleExpander {
 public:
  OleExpander(int N, int c, int t, int w = 1, DpfGpuPrg prg = DpfGpuPrg::CHACHA8,
              uint64_t pub_seed = 0, size_t hbm_budget = 0)
      : N_(N) {
    pcgsq_params p{N, c, t, w, detail::prg_code(prg)};
    std::lock_guard<std::mutex> lk(detail::mu());
    detail::check(pcgsq_ole_plan_create(detail::ctx(), &p, pub_seed, hbm_budget, &plan_));
  }
  ~OleExpander() {
    if (plan_) pcgsq_ole_plan_destroy(plan_);
  }
  OleExpander(const OleExpander&) = delete;
  OleExpander& operator=(const OleExpander&) = delete;

  struct Output {
    std::vector<uint64_t> x0, z0, x1, z1;   // z0 + z1 = x0 * x1 in R_p
  };
  // pcg_ole::generate_triples (pcg_ole_2pc/triple.h:45-110) for both parties:
  // N Z_p triple shares each, from the expansions of seed1 and seed2.
  struct Triples {
    std::vector<uint64_t> a0, b0, c0, a1, b1, c1;  // c0 + c1 = (a0 + a1)(b0 + b1) per slot
  };
  Triples triples(uint64_t seed1, uint64_t seed2) {
    Triples t;
    for (auto* v : {&t.a0, &t.b0, &t.c0, &t.a1, &t.b1, &t.c1}) v->resize(N_);
    std::lock_guard<std::mutex> lk(detail::mu());
    detail::check(pcgsq_triples_generate(plan_, seed1, seed2, t.a0.data(), t.b0.data(),
                                         t.c0.data(), t.a1.data(), t.b1.data(), t.c1.data()));
    return t;
  }

  // gmw: run the two parties' 2PC preprocessing (pcg_ole_impl.h:114-199) in
  // this process instead of taking DltSft, beta and rs from the dealer: base
  // OTs and IKNP random OTs, the position adder and DltSft circuits under
  // GMW with AND triples from the OTs, and Gilboa OT products for beta and
  // the Beaver triples; the outputs are the same.
  Output expand(uint64_t seed, bool gmw = false) {
    Output o;
    o.x0.resize(N_); o.z0.resize(N_); o.x1.resize(N_); o.z1.resize(N_);
    std::lock_guard<std::mutex> lk(detail::mu());
    detail::check(gmw ? pcgsq_ole_deal_gmw(plan_, seed) : pcgsq_ole_deal(plan_, seed));
    detail::check(pcgsq_ole_run(plan_));
    detail::check(pcgsq_ole_outputs(plan_, o.x0.data(), o.z0.data(), o.x1.data(), o.z1.data()));
    return o;
  }

 private:
  int N_;
  pcgsq_ole_plan* plan_ = nullptr;
};

// One party of the 2PC expansion (pcg_ole_impl.h:46-490), as one party's
// process runs it with emp::NetIO replaced by `exchange`: every message of
// the protocol goes through it, in the reference's order, with the columns
// of the reference's comm_trace.h rows (phase, level, batch, bytes sent,
// bytes received).  `exchange` sends send_bytes from `send` and receives
// recv_bytes from the other party into `recv` (either may be 0; party 0
// sends first).  An exception thrown by `exchange` aborts the expansion and
// reaches the caller unchanged.  Each party owns its own device context, so
// the two parties may run at once from two threads of one process, or as two
// processes; the outputs of the two together equal OleExpander's.
class OleParty {
 public:
  using ExchangeFn = std::function<void(const char* phase, int level, int batch, const void* send,
                                        size_t send_bytes, void* recv, size_t recv_bytes)>;
  struct Output {
    std::vector<uint64_t> x, z;   // this party's shares: z0 + z1 = x0 * x1 in R_p
  };

  OleParty(int party, int N, int c, int t, int w = 1, DpfGpuPrg prg = DpfGpuPrg::CHACHA8,
           uint64_t pub_seed = 0, size_t hbm_budget = 0, int device = 0)
      : N_(N) {
    detail::check(pcgsq_ctx_create(device, &ctx_));
    pcgsq_params p{N, c, t, w, detail::prg_code(prg)};
    const int rc = pcgsq_ole_party_create(ctx_, &p, party, pub_seed, hbm_budget, &plan_);
    if (rc != PCGSQ_OK) {
      const std::string why = pcgsq_last_error();
      pcgsq_ctx_destroy(ctx_);
      throw std::runtime_error("pcg_hip: " + why);
    }
  }
  ~OleParty() {
    if (plan_) pcgsq_ole_plan_destroy(plan_);
    if (ctx_) pcgsq_ctx_destroy(ctx_);
  }
  OleParty(const OleParty&) = delete;
  OleParty& operator=(const OleParty&) = delete;

  // The reference's deployment: this party's noise, Delta share and Beaver
  // randomness from its own seed; alpha, DltSft, beta and rs computed with
  // the other party (base OTs on the first call, then OT extension, GMW and
  // Gilboa products: "base_ot", "ot_ext", "gilboa", "circuit_and" rows before
  // the expansion's own).
  Output expand_own(uint64_t own_seed, const ExchangeFn& exchange) {
    detail::check(pcgsq_ole_party_deal_own(plan_, own_seed));
    return run(exchange);
  }
  // Trusted-dealer simulation: both parties' shares from one shared seed
  // (either party could recompute the other's), the run starting at the DPF.
  Output expand(uint64_t seed, const ExchangeFn& exchange) {
    detail::check(pcgsq_ole_deal(plan_, seed));
    return run(exchange);
  }
  // One exchange per DPF level, Beaver open and circuit depth for a whole
  // group of DPF blocks (the reference's batched arm) instead of per block.
  void set_batched(bool batched) {
    detail::check(pcgsq_ole_party_set_exchange_mode(
        plan_, batched ? PCGSQ_EXCHANGE_BATCHED : PCGSQ_EXCHANGE_PER_BLOCK));
  }

 private:
  struct Call {
    const ExchangeFn* fn;
    std::exception_ptr err;
  };
  Output run(const ExchangeFn& exchange) {
    Call call{&exchange, nullptr};
    const int rc = pcgsq_ole_party_run(plan_, tramp, &call);
    if (call.err) std::rethrow_exception(call.err);
    detail::check(rc);
    Output o;
    o.x.resize(N_);
    o.z.resize(N_);
    detail::check(pcgsq_ole_party_outputs(plan_, o.x.data(), o.z.data()));
    return o;
  }
  static int tramp(void* user, const char* phase, int level, int batch, const void* send,
                   size_t send_bytes, void* recv, size_t recv_bytes) {
    Call* c = static_cast<Call*>(user);
    try {
      (*c->fn)(phase, level, batch, send, send_bytes, recv, recv_bytes);
      return 0;
    } catch (...) {
      c->err = std::current_exception();
      return 1;
    }
  }
  int N_;
  pcgsq_ctx* ctx_ = nullptr;
  pcgsq_ole_plan* plan_ = nullptr;
};

}  // namespace pcg_hip

#endif  // PCGSQ_PCG_HIP_HPP
