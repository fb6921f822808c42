// NTT kernels and plans (see ntt.hpp).
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "ntt.hpp"
#include "ntt_pass.hpp"

namespace pcgsq {

namespace {

constexpr int kPassThreads = 512;
constexpr int kEwThreads = 256;

inline int brv_bits(uint64_t x, int bits) {
  uint64_t r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1ULL) << (bits - 1 - i);
  return (int)r;
}

__device__ __forceinline__ int brv_dev(int x, int bits) {
  return (int)(__brev((unsigned)x) >> (32 - bits));
}

// One four-step (or single-pass) NTT pass as a grid of tiles.  Tiles of one
// transform are dealt to the eight XCDs in contiguous runs, so tiles that
// share 128-byte lines (C < 16 columns of 8 bytes) run under one L2.
template <int K, int RA, int RB, int RC>
__global__ __launch_bounds__(kPassThreads) void ntt_pass_kernel(NttPassDesc d, uint32_t tiles) {
  extern __shared__ uint64_t lds[];
  uint32_t bid = blockIdx.x;
  if ((tiles & 7) == 0) bid = (bid & 7) * (tiles >> 3) + (bid >> 3);
  nttpass::pass_tile<K, RA, RB, RC>(d, lds, blockIdx.y, bid);
}

// ---- naive per-stage kernels ------------------------------------------------

__global__ void twist_kernel(uint64_t* a, const uint64_t* tw, int N_log, size_t total) {
  const size_t i = (size_t)blockIdx.x * kEwThreads + threadIdx.x;
  if (i >= total) return;
  const size_t k = i & ((1ULL << N_log) - 1);
  a[i] = shoup(a[i], tw[2 * k], tw[2 * k + 1]);
}

// DIF stage with half-size m: X[i] = u + v, X[i+m] = (u - v) * w^(off*N/2m).
__global__ void dif_stage_kernel(uint64_t* a, const uint64_t* w, int N_log, int m_log,
                                 size_t pairs) {
  const size_t p = (size_t)blockIdx.x * kEwThreads + threadIdx.x;
  if (p >= pairs) return;
  const size_t half = 1ULL << (N_log - 1);
  const size_t b = p >> (N_log - 1), q = p & (half - 1);
  const size_t m = 1ULL << m_log;
  const size_t blk = q >> m_log, off = q & (m - 1);
  const size_t i = (b << N_log) + blk * 2 * m + off;
  const uint64_t u = a[i], v = a[i + m];
  a[i] = fadd(u, v);
  const size_t wi = off << (N_log - 1 - m_log);
  a[i + m] = shoup(fsub(u, v), w[2 * wi], w[2 * wi + 1]);
}

// DIT stage (mirror of the DIF stage): t = a[i+m] w^-(...), a[i] = u + t,
// a[i+m] = u - t.
__global__ void dit_stage_kernel(uint64_t* a, const uint64_t* winv, int N_log, int m_log,
                                 size_t pairs) {
  const size_t p = (size_t)blockIdx.x * kEwThreads + threadIdx.x;
  if (p >= pairs) return;
  const size_t half = 1ULL << (N_log - 1);
  const size_t b = p >> (N_log - 1), q = p & (half - 1);
  const size_t m = 1ULL << m_log;
  const size_t blk = q >> m_log, off = q & (m - 1);
  const size_t i = (b << N_log) + blk * 2 * m + off;
  const size_t wi = off << (N_log - 1 - m_log);
  const uint64_t u = a[i];
  const uint64_t t = shoup(a[i + m], winv[2 * wi], winv[2 * wi + 1]);
  a[i] = fadd(u, t);
  a[i + m] = fsub(u, t);
}

__global__ void bitrev_kernel(const uint64_t* in, uint64_t* out, int N_log, size_t total) {
  const size_t i = (size_t)blockIdx.x * kEwThreads + threadIdx.x;
  if (i >= total) return;
  const size_t b = i >> N_log, k = i & ((1ULL << N_log) - 1);
  out[(b << N_log) + brv_dev((int)k, N_log)] = in[i];
}

__global__ void pointwise_kernel(uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n) {
  const size_t i = (size_t)blockIdx.x * kEwThreads + threadIdx.x;
  if (i >= n) return;
  out[i] = fmul(a[i], b[i]);
}

std::vector<uint64_t> with_companions(const std::vector<uint64_t>& w) {
  std::vector<uint64_t> t(2 * w.size());
  for (size_t i = 0; i < w.size(); ++i) {
    t[2 * i] = w[i];
    t[2 * i + 1] = shoup_companion(w[i]);
  }
  return t;
}

void upload(DevBuf& buf, const std::vector<uint64_t>& v, const char* what) {
  buf.ensure(v.size() * sizeof(uint64_t), what);
  hip_check(hipMemcpy(buf.ptr, v.data(), v.size() * sizeof(uint64_t), hipMemcpyHostToDevice),
            what);
}

unsigned grid_for(size_t n) { return (unsigned)((n + kEwThreads - 1) / kEwThreads); }

}  // namespace

struct NttPlan {
  int N = 0, L = 0, n1_log = 0, n2_log = 0, tw_lo_bits = 0;
  int variant = 0;           // bit 0: alternative stage grouping of the n1 passes, bit 1: n2
  int cols1 = 0, cols2 = 0;  // columns per tile of the n1 / n2 passes (0 = automatic)
  uint64_t psi = 0, ninv = 0, ninv_q = 0;
  DevBuf n1_neg_fwd, n1_neg_inv, n2_cyc_fwd, n2_cyc_inv;
  DevBuf tw_fwd_lo, tw_fwd_hi, tw_inv_lo, tw_inv_hi;
  bool naive_ready = false;
  DevBuf nv_w, nv_winv, nv_twf, nv_twi;

  bool single() const { return n2_log == 0; }

  // n1_log == L: one pass (whole transform in one tile); otherwise the
  // four-step split N = n1 * n2.
  NttPlan(int N_, int n1_log_) : N(N_) {
    L = ilog2_exact((uint64_t)N);
    n1_log = n1_log_;
    n2_log = L - n1_log;
    psi = hpow(3, (kPrime - 1) / (2 * (uint64_t)N));
    const uint64_t psi_inv = hinv(psi);
    ninv = hinv((uint64_t)N);
    ninv_q = shoup_companion(ninv);
    const int n1 = 1 << n1_log, n2 = 1 << n2_log;
    // length-n1 negacyclic sub-transform, root psi1 = psi^n2 (order 2*n1)
    const uint64_t psi1 = hpow(psi, (uint64_t)n2), psi1_inv = hinv(psi1);
    std::vector<uint64_t> f(n1, 0), fi(n1, 0);
    for (int k = 1; k < n1; ++k) {
      f[k] = hpow(psi1, (uint64_t)brv_bits(k, n1_log));
      fi[k] = hpow(psi1_inv, (uint64_t)brv_bits(k, n1_log));
    }
    upload(n1_neg_fwd, with_companions(f), "ntt tables");
    upload(n1_neg_inv, with_companions(fi), "ntt tables");
    if (single()) return;
    // length-n2 cyclic sub-transform, root w2 = psi^(2 n1) (order n2)
    const uint64_t w2 = hpow(psi, 2 * (uint64_t)n1), w2_inv = hinv(w2);
    std::vector<uint64_t> g(n2, 0), gi(n2, 0);
    for (int k = 1; k < n2; ++k) {
      int s = 0;
      while ((2 << s) <= k) ++s;
      const int m = 1 << s, i = k - m;
      const uint64_t e = (uint64_t)brv_bits(i, s) * (uint64_t)(n2 / (2 * m));
      g[k] = hpow(w2, e);
      gi[k] = hpow(w2_inv, e);
    }
    upload(n2_cyc_fwd, with_companions(g), "ntt tables");
    upload(n2_cyc_inv, with_companions(gi), "ntt tables");
    // two-level psi^(+-e), e < 2N; the inverse high table carries N^-1
    tw_lo_bits = (L + 2) / 2;
    const int hi_bits = L + 1 - tw_lo_bits;
    std::vector<uint64_t> lo(1 << tw_lo_bits), hi(1 << hi_bits), loi(1 << tw_lo_bits),
        hii(1 << hi_bits);
    for (size_t x = 0; x < lo.size(); ++x) {
      lo[x] = hpow(psi, x);
      loi[x] = hpow(psi_inv, x);
    }
    for (size_t y = 0; y < hi.size(); ++y) {
      hi[y] = hpow(psi, y << tw_lo_bits);
      hii[y] = hmul(hpow(psi_inv, y << tw_lo_bits), ninv);
    }
    upload(tw_fwd_lo, with_companions(lo), "ntt twiddles");
    upload(tw_fwd_hi, with_companions(hi), "ntt twiddles");
    upload(tw_inv_lo, with_companions(loi), "ntt twiddles");
    upload(tw_inv_hi, with_companions(hii), "ntt twiddles");
  }

  void ensure_naive() {
    if (naive_ready) return;
    const uint64_t omega = hmul(psi, psi), omega_inv = hinv(omega);
    const uint64_t psi_inv = hinv(psi);
    std::vector<uint64_t> w(std::max(N / 2, 1)), wi(std::max(N / 2, 1)), tf(N), ti(N);
    w[0] = wi[0] = 1;
    for (int k = 1; k < N / 2; ++k) {
      w[k] = hmul(w[k - 1], omega);
      wi[k] = hmul(wi[k - 1], omega_inv);
    }
    tf[0] = 1;
    ti[0] = ninv;
    for (int k = 1; k < N; ++k) {
      tf[k] = hmul(tf[k - 1], psi);
      ti[k] = hmul(ti[k - 1], psi_inv);
    }
    upload(nv_w, with_companions(w), "naive tables");
    upload(nv_winv, with_companions(wi), "naive tables");
    upload(nv_twf, with_companions(tf), "naive tables");
    upload(nv_twi, with_companions(ti), "naive tables");
    naive_ready = true;
  }
};

struct NttEngine::Impl {
  std::mutex mu;
  std::map<int, std::unique_ptr<NttPlan>> plans;
  DevBuf scratch[3];
};

NttEngine::NttEngine() : impl_(new Impl) {}
NttEngine::~NttEngine() = default;

bool NttEngine::supported(int N, int backend, std::string* reason) {
  auto fail = [&](const char* why) {
    if (reason) *reason = why;
    return false;
  };
  if (N <= 0 || (N & (N - 1))) return fail("N must be a power of two");
  const int L = ilog2_exact((uint64_t)N);
  if (L > 24) return fail("N > 2^24: 2N does not divide P-1");
  if (backend != kNttFourStep && backend != kNttNaive) return fail("unknown backend");
  if (L < 1) return fail("N must be at least 2");
  return true;
}

int NttEngine::default_split(int L) { return L <= kNttSingleMax ? L : (L + 1) / 2; }

NttPlan& NttEngine::plan(int N) {
  std::lock_guard<std::mutex> lk(impl_->mu);
  auto& p = impl_->plans[N];
  if (!p) p.reset(new NttPlan(N, default_split(ilog2_exact((uint64_t)N))));
  return *p;
}

void NttEngine::tune(int N, int n1_log, int variant, int cols1, int cols2) {
  const int L = ilog2_exact((uint64_t)N);
  require(L >= 1 && L <= 24, "ntt tune: N must be 2^1..2^24");
  const int want = n1_log > 0 ? n1_log : default_split(L);
  require(want == L ? L <= kNttSingleMax
                    : (want >= 6 && L - want >= 6 && want <= kNttSingleMax &&
                       L - want <= kNttSingleMax),
          "ntt tune: one pass needs N <= 2^13; a split needs passes of length 2^6..2^13");
  require(variant >= 0 && variant <= 3, "ntt tune: variant must be 0..3");
  auto okc = [](int c) { return c == 0 || (c >= 1 && c <= 64 && (c & (c - 1)) == 0); };
  require(okc(cols1) && okc(cols2), "ntt tune: cols must be 0 or a power of two <= 64");
  std::lock_guard<std::mutex> lk(impl_->mu);
  auto& p = impl_->plans[N];
  if (!p || p->n1_log != want) p.reset(new NttPlan(N, want));
  p->variant = variant;
  p->cols1 = cols1;
  p->cols2 = cols2;
}

uint64_t* NttEngine::scratch(size_t words, int which) {
  impl_->scratch[which].ensure(words * sizeof(uint64_t), "hipMalloc ntt scratch");
  return impl_->scratch[which].as<uint64_t>();
}

namespace {

size_t g_lds_max = 0;  // bytes of LDS one workgroup may use

size_t lds_max() {
  if (g_lds_max == 0) {
    int dev = 0;
    size_t b = 65536;
    int v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) ==
            hipSuccess &&
        v > 0)
      b = (size_t)v;
    g_lds_max = std::min<size_t>(b, 160 * 1024);
  }
  return g_lds_max;
}

struct PassKernel {
  const void* fn;
  int nlog, rmax;
};

template <int K, int RA, int RB, int RC>
PassKernel pk() {
  using P = nttpass::Part<RA, RB, RC>;
  return PassKernel{(const void*)ntt_pass_kernel<K, RA, RB, RC>, P::NLOG, P::RMAX};
}

// Stage grouping per pass length (variant 1 = more, shorter groups).  Four-step
// passes are 6..13 long; whole-transform passes 1..13.
template <int K>
PassKernel pass_kernel(int nlog, int variant) {
  constexpr bool whole = K == kFwdWhole || K == kInvWhole || K == kInvWholeProd;
  if constexpr (whole) {
    switch (nlog) {
      case 1: return pk<K, 1, 0, 0>();
      case 2: return pk<K, 2, 0, 0>();
      case 3: return pk<K, 3, 0, 0>();
      case 4: return pk<K, 4, 0, 0>();
      case 5: return pk<K, 5, 0, 0>();
      default: break;
    }
  }
  switch (nlog) {
    case 6: return pk<K, 3, 3, 0>();
    case 7: return pk<K, 4, 3, 0>();
    case 8: return pk<K, 4, 4, 0>();
    case 9: return variant ? pk<K, 3, 3, 3>() : pk<K, 5, 4, 0>();
    case 10: return variant ? pk<K, 4, 3, 3>() : pk<K, 5, 5, 0>();
    case 11: return pk<K, 4, 4, 3>();
    case 12: return pk<K, 4, 4, 4>();
    case 13: return pk<K, 5, 4, 4>();
    default: throw Error(kErrUnsupported, "ntt: pass length out of range");
  }
}

// Columns per tile: requested, else about 256 threads; bounded by 512
// threads, two tiles' LDS per CU and the column count.
int tile_cols(int want, int nlog, int rmax, int ncols) {
  int C = want > 0 ? want : std::max(1, 256 >> (nlog - rmax));
  if (want <= 0) C = std::min(C, 16);
  while (C > 1 && (C << (nlog - rmax)) > kPassThreads) C >>= 1;
  while (C > 1 && ntt_tile_lds_bytes(nlog, C) > lds_max() / (want > 0 ? 1 : 2)) C >>= 1;
  while (C > 1 && C > ncols) C >>= 1;
  return C;
}

template <int K>
void launch_pass(hipStream_t st, NttPassDesc d, int nlog, int variant, int want_cols, int ncols,
                 int batch, Profiler* prof, const char* zone) {
  ProfZone pz(prof, st, zone);
  const PassKernel k = pass_kernel<K>(nlog, variant);
  const int C = tile_cols(want_cols, nlog, k.rmax, ncols);
  d.c_log = ilog2_exact((uint64_t)C);
  d.ncols = ncols;
  const uint32_t tiles = (uint32_t)((ncols + C - 1) / C);
  const size_t lds = ntt_tile_lds_bytes(nlog, C);
  require(lds <= lds_max(), "ntt: tile exceeds LDS");
  if (lds > 65536) {
    static std::mutex mu;
    static std::vector<const void*> raised;
    std::lock_guard<std::mutex> lk(mu);
    if (std::find(raised.begin(), raised.end(), k.fn) == raised.end()) {
      hip_check(hipFuncSetAttribute(k.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds_max()),
                "ntt: raise LDS limit");
      raised.push_back(k.fn);
    }
  }
  uint32_t tl = tiles;
  void* args[] = {&d, &tl};
  hip_check(hipLaunchKernel(k.fn, dim3(tiles, (unsigned)batch), dim3((unsigned)(C << (nlog - k.rmax))),
                            args, lds, st),
            "launch ntt_pass_kernel");
}

NttPassDesc base_desc(const NttPlan& pl) {
  NttPassDesc d{};
  d.N_log = pl.L;
  d.tw_lo_bits = pl.tw_lo_bits;
  d.scale_w = pl.ninv;
  d.scale_wq = pl.ninv_q;
  return d;
}

void four_step_forward(hipStream_t st, NttPlan& pl, const uint64_t* in, uint64_t* tmp,
                       uint64_t* out, int batch, Profiler* prof) {
  if (pl.single()) {
    // whole transforms as columns: one HBM round trip
    NttPassDesc a = base_desc(pl);
    a.in = in; a.out = out;
    a.tab = pl.n1_neg_fwd.as<uint64_t>();
    a.in_ps_log = 0; a.in_cs_log = pl.L; a.out_ps_log = 0; a.out_cs_log = pl.L;
    launch_pass<kFwdWhole>(st, a, pl.L, pl.variant & 1, pl.cols1, batch, 1, prof,
                           "ntt.fwd_single");
    return;
  }
  // pass 1: columns i2, length-n1 negacyclic over i1, twiddle psi^(i2(2k1+1))
  NttPassDesc a = base_desc(pl);
  a.in = in; a.out = tmp;
  a.tab = pl.n1_neg_fwd.as<uint64_t>();
  a.tw_lo = pl.tw_fwd_lo.as<uint64_t>(); a.tw_hi = pl.tw_fwd_hi.as<uint64_t>();
  a.in_ps_log = pl.n2_log; a.in_cs_log = 0; a.out_ps_log = pl.n2_log; a.out_cs_log = 0;
  launch_pass<kFwdCols>(st, a, pl.n1_log, pl.variant & 1, pl.cols1, 1 << pl.n2_log, batch, prof,
                     "ntt.fwd_pass1");
  // pass 2: rows k1, length-n2 cyclic over i2, written transposed (natural k)
  NttPassDesc b = base_desc(pl);
  b.in = tmp; b.out = out;
  b.tab = pl.n2_cyc_fwd.as<uint64_t>();
  b.in_ps_log = 0; b.in_cs_log = pl.n2_log; b.out_ps_log = pl.n1_log; b.out_cs_log = 0;
  launch_pass<kFwdRows>(st, b, pl.n2_log, (pl.variant >> 1) & 1, pl.cols2, 1 << pl.n1_log, batch,
                     prof, "ntt.fwd_pass2");
}

void four_step_inverse(hipStream_t st, NttPlan& pl, const uint64_t* in, const uint64_t* in2,
                       uint64_t* tmp, uint64_t* out, int batch, Profiler* prof) {
  if (pl.single()) {
    NttPassDesc a = base_desc(pl);
    a.in = in; a.in2 = in2; a.out = out;
    a.tab = pl.n1_neg_inv.as<uint64_t>();
    a.in_ps_log = 0; a.in_cs_log = pl.L; a.out_ps_log = 0; a.out_cs_log = pl.L;
    if (in2)
      launch_pass<kInvWholeProd>(st, a, pl.L, pl.variant & 1, pl.cols1, batch, 1, prof,
                                 "ntt.inv_single");
    else
      launch_pass<kInvWhole>(st, a, pl.L, pl.variant & 1, pl.cols1, batch, 1, prof,
                             "ntt.inv_single");
    return;
  }
  // pass I1: columns k1 of the [k2][k1] input, inverse cyclic over k2,
  // twiddle psi^-(i2(2k1+1)) N^-1, rows k1 out
  NttPassDesc a = base_desc(pl);
  a.in = in; a.in2 = in2; a.out = tmp;
  a.tab = pl.n2_cyc_inv.as<uint64_t>();
  a.tw_lo = pl.tw_inv_lo.as<uint64_t>(); a.tw_hi = pl.tw_inv_hi.as<uint64_t>();
  a.in_ps_log = pl.n1_log; a.in_cs_log = 0; a.out_ps_log = 0; a.out_cs_log = pl.n2_log;
  if (in2)
    launch_pass<kInvRowsProd>(st, a, pl.n2_log, (pl.variant >> 1) & 1, pl.cols2,
                              1 << pl.n1_log, batch, prof, "ntt.inv_pass1");
  else
    launch_pass<kInvRows>(st, a, pl.n2_log, (pl.variant >> 1) & 1, pl.cols2, 1 << pl.n1_log,
                          batch, prof, "ntt.inv_pass1");
  // pass I2: columns i2, inverse negacyclic over k1
  NttPassDesc b = base_desc(pl);
  b.in = tmp; b.out = out;
  b.tab = pl.n1_neg_inv.as<uint64_t>();
  b.in_ps_log = pl.n2_log; b.in_cs_log = 0; b.out_ps_log = pl.n2_log; b.out_cs_log = 0;
  launch_pass<kInvCols>(st, b, pl.n1_log, pl.variant & 1, pl.cols1, 1 << pl.n2_log, batch, prof,
                    "ntt.inv_pass2");
}

void naive_forward_brv(hipStream_t st, NttPlan& pl, uint64_t* d, int batch) {
  pl.ensure_naive();
  const size_t total = (size_t)batch * pl.N, pairs = total / 2;
  hipLaunchKernelGGL(twist_kernel, dim3(grid_for(total)), dim3(kEwThreads), 0, st, d,
                     pl.nv_twf.as<uint64_t>(), pl.L, total);
  for (int m_log = pl.L - 1; m_log >= 0; --m_log)
    hipLaunchKernelGGL(dif_stage_kernel, dim3(grid_for(pairs)), dim3(kEwThreads), 0, st, d,
                       pl.nv_w.as<uint64_t>(), pl.L, m_log, pairs);
  hip_check(hipGetLastError(), "launch naive forward");
}

void naive_inverse_brv(hipStream_t st, NttPlan& pl, uint64_t* d, int batch) {
  pl.ensure_naive();
  const size_t total = (size_t)batch * pl.N, pairs = total / 2;
  for (int m_log = 0; m_log < pl.L; ++m_log)
    hipLaunchKernelGGL(dit_stage_kernel, dim3(grid_for(pairs)), dim3(kEwThreads), 0, st, d,
                       pl.nv_winv.as<uint64_t>(), pl.L, m_log, pairs);
  hipLaunchKernelGGL(twist_kernel, dim3(grid_for(total)), dim3(kEwThreads), 0, st, d,
                     pl.nv_twi.as<uint64_t>(), pl.L, total);
  hip_check(hipGetLastError(), "launch naive inverse");
}

}  // namespace

void launch_pointwise_mul(hipStream_t st, uint64_t* out, const uint64_t* a, const uint64_t* b,
                          size_t n) {
  hipLaunchKernelGGL(pointwise_kernel, dim3(grid_for(n)), dim3(kEwThreads), 0, st, out, a, b, n);
  hip_check(hipGetLastError(), "launch pointwise");
}

void NttEngine::forward(hipStream_t st, uint64_t* d, int batch, int N, int backend) {
  std::string why;
  require(supported(N, backend, &why), "ntt forward: " + why, kErrUnsupported);
  require(batch >= 1 && batch <= 65535, "ntt: batch out of range");
  NttPlan& pl = plan(N);
  const size_t total = (size_t)batch * N;
  if (backend == kNttFourStep) {
    uint64_t* tmp = scratch(total, 0);
    four_step_forward(st, pl, d, tmp, d, batch, prof);
  } else {
    naive_forward_brv(st, pl, d, batch);
    uint64_t* tmp = scratch(total, 0);
    hipLaunchKernelGGL(bitrev_kernel, dim3(grid_for(total)), dim3(kEwThreads), 0, st, d, tmp,
                       pl.L, total);
    hip_check(hipMemcpyAsync(d, tmp, total * 8, hipMemcpyDeviceToDevice, st), "ntt copy");
  }
}

void NttEngine::inverse(hipStream_t st, uint64_t* d, int batch, int N, int backend) {
  std::string why;
  require(supported(N, backend, &why), "ntt inverse: " + why, kErrUnsupported);
  require(batch >= 1 && batch <= 65535, "ntt: batch out of range");
  NttPlan& pl = plan(N);
  const size_t total = (size_t)batch * N;
  if (backend == kNttFourStep) {
    uint64_t* tmp = scratch(total, 0);
    four_step_inverse(st, pl, d, nullptr, tmp, d, batch, prof);
  } else {
    uint64_t* tmp = scratch(total, 0);
    hipLaunchKernelGGL(bitrev_kernel, dim3(grid_for(total)), dim3(kEwThreads), 0, st, d, tmp,
                       pl.L, total);
    hip_check(hipMemcpyAsync(d, tmp, total * 8, hipMemcpyDeviceToDevice, st), "ntt copy");
    naive_inverse_brv(st, pl, d, batch);
  }
}

void NttEngine::inverse_of_product(hipStream_t st, uint64_t* out, const uint64_t* x,
                                   const uint64_t* y, int batch, int N) {
  NttPlan& pl = plan(N);
  uint64_t* tmp = scratch((size_t)batch * N, 0);
  four_step_inverse(st, pl, x, y, tmp, out, batch, prof);
}

void NttEngine::poly_mul(hipStream_t st, uint64_t* out, const uint64_t* a, const uint64_t* b,
                         int batch, int N, int backend) {
  std::string why;
  require(supported(N, backend, &why), "poly_mul: " + why, kErrUnsupported);
  require(batch >= 1 && batch <= 65535, "poly_mul: batch out of range");
  NttPlan& pl = plan(N);
  const size_t total = (size_t)batch * N;
  uint64_t* fa = scratch(total, 1);
  uint64_t* fb = scratch(total, 2);
  if (backend == kNttFourStep) {
    uint64_t* tmp = scratch(total, 0);
    four_step_forward(st, pl, a, tmp, fa, batch, prof);
    four_step_forward(st, pl, b, tmp, fb, batch, prof);
    four_step_inverse(st, pl, fa, fb, tmp, out, batch, prof);
  } else {
    // textbook: twist, DIF stages, pointwise in the bit-reversed domain,
    // DIT stages, twist back (naive_ntt_bench.cu:147-159)
    hip_check(hipMemcpyAsync(fa, a, total * 8, hipMemcpyDeviceToDevice, st), "copy");
    hip_check(hipMemcpyAsync(fb, b, total * 8, hipMemcpyDeviceToDevice, st), "copy");
    naive_forward_brv(st, pl, fa, batch);
    naive_forward_brv(st, pl, fb, batch);
    launch_pointwise_mul(st, fa, fa, fb, total);
    naive_inverse_brv(st, pl, fa, batch);
    hip_check(hipMemcpyAsync(out, fa, total * 8, hipMemcpyDeviceToDevice, st), "copy");
  }
}

}  // namespace pcgsq
