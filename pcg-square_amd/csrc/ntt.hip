// NTT kernels and plans (see ntt.hpp).
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "ntt.hpp"

namespace pcgsq {

namespace {

constexpr int kPassThreads = 1024;
constexpr int kEwThreads = 256;

inline int brv_bits(uint64_t x, int bits) {
  uint64_t r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1ULL) << (bits - 1 - i);
  return (int)r;
}

__device__ __forceinline__ int brv_dev(int x, int bits) {
  return (int)(__brev((unsigned)x) >> (32 - bits));
}

// Lazy-reduction helpers (Harvey): values live in [0, 4P) (< 2^64).
__device__ __forceinline__ uint64_t red2p(uint64_t x) { return x >= kTwoP ? x - kTwoP : x; }
__device__ __forceinline__ uint64_t red_full(uint64_t x) {  // [0,4P) -> [0,P)
  x = x >= kTwoP ? x - kTwoP : x;
  return x >= kP ? x - kP : x;
}

struct PassArgs {
  const uint64_t* in;
  const uint64_t* in2;      // product pass: second factor (same layout as in)
  uint64_t* out;
  const uint64_t* tab;      // sub-transform table: [n][2] = (w, w')
  const uint64_t* tw_lo;    // two-level twiddle tables [.][2] (or null)
  const uint64_t* tw_hi;
  int tw_lo_bits;
  int tw_swap;              // 0: e = c*(2p+1); 1: e = p*(2c+1)
  int n_log;                // sub-transform length n = 2^n_log
  int C;                    // sub-transforms per workgroup
  int ncols;                // sub-transforms per transform
  size_t in_ps, in_cs, out_ps, out_cs;  // element strides (position, column)
  int in_brv, out_brv;
  int N_log;                // whole transform length (for the twiddle modulus 2N)
};

// LDS element (c, e) of a tile of C sub-transforms of length n: column c is
// rotated by c words (the C lanes that load one row segment, same e and
// consecutive c, hit different banks) and the low 4 index bits are XORed with
// the next 4 (the 16 elements a thread gathers for one radix-16 set, which
// are 16 apart in the last stage group, land in different banks).
__device__ __forceinline__ int lidx(int c, int e, int n_log) {
  const int x = (e + c) & ((1 << n_log) - 1);
  return (c << n_log) + (x ^ ((x >> 4) & 15));
}

// One group of r consecutive butterfly stages [s0, s0 + r) of every
// sub-transform in the tile, done in registers: each thread gathers sets of
// 2^r elements spaced t_min = n >> (s0 + r) apart, runs r stages on them and
// writes them back.  Forward: Cooley-Tukey with merged roots, twiddle
// tab[2^s + i] on the upper input, values lazily reduced in [0, 4P).
// Inverse: Gentleman-Sande in the reverse stage order, values in [0, 2P).
template <bool INVERSE, int R>
__device__ __forceinline__ void stage_group(uint64_t* __restrict__ lds,
                                            const uint64_t* __restrict__ tab, int n_log, int C,
                                            int s0) {
  constexpr int M = 1 << R;
  const int tmin_log = n_log - (s0 + R);
  const int sets_log = n_log - R;
  const int nsets = C << sets_log;
  for (int gs = threadIdx.x; gs < nsets; gs += blockDim.x) {
    const int c = gs >> sets_log;
    const int sigma = gs & ((1 << sets_log) - 1);
    const int L = sigma & ((1 << tmin_log) - 1);
    const int Hh = sigma >> tmin_log;
    const int base = (Hh << (tmin_log + R)) | L;
    uint64_t v[M];
#pragma unroll
    for (int w = 0; w < M; ++w) v[w] = lds[lidx(c, base | (w << tmin_log), n_log)];
    if constexpr (!INVERSE) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int bit = R - 1 - q;
#pragma unroll
        for (int w = 0; w < M; ++w) {
          if ((w >> bit) & 1) continue;
          const int w2 = w | (1 << bit);
          const int ti = 2 * ((1 << (s0 + q)) + ((Hh << q) | (w >> (R - q))));
          const uint64_t U = red2p(v[w]);
          const uint64_t T = shoup_lazy(v[w2], tab[ti], tab[ti + 1]);
          v[w] = U + T;
          v[w2] = U - T + kTwoP;
        }
      }
    } else {
#pragma unroll
      for (int q = R - 1; q >= 0; --q) {
        const int bit = R - 1 - q;
#pragma unroll
        for (int w = 0; w < M; ++w) {
          if ((w >> bit) & 1) continue;
          const int w2 = w | (1 << bit);
          const int ti = 2 * ((1 << (s0 + q)) + ((Hh << q) | (w >> (R - q))));
          const uint64_t U = v[w], V = v[w2];   // both < 2P
          v[w] = red2p(U + V);
          v[w2] = shoup_lazy(U - V + kTwoP, tab[ti], tab[ti + 1]);
        }
      }
    }
#pragma unroll
    for (int w = 0; w < M; ++w) lds[lidx(c, base | (w << tmin_log), n_log)] = v[w];
  }
}

template <bool INVERSE, int R>
__device__ __forceinline__ void run_group(uint64_t* lds, const uint64_t* tab, int n_log, int C,
                                          int s0, int r) {
  // r <= R: the last group of a length that R does not divide
  if constexpr (R >= 1) {
    if (r == R) {
      stage_group<INVERSE, R>(lds, tab, n_log, C, s0);
      return;
    }
    if constexpr (R > 1) run_group<INVERSE, R - 1>(lds, tab, n_log, C, s0, r);
  }
}

// One pass of the four-step transform: C whole sub-transforms of length n are
// loaded into LDS, transformed in groups of R stages in registers, optionally
// multiplied by a twiddle and written back with arbitrary strides (which is
// where the transposition of the four-step happens).
template <bool INVERSE, bool TWIDDLE, bool PRODUCT, int R>
__global__ __launch_bounds__(kPassThreads) void ntt_pass_kernel(PassArgs a) {
  extern __shared__ uint64_t lds[];
  const int n = 1 << a.n_log;
  const int C = a.C;
  const size_t batch_off = (size_t)blockIdx.y << a.N_log;
  const int c0 = blockIdx.x * C;
  const uint64_t* in = a.in + batch_off;
  const uint64_t* in2 = PRODUCT ? a.in2 + batch_off : nullptr;
  uint64_t* out = a.out + batch_off;
  const int tile = n * C;

  // ---- load (coalesced along whichever stride is 1) ----
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    int c, p;
    if (a.in_cs == 1) { c = e % C; p = e / C; } else { c = e / n; p = e % n; }
    const size_t addr = (size_t)p * a.in_ps + (size_t)(c0 + c) * a.in_cs;
    uint64_t v = in[addr];
    if constexpr (PRODUCT) v = fmul(v, in2[addr]);
    const int q = a.in_brv ? brv_dev(p, a.n_log) : p;
    lds[lidx(c, q, a.n_log)] = v;
  }
  __syncthreads();

  const int ngroups = (a.n_log + R - 1) / R;
  if constexpr (!INVERSE) {
    for (int gi = 0; gi < ngroups; ++gi) {
      const int s0 = gi * R;
      const int r = a.n_log - s0 < R ? a.n_log - s0 : R;
      run_group<false, R>(lds, a.tab, a.n_log, C, s0, r);
      __syncthreads();
    }
  } else {
    for (int gi = ngroups - 1; gi >= 0; --gi) {
      const int s0 = gi * R;
      const int r = a.n_log - s0 < R ? a.n_log - s0 : R;
      run_group<true, R>(lds, a.tab, a.n_log, C, s0, r);
      __syncthreads();
    }
  }

  // ---- store ----
  const uint64_t two_n_mask = (2ULL << a.N_log) - 1;
  const uint64_t lo_mask = (1ULL << a.tw_lo_bits) - 1;
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    int c, p;
    if (a.out_cs == 1) { c = e % C; p = e / C; } else { c = e / n; p = e % n; }
    const int q = a.out_brv ? brv_dev(p, a.n_log) : p;
    uint64_t v = lds[lidx(c, q, a.n_log)];
    if constexpr (TWIDDLE) {
      const uint64_t cc = (uint64_t)(c0 + c), pp = (uint64_t)p;
      const uint64_t ex = (a.tw_swap ? pp * (2 * cc + 1) : cc * (2 * pp + 1)) & two_n_mask;
      const uint64_t xl = ex & lo_mask, xh = ex >> a.tw_lo_bits;
      v = shoup_lazy(v, a.tw_lo[2 * xl], a.tw_lo[2 * xl + 1]);
      v = shoup_lazy(v, a.tw_hi[2 * xh], a.tw_hi[2 * xh + 1]);
    }
    out[(size_t)p * a.out_ps + (size_t)(c0 + c) * a.out_cs] = red_full(v);
  }
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
  int radix = 4;       // butterfly stages per register group in a pass
  int cols_max = 16;   // sub-transforms per workgroup (LDS tile width)
  uint64_t psi = 0;
  DevBuf n1_neg_fwd, n1_neg_inv, n2_cyc_fwd, n2_cyc_inv;
  DevBuf tw_fwd_lo, tw_fwd_hi, tw_inv_lo, tw_inv_hi;
  bool naive_ready = false;
  DevBuf nv_w, nv_winv, nv_twf, nv_twi;

  NttPlan(int N_, int n1_log_) : N(N_) {
    L = ilog2_exact((uint64_t)N);
    n1_log = n1_log_ > 0 ? n1_log_ : (L + 1) / 2;
    n2_log = L - n1_log;
    psi = hpow(3, (kPrime - 1) / (2 * (uint64_t)N));
    const uint64_t psi_inv = hinv(psi);
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
    const uint64_t ninv = hinv((uint64_t)N);
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
    const uint64_t psi_inv = hinv(psi), ninv = hinv((uint64_t)N);
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
  if (backend == kNttFourStep && L < 2) return fail("four-step needs N >= 4");
  if (backend != kNttFourStep && backend != kNttNaive) return fail("unknown backend");
  if (backend == kNttNaive && L < 1) return fail("naive needs N >= 2");
  return true;
}

NttPlan& NttEngine::plan(int N) {
  std::lock_guard<std::mutex> lk(impl_->mu);
  auto& p = impl_->plans[N];
  if (!p) p.reset(new NttPlan(N, 0));
  return *p;
}

void NttEngine::tune(int N, int n1_log, int radix, int cols_max) {
  const int L = ilog2_exact((uint64_t)N);
  require(L >= 2 && L <= 24, "ntt tune: N must be 2^2..2^24");
  require(n1_log == 0 || (n1_log >= 1 && n1_log < L), "ntt tune: bad split");
  require(radix >= 2 && radix <= 5, "ntt tune: radix must be 2..5");
  require(cols_max >= 1 && cols_max <= 64 && (cols_max & (cols_max - 1)) == 0,
          "ntt tune: cols must be a power of two <= 64");
  std::lock_guard<std::mutex> lk(impl_->mu);
  auto& p = impl_->plans[N];
  const int want = n1_log > 0 ? n1_log : (L + 1) / 2;
  if (!p || p->n1_log != want) p.reset(new NttPlan(N, want));
  p->radix = radix;
  p->cols_max = cols_max;
}

uint64_t* NttEngine::scratch(size_t words, int which) {
  impl_->scratch[which].ensure(words * sizeof(uint64_t), "hipMalloc ntt scratch");
  return impl_->scratch[which].as<uint64_t>();
}

namespace {

size_t g_lds_budget = 0;  // bytes of LDS one workgroup may use for a tile

size_t lds_budget() {
  if (g_lds_budget == 0) {
    int dev = 0;
    hipDeviceProp_t prop;
    size_t b = 65536;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      b = prop.sharedMemPerBlock;
    g_lds_budget = std::min<size_t>(b, 131072);
  }
  return g_lds_budget;
}

int cols_per_wg(int n_log, int ncols, int cols_max) {
  int C = cols_max;
  while (C > 1 && (((size_t)C << n_log) * sizeof(uint64_t)) > lds_budget()) C >>= 1;
  while (C > ncols) C >>= 1;
  return C;
}

template <bool INV, bool TW, bool PROD, int R>
void launch_pass_r(hipStream_t st, const PassArgs& a, int batch, size_t lds) {
  static bool attr_set = false;
  if (!attr_set && lds > 65536) {
    hip_check(hipFuncSetAttribute((const void*)ntt_pass_kernel<INV, TW, PROD, R>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_budget()),
              "ntt: raise LDS limit");
    attr_set = true;
  }
  hipLaunchKernelGGL((ntt_pass_kernel<INV, TW, PROD, R>), dim3(a.ncols / a.C, batch),
                     dim3(kPassThreads), lds, st, a);
}

template <bool INV, bool TW, bool PROD>
void launch_pass(hipStream_t st, PassArgs a, int batch, Profiler* prof, int radix) {
  ProfZone pz(prof, st, INV ? (TW ? "ntt.inv_pass1" : "ntt.inv_pass2")
                                  : (TW ? "ntt.fwd_pass1" : "ntt.fwd_pass2"));
  const size_t lds = ((size_t)a.C << a.n_log) * sizeof(uint64_t);
  require(a.ncols % a.C == 0, "ntt: tile does not divide the columns");
  require(lds <= lds_budget(), "ntt: tile exceeds LDS");
  switch (radix) {
    case 2: launch_pass_r<INV, TW, PROD, 2>(st, a, batch, lds); break;
    case 3: launch_pass_r<INV, TW, PROD, 3>(st, a, batch, lds); break;
    case 5: launch_pass_r<INV, TW, PROD, 5>(st, a, batch, lds); break;
    default: launch_pass_r<INV, TW, PROD, 4>(st, a, batch, lds); break;
  }
  hip_check(hipGetLastError(), "launch ntt_pass_kernel");
}

void four_step_forward(hipStream_t st, NttPlan& pl, const uint64_t* in, uint64_t* tmp,
                       uint64_t* out, int batch, Profiler* prof) {
  const int n1 = 1 << pl.n1_log, n2 = 1 << pl.n2_log;
  PassArgs a{};
  // pass 1: columns i2, length-n1 negacyclic over i1, twiddle psi^(i2(2k1+1))
  a.in = in; a.out = tmp;
  a.tab = pl.n1_neg_fwd.as<uint64_t>();
  a.tw_lo = pl.tw_fwd_lo.as<uint64_t>(); a.tw_hi = pl.tw_fwd_hi.as<uint64_t>();
  a.tw_lo_bits = pl.tw_lo_bits; a.tw_swap = 0;
  a.n_log = pl.n1_log; a.ncols = n2; a.C = cols_per_wg(pl.n1_log, n2, pl.cols_max);
  a.in_ps = n2; a.in_cs = 1; a.out_ps = n2; a.out_cs = 1;
  a.in_brv = 0; a.out_brv = 1; a.N_log = pl.L;
  launch_pass<false, true, false>(st, a, batch, prof, pl.radix);
  // pass 2: rows k1, length-n2 cyclic over i2, written transposed (natural k)
  PassArgs b{};
  b.in = tmp; b.out = out;
  b.tab = pl.n2_cyc_fwd.as<uint64_t>();
  b.n_log = pl.n2_log; b.ncols = n1; b.C = cols_per_wg(pl.n2_log, n1, pl.cols_max);
  b.in_ps = 1; b.in_cs = n2; b.out_ps = n1; b.out_cs = 1;
  b.in_brv = 0; b.out_brv = 1; b.N_log = pl.L;
  launch_pass<false, false, false>(st, b, batch, prof, pl.radix);
}

void four_step_inverse(hipStream_t st, NttPlan& pl, const uint64_t* in, const uint64_t* in2,
                       uint64_t* tmp, uint64_t* out, int batch, Profiler* prof) {
  const int n1 = 1 << pl.n1_log, n2 = 1 << pl.n2_log;
  PassArgs a{};
  // pass I1: columns k1 of the [k2][k1] input, inverse cyclic over k2,
  // twiddle psi^-(i2(2k1+1)) N^-1, rows k1 out
  a.in = in; a.in2 = in2; a.out = tmp;
  a.tab = pl.n2_cyc_inv.as<uint64_t>();
  a.tw_lo = pl.tw_inv_lo.as<uint64_t>(); a.tw_hi = pl.tw_inv_hi.as<uint64_t>();
  a.tw_lo_bits = pl.tw_lo_bits; a.tw_swap = 1;
  a.n_log = pl.n2_log; a.ncols = n1; a.C = cols_per_wg(pl.n2_log, n1, pl.cols_max);
  a.in_ps = n1; a.in_cs = 1; a.out_ps = 1; a.out_cs = n2;
  a.in_brv = 1; a.out_brv = 0; a.N_log = pl.L;
  if (in2) launch_pass<true, true, true>(st, a, batch, prof, pl.radix);
  else launch_pass<true, true, false>(st, a, batch, prof, pl.radix);
  // pass I2: columns i2, inverse negacyclic over k1
  PassArgs b{};
  b.in = tmp; b.out = out;
  b.tab = pl.n1_neg_inv.as<uint64_t>();
  b.n_log = pl.n1_log; b.ncols = n2; b.C = cols_per_wg(pl.n1_log, n2, pl.cols_max);
  b.in_ps = n2; b.in_cs = 1; b.out_ps = n2; b.out_cs = 1;
  b.in_brv = 1; b.out_brv = 0; b.N_log = pl.L;
  launch_pass<true, false, false>(st, b, batch, prof, pl.radix);
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
