// NTT kernels and plans (see ntt.hpp).
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "ntt.hpp"
#include "ntt_pass.hpp"
#include "ntt_plan.hpp"

namespace pcgsq {

namespace {

constexpr int kPassThreads = 512;
constexpr int kEwThreads = 256;


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
  const size_t base = d.boff ? (size_t)d.boff[blockIdx.y] : (size_t)blockIdx.y << d.N_log;
  nttpass::pass_tile<K, RA, RB, RC>(d, lds, base, bid);
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

unsigned grid_for(size_t n) { return (unsigned)((n + kEwThreads - 1) / kEwThreads); }

// Per-element four-step twiddles (ntt_plan.hpp), one thread per element:
// both tables' entry idx = a 2^n2_log + b holds psi^(+-e) with e = b (2a + 1)
// mod 2N, a product of one entry of each two-level table.
__global__ void build_twiddles_kernel(uint64_t* fwd, uint64_t* inv, const uint64_t* lo,
                                      const uint64_t* hi, const uint64_t* loi,
                                      const uint64_t* hii, int n1_log, int n2_log, int lo_bits) {
  const uint32_t idx = blockIdx.x * kEwThreads + threadIdx.x;
  const int L = n1_log + n2_log;
  if (idx >> L) return;
  const uint32_t a = idx >> n2_log, b = idx & ((1u << n2_log) - 1);
  const uint32_t e = (b * (2 * a + 1)) & ((2u << L) - 1);
  const uint32_t el = e & ((1u << lo_bits) - 1), eh = e >> lo_bits;
  fwd[idx] = shoup(lo[2 * el], hi[2 * eh], hi[2 * eh + 1]);
  inv[idx] = shoup(loi[2 * el], hii[2 * eh], hii[2 * eh + 1]);
}

}  // namespace

void ntt_build_twiddles(uint64_t* fwd, uint64_t* inv, const uint64_t* lo, const uint64_t* hi,
                        const uint64_t* loi, const uint64_t* hii, int n1_log, int n2_log,
                        int lo_bits) {
  hipStream_t st = nullptr;
  hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "ntt twiddles: stream");
  const size_t n = 1ULL << (n1_log + n2_log);
  hipLaunchKernelGGL(build_twiddles_kernel, dim3(grid_for(n)), dim3(kEwThreads), 0, st, fwd, inv,
                     lo, hi, loi, hii, n1_log, n2_log, lo_bits);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  hip_check(e, "ntt twiddles");
}

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

void NttEngine::trim() {
  for (DevBuf& b : impl_->scratch) b.release();
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
  int nlog, rmax, groups;
};

template <int K, int RA, int RB, int RC>
PassKernel pk() {
  using P = nttpass::Part<RA, RB, RC>;
  return PassKernel{(const void*)ntt_pass_kernel<K, RA, RB, RC>, P::NLOG, P::RMAX, P::G};
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
int tile_cols(int want, int nlog, int rmax, int groups, int ncols) {
  int C = want > 0 ? want : std::max(1, 256 >> (nlog - rmax));
  if (want <= 0) C = std::min(C, 16);
  while (C > 1 && (C << (nlog - rmax)) > kPassThreads) C >>= 1;
  while (C > 1 && ntt_tile_lds_bytes(nlog, C, groups) > lds_max() / (want > 0 ? 1 : 2)) C >>= 1;
  while (C > 1 && C > ncols) C >>= 1;
  return C;
}

template <int K>
void launch_pass(hipStream_t st, NttPassDesc d, int nlog, int variant, int want_cols, int ncols,
                 int batch, Profiler* prof, const char* zone) {
  ProfZone pz(prof, st, zone);
  const PassKernel k = pass_kernel<K>(nlog, variant);
  const int C = tile_cols(want_cols, nlog, k.rmax, k.groups, ncols);
  d.c_log = ilog2_exact((uint64_t)C);
  d.ncols = ncols;
  const uint32_t tiles = (uint32_t)((ncols + C - 1) / C);
  const size_t lds = ntt_tile_lds_bytes(nlog, C, k.groups);
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

void launch_spec(hipStream_t st, PassSpec ps, int batch, Profiler* prof, const char* zone) {
  // whole transforms back to back are the columns of one launch; with an
  // offset table each is a launch row of its own
  const bool whole = ps.kind == kFwdWhole || ps.kind == kInvWhole || ps.kind == kInvWholeProd;
  const int grid_y = whole && !ps.d.boff ? 1 : batch;
  switch (ps.kind) {
    case kFwdCols: return launch_pass<kFwdCols>(st, ps.d, ps.nlog, ps.variant, ps.cols, ps.ncols, grid_y, prof, zone);
    case kFwdRows: return launch_pass<kFwdRows>(st, ps.d, ps.nlog, ps.variant, ps.cols, ps.ncols, grid_y, prof, zone);
    case kInvRows: return launch_pass<kInvRows>(st, ps.d, ps.nlog, ps.variant, ps.cols, ps.ncols, grid_y, prof, zone);
    case kInvRowsProd: return launch_pass<kInvRowsProd>(st, ps.d, ps.nlog, ps.variant, ps.cols, ps.ncols, grid_y, prof, zone);
    case kInvCols: return launch_pass<kInvCols>(st, ps.d, ps.nlog, ps.variant, ps.cols, ps.ncols, grid_y, prof, zone);
    case kFwdWhole: return launch_pass<kFwdWhole>(st, ps.d, ps.nlog, ps.variant, ps.cols, ps.ncols, grid_y, prof, zone);
    case kInvWhole: return launch_pass<kInvWhole>(st, ps.d, ps.nlog, ps.variant, ps.cols, ps.ncols, grid_y, prof, zone);
    case kInvWholeProd: return launch_pass<kInvWholeProd>(st, ps.d, ps.nlog, ps.variant, ps.cols, ps.ncols, grid_y, prof, zone);
    default: throw Error(kErrInternal, "ntt: no such pass");
  }
}

void four_step_forward(hipStream_t st, NttPlan& pl, const uint64_t* in, uint64_t* tmp,
                       uint64_t* out, int batch, Profiler* prof) {
  if (pl.single()) {
    // whole transforms as columns: one HBM round trip
    launch_spec(st, plan_pass(pl, kPhFwd1, in, nullptr, out, batch), batch, prof, "ntt.fwd_single");
    return;
  }
  launch_spec(st, plan_pass(pl, kPhFwd1, in, nullptr, tmp, batch), batch, prof, "ntt.fwd_pass1");
  launch_spec(st, plan_pass(pl, kPhFwd2, tmp, nullptr, out, batch), batch, prof, "ntt.fwd_pass2");
}

void four_step_inverse(hipStream_t st, NttPlan& pl, const uint64_t* in, const uint64_t* in2,
                       uint64_t* tmp, uint64_t* out, int batch, Profiler* prof) {
  if (pl.single()) {
    launch_spec(st, plan_pass(pl, kPhInv1, in, in2, out, batch), batch, prof, "ntt.inv_single");
    return;
  }
  launch_spec(st, plan_pass(pl, kPhInv1, in, in2, tmp, batch), batch, prof, "ntt.inv_pass1");
  launch_spec(st, plan_pass(pl, kPhInv2, tmp, nullptr, out, batch), batch, prof, "ntt.inv_pass2");
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

void ntt_launch_pass(hipStream_t st, const PassSpec& ps, int batch, Profiler* prof,
                     const char* zone) {
  launch_spec(st, ps, batch, prof, zone);
}

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
