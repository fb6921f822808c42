// C ABI entry points (include/pcgsq.h) over the context, DPF, output-layer,
// NTT and AES engines.  Every function runs inside guarded(): exceptions
// become status codes and a per-thread last-error string.
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "context.hpp"
#include "pcgsq.h"
#include "prg.hpp"

// The no-allocation entry point's scratch array (see scratch_address below).
static __device__ __align__(256) uint64_t scratch[PCGSQ_NOALLOC_SCRATCH_BYTES / sizeof(uint64_t)];

namespace pcgsq {

namespace {
thread_local std::string t_last_error;

__global__ void aes_ecb_kernel(const uint8_t* __restrict__ keys, const uint8_t* __restrict__ in,
                               uint8_t* __restrict__ out, int count, const uint32_t* t0_g) {
  __shared__ uint32_t t0[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) t0[i] = t0_g[i];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count) return;
  // key schedule per block from the 16-byte key (on device, FIPS-197 5.2)
  uint32_t rk[44];
  for (int i = 0; i < 4; ++i) {
    const uint8_t* p = keys + 16 * (size_t)b + 4 * i;
    rk[i] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
  }
  uint32_t rcon = 1;
  for (int i = 4; i < 44; ++i) {
    uint32_t tmp = rk[i - 1];
    if (i % 4 == 0) {
      // RotWord then SubWord (bytes little endian in the word)
      tmp = (tmp >> 8) | (tmp << 24);
      const uint32_t s0 = (t0[tmp & 0xff] >> 8) & 0xff, s1 = (t0[(tmp >> 8) & 0xff] >> 8) & 0xff,
                     s2 = (t0[(tmp >> 16) & 0xff] >> 8) & 0xff, s3 = (t0[tmp >> 24] >> 8) & 0xff;
      tmp = (s0 | (s1 << 8) | (s2 << 16) | (s3 << 24)) ^ rcon;
      rcon = (rcon << 1) ^ ((rcon & 0x80) ? 0x11b : 0);
    }
    rk[i] = rk[i - 4] ^ tmp;
  }
  const uint8_t* p = in + 16 * (size_t)b;
  uint32_t s[4];
  for (int i = 0; i < 4; ++i)
    s[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
           ((uint32_t)p[4 * i + 3] << 24);
  aes128_encrypt_words(AesTable{t0}, rk, s[0], s[1], s[2], s[3]);
  uint8_t* o = out + 16 * (size_t)b;
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) o[4 * i + k] = (uint8_t)(s[i] >> (8 * k));
}

// Register-only ChaCha8 block loop: the compute ceiling of the tree PRG on
// this device.  Like a tree node keying its child, all four key words of
// each block come from the previous block's output and counter and nonce are
// zero, so the compiler can fold exactly what it folds in the tree kernels
// (the zero key half, the constants) and nothing more: no key word is
// loop-invariant, so no first-round quarter-round can be hoisted.
__global__ void chacha8_probe_kernel(uint32_t* out, int iters) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t k0 = t, k1 = t * 0x9e3779b9u, k2 = t * 3u, k3 = ~t;
  for (int i = 0; i < iters; ++i) {
    uint32_t o[4];
    chacha8<4>(k0, k1, k2, k3, 0, 0, o);
    k0 = o[0];
    k1 = o[1];
    k2 = o[2];
    k3 = o[3];
  }
  out[t] = k0 ^ k1 ^ k2 ^ k3;
}

// Device F_p arithmetic applied elementwise, for checking it against exact
// host arithmetic: op 0 mulhi64_dev(a, b), 1 shoup_lazy_dev(a, b, c),
// 2 shoup_lazy(a, b, c), 3 fmul(a, b).
__global__ void field_op_kernel(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                const uint64_t* __restrict__ c, uint64_t* __restrict__ o,
                                size_t n, int op) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (op) {
    case 0: o[i] = mulhi64_dev(a[i], b[i]); break;
    case 1: o[i] = shoup_lazy_dev(a[i], b[i], c[i]); break;
    case 2: o[i] = shoup_lazy(a[i], b[i], c[i]); break;
    default: o[i] = fmul(a[i], b[i]); break;
  }
}

// Plain 16-byte-per-thread copy: the known-size reference for HBM counters.
// HBM probes.  Each workgroup walks chunks of U * blockDim 16-byte words
// (grid-stride over chunks); a lane first issues its U loads, then its U
// stores, so U loads per lane are in flight at once.  NT: non-temporal loads
// and stores (no reuse expected, do not keep the lines in cache).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ void copy_probe_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                  size_t n) {
  const size_t chunk = (size_t)U * blockDim.x;
  for (size_t base = (size_t)blockIdx.x * chunk; base < n; base += (size_t)gridDim.x * chunk) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * blockDim.x + threadIdx.x;
      if (i < n) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * blockDim.x + threadIdx.x;
      if (i < n) {
        if (NT)
          __builtin_nontemporal_store(v[u], dst + i);
        else
          dst[i] = v[u];
      }
    }
  }
}

// Read-only: XOR of everything read, one word per lane written at the end
// (out has gridDim * blockDim words) so that the loads stay live.
template <int U>
__global__ void read_probe_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ out,
                                  size_t n) {
  const size_t chunk = (size_t)U * blockDim.x;
  u32x4 acc = {0u, 0u, 0u, 0u};
  for (size_t base = (size_t)blockIdx.x * chunk; base < n; base += (size_t)gridDim.x * chunk) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * blockDim.x + threadIdx.x;
      v[u] = i < n ? __builtin_nontemporal_load(src + i) : u32x4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc ^= v[u];
  }
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Write-only: a pattern derived from the index.
template <int U>
__global__ void write_probe_kernel(u32x4* __restrict__ dst, size_t n) {
  const size_t chunk = (size_t)U * blockDim.x;
  for (size_t base = (size_t)blockIdx.x * chunk; base < n; base += (size_t)gridDim.x * chunk) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * blockDim.x + threadIdx.x;
      const unsigned x = (unsigned)i;
      if (i < n) __builtin_nontemporal_store(u32x4{x, x ^ 1u, x ^ 2u, x ^ 3u}, dst + i);
    }
  }
}

// Best of `reps` timed launches (the first one untimed) of `launch` on st, ms.
template <class F>
float best_ms(hipStream_t st, int reps, F launch) {
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  float best = 1e30f;
  for (int rep = 0; rep <= reps; ++rep) {
    hip_check(hipEventRecord(e0, st), "event");
    launch();
    hip_check(hipGetLastError(), "launch probe");
    hip_check(hipEventRecord(e1, st), "event");
    hip_check(hipEventSynchronize(e1), "event");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
    if (rep > 0 && ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// Copy rate (bytes read + written per second) of the best probe shape, and
// that shape: unroll U in {1, 2, 4, 8}, plain or non-temporal, and 2..32
// workgroups of 256 lanes per CU.
struct CopyBest {
  double bps = 0;
  int U = 0, nt = 0, wg_per_cu = 0;
};

template <int U, bool NT>
void copy_sweep_one(hipStream_t st, int cus, const u32x4* a, u32x4* b, size_t n, CopyBest& best) {
  for (int wpc : {2, 4, 8, 16, 32}) {
    const dim3 grid((unsigned)(cus * wpc));
    const float ms = best_ms(st, 3, [&] {
      hipLaunchKernelGGL((copy_probe_kernel<U, NT>), grid, dim3(256), 0, st, a, b, n);
    });
    const double bps = 2.0 * (double)n * 16 / (ms * 1e-3);
    if (bps > best.bps) best = CopyBest{bps, U, NT ? 1 : 0, wpc};
  }
}

void upload(DevBuf& buf, const void* src, size_t bytes, hipStream_t st, const char* what) {
  buf.ensure(bytes, what);
  if (bytes) hip_check(hipMemcpyAsync(buf.ptr, src, bytes, hipMemcpyHostToDevice, st), what);
}

// Device copies of the AES T-table and round keys when a host schedule is given.
void stage_aes(pcgsq_ctx* ctx, const uint8_t* round_keys, int nk, DpfRun& r) {
  if (r.prg != kPrgAes) return;
  require(round_keys != nullptr, "AES PRG requires round_keys", kErrInvalid);
  uint32_t t0[256];
  aes_ttable_host(t0);
  upload(ctx->t0, t0, sizeof(t0), ctx->stream, "aes t-table");
  upload(ctx->rks, round_keys, (size_t)nk * 176, ctx->stream, "aes round keys");
  r.aes_t0 = ctx->t0.as<uint32_t>();
  r.aes_rks = ctx->rks.as<uint8_t>();
}

// Scratch of the no-allocation entry point (pcgsq_dpf_out_sums_noalloc): the
// leaf-sum partials, [B][nparts][2] words.  Each device holds its own copy
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory: the device address comes from hipGetSymbolAddress, once
// per device (scratch_address), and only that address is ever passed to a
// kernel.
constexpr size_t kNoAllocScratchBytes = (size_t)PCGSQ_NOALLOC_SCRATCH_BYTES;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[kMaxDevices] = {};

// This device's address of `scratch` (device current; looked up once).
uint64_t* scratch_address(int device) {
  require(device >= 0 && device < kMaxDevices, "scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch");
    require(p != nullptr, "hipGetSymbolAddress returned null for scratch", kErrDevice);
    g_scratch_ptr[device] = static_cast<uint64_t*>(p);
  }
  return g_scratch_ptr[device];
}

struct HostEx {
  pcgsq_exchange_fn fn;
  void* user;
};
int host_ex_tramp(void* u, int level, const Blk* mine, Blk* peer, size_t count) {
  HostEx* h = static_cast<HostEx*>(u);
  return h->fn(h->user, level, reinterpret_cast<const pcgsq_blk*>(mine),
               reinterpret_cast<pcgsq_blk*>(peer), count);
}

Blk* dpf_single(pcgsq_ctx* ctx, int B, int n, int w, const pcgsq_blk* roots,
                const pcgsq_blk* dltsft, const uint8_t* round_keys, pcgsq_exchange_fn exchange,
                void* user, int prg) {
  require(B > 0, "dpf: B must be positive");
  require(roots && dltsft && exchange, "dpf: null argument");
  require(w >= 1 && w <= 4 && n >= w && n % w == 0, "dpf: w must be in [1,4] and divide n");
  const int levels = n / w, nk = (1 << w) - 1;
  hipStream_t st = ctx->stream;
  upload(ctx->in_a, roots, sizeof(Blk) * B, st, "dpf roots");
  upload(ctx->in_b, dltsft, sizeof(Blk) * (size_t)B * levels * nk, st, "dpf dltsft");
  ctx->leaves.ensure(sizeof(Blk) * ((size_t)B << n), "dpf leaves");
  DpfRun r;
  r.ntrees = B;
  r.n = n;
  r.w = w;
  r.prg = prg;
  r.roots = ctx->in_a.as<Blk>();
  r.dltsft = ctx->in_b.as<Blk>();
  stage_aes(ctx, round_keys, nk, r);
  HostEx hx{exchange, user};
  r.host_ex = host_ex_tramp;
  r.host_user = &hx;
  r.mode = kLeafStore;
  r.leaves = ctx->leaves.as<Blk>();
  dpf_expand(st, r, ctx->dpf);
  return r.leaves;
}

int status_of_hip(hipError_t e) { return e == hipErrorOutOfMemory ? kErrNoMemory : kErrDevice; }

}  // namespace

void set_last_error(const std::string& msg) { t_last_error = msg; }

}  // namespace pcgsq

using namespace pcgsq;

extern "C" {

const char* pcgsq_last_error(void) { return t_last_error.c_str(); }

int pcgsq_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
  hipDeviceProp_t prop;
  for (int d = 0; d < n; ++d)
    if (hipGetDeviceProperties(&prop, d) == hipSuccess &&
        strncmp(prop.gcnArchName, "gfx950", 6) == 0)
      return 1;
  return 0;
}

int pcgsq_device_name(char* buf, size_t len) {
  return guarded([&] {
    require(buf && len, "device_name: null buffer");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    std::string s = std::string(prop.name) + " (" + prop.gcnArchName + ", HIP runtime " +
                    std::to_string(rt) + ")";
    strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
  });
}

int pcgsq_ctx_create(int device, pcgsq_ctx** out) {
  return guarded([&] {
    require(out != nullptr, "ctx_create: null out");
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    require(device >= 0 && device < n, "ctx_create: no such device", kErrUnsupported);
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    require(strncmp(prop.gcnArchName, "gfx950", 6) == 0,
            std::string("ctx_create: device is ") + prop.gcnArchName + ", built for gfx950",
            kErrUnsupported);
    std::unique_ptr<pcgsq_ctx> c(new pcgsq_ctx);
    c->device = device;
    c->ntt.prof = &c->prof;
    c->activate();
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    c->noalloc_scratch = scratch_address(device);
    *out = c.release();
  });
}

int pcgsq_ctx_destroy(pcgsq_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    if (ctx->owns_stream) {
      CtxGuard g(ctx);
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
  });
}

void* pcgsq_ctx_stream(pcgsq_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int pcgsq_ctx_synchronize(pcgsq_ctx* ctx) {
  return guarded([&] {
    require(ctx, "null ctx");
    CtxGuard g(ctx);
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  });
}

int pcgsq_ctx_profile(pcgsq_ctx* ctx, int enable) {
  return guarded([&] {
    require(ctx, "null ctx");
    CtxGuard g(ctx);
    ctx->prof.reset();
    ctx->prof.enabled = enable != 0;
  });
}

int pcgsq_ctx_profile_read(pcgsq_ctx* ctx, char* buf, size_t len) {
  return guarded([&] {
    require(ctx && buf && len, "profile_read: null argument");
    CtxGuard g(ctx);
    ctx->prof.resolve();
    std::string s = "{";
    bool first = true;
    for (const auto& kv : ctx->prof.totals) {
      char tmp[160];
      snprintf(tmp, sizeof tmp, "%s\"%s\": [%ld, %.6f]", first ? "" : ", ", kv.first.c_str(),
               kv.second.count, kv.second.ms);
      s += tmp;
      first = false;
    }
    s += "}";
    require(s.size() < len, "profile_read: buffer too small");
    memcpy(buf, s.c_str(), s.size() + 1);
  });
}

int pcgsq_probe_chacha8(pcgsq_ctx* ctx, int iters, double* blocks_per_s) {
  return guarded([&] {
    require(ctx && blocks_per_s && iters > 0, "probe: bad argument");
    CtxGuard g(ctx);
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, ctx->device), "props");
    const int grid = prop.multiProcessorCount * 8, tpb = 256;
    ctx->in_d.ensure((size_t)grid * tpb * 4, "probe");
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      hip_check(hipEventRecord(e0, ctx->stream), "event");
      hipLaunchKernelGGL(chacha8_probe_kernel, dim3(grid), dim3(tpb), 0, ctx->stream,
                         ctx->in_d.as<uint32_t>(), iters);
      hip_check(hipGetLastError(), "launch probe");
      hip_check(hipEventRecord(e1, ctx->stream), "event");
      hip_check(hipEventSynchronize(e1), "event");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
      if (rep > 0 && ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *blocks_per_s = (double)grid * tpb * iters / (best * 1e-3);
  });
}

int pcgsq_field_op(pcgsq_ctx* ctx, int op, size_t n, const uint64_t* a, const uint64_t* b,
                   const uint64_t* c, uint64_t* out) {
  return guarded([&] {
    require(ctx && a && b && out && n > 0, "field_op: bad argument");
    require(op >= 0 && op <= 3, "field_op: op must be 0..3");
    require(c || (op != 1 && op != 2), "field_op: Shoup needs the quotients");
    CtxGuard g(ctx);
    DevBuf da, db, dc, dout;
    da.ensure(n * 8, "field_op");
    db.ensure(n * 8, "field_op");
    dc.ensure(n * 8, "field_op");
    dout.ensure(n * 8, "field_op");
    hip_check(hipMemcpyAsync(da.ptr, a, n * 8, hipMemcpyHostToDevice, ctx->stream), "field_op");
    hip_check(hipMemcpyAsync(db.ptr, b, n * 8, hipMemcpyHostToDevice, ctx->stream), "field_op");
    hip_check(hipMemcpyAsync(dc.ptr, c ? c : b, n * 8, hipMemcpyHostToDevice, ctx->stream),
              "field_op");
    hipLaunchKernelGGL(field_op_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, da.as<uint64_t>(), db.as<uint64_t>(), dc.as<uint64_t>(),
                       dout.as<uint64_t>(), n, op);
    hip_check(hipGetLastError(), "launch field_op");
    hip_check(hipMemcpyAsync(out, dout.ptr, n * 8, hipMemcpyDeviceToHost, ctx->stream), "field_op");
    hip_check(hipStreamSynchronize(ctx->stream), "field_op");
  });
}

int pcgsq_probe_copy(pcgsq_ctx* ctx, size_t bytes, double* bytes_per_s) {
  return pcgsq_probe_hbm(ctx, bytes, bytes_per_s, nullptr, nullptr, nullptr);
}

int pcgsq_probe_hbm(pcgsq_ctx* ctx, size_t bytes, double* copy_bps, double* read_bps,
                    double* write_bps, int* copy_shape) {
  return guarded([&] {
    require(ctx && bytes >= 16, "probe_hbm: bad argument");
    CtxGuard g(ctx);
    const size_t n = bytes / 16;
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, ctx->device), "props");
    const int cus = prop.multiProcessorCount;
    DevBuf a, b;
    a.ensure(n * 16, "probe_hbm");
    b.ensure(std::max<size_t>(n, (size_t)cus * 32 * 256) * 16, "probe_hbm");
    hip_check(hipMemsetAsync(a.ptr, 1, n * 16, ctx->stream), "probe_hbm");
    const u32x4* src = a.as<u32x4>();
    u32x4* dst = b.as<u32x4>();
    hipStream_t st = ctx->stream;
    if (copy_bps) {
      CopyBest best;
      copy_sweep_one<1, false>(st, cus, src, dst, n, best);
      copy_sweep_one<4, false>(st, cus, src, dst, n, best);
      copy_sweep_one<8, false>(st, cus, src, dst, n, best);
      copy_sweep_one<1, true>(st, cus, src, dst, n, best);
      copy_sweep_one<2, true>(st, cus, src, dst, n, best);
      copy_sweep_one<4, true>(st, cus, src, dst, n, best);
      copy_sweep_one<8, true>(st, cus, src, dst, n, best);
      *copy_bps = best.bps;
      if (copy_shape) {
        copy_shape[0] = best.U;
        copy_shape[1] = best.nt;
        copy_shape[2] = best.wg_per_cu;
      }
    }
    if (read_bps) {
      double bestr = 0;
      for (int wpc : {4, 8, 16, 32}) {
        const float ms = best_ms(st, 3, [&] {
          hipLaunchKernelGGL(read_probe_kernel<8>, dim3((unsigned)(cus * wpc)), dim3(256), 0, st,
                             src, dst, n);
        });
        bestr = std::max(bestr, (double)n * 16 / (ms * 1e-3));
      }
      *read_bps = bestr;
    }
    if (write_bps) {
      double bestw = 0;
      for (int wpc : {4, 8, 16, 32}) {
        const float ms = best_ms(st, 3, [&] {
          hipLaunchKernelGGL(write_probe_kernel<8>, dim3((unsigned)(cus * wpc)), dim3(256), 0,
                             st, dst, n);
        });
        bestw = std::max(bestw, (double)n * 16 / (ms * 1e-3));
      }
      *write_bps = bestw;
    }
  });
}

int pcgsq_dev_alloc(pcgsq_ctx* ctx, size_t bytes, void** dev_out) {
  return guarded([&] {
    require(ctx && dev_out, "dev_alloc: null argument");
    CtxGuard g(ctx);
    *dev_out = nullptr;
    hip_check(hipMalloc(dev_out, bytes ? bytes : 1), "hipMalloc");
  });
}

int pcgsq_host_alloc(size_t bytes, void** host_out) {
  return guarded([&] {
    require(host_out, "host_alloc: null argument");
    *host_out = nullptr;
    hip_check(hipHostMalloc(host_out, bytes ? bytes : 1, hipHostMallocDefault),
              "hipHostMalloc");
  });
}

int pcgsq_host_free(void* host) {
  return guarded([&] {
    if (host) hip_check(hipHostFree(host), "hipHostFree");
  });
}

int pcgsq_dev_free(pcgsq_ctx* ctx, void* dev) {
  return guarded([&] {
    require(ctx, "dev_free: null ctx");
    CtxGuard g(ctx);
    hip_check(hipStreamSynchronize(ctx->stream), "dev_free");
    if (dev) hip_check(hipFree(dev), "hipFree");
  });
}

int pcgsq_memcpy_h2d(pcgsq_ctx* ctx, void* dev_dst, const void* host_src, size_t bytes) {
  return guarded([&] {
    require(ctx && dev_dst && host_src, "memcpy_h2d: null argument");
    CtxGuard g(ctx);
    hip_check(hipMemcpyAsync(dev_dst, host_src, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
    hip_check(hipStreamSynchronize(ctx->stream), "H2D");
  });
}

int pcgsq_memcpy_d2h(pcgsq_ctx* ctx, void* host_dst, const void* dev_src, size_t bytes) {
  return guarded([&] {
    require(ctx && host_dst && dev_src, "memcpy_d2h: null argument");
    CtxGuard g(ctx);
    hip_check(hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "D2H");
  });
}

int pcgsq_dpf_batch_full_eval(pcgsq_ctx* ctx, int B, int n, int w, const pcgsq_blk* roots,
                              const pcgsq_blk* dltsft, const uint8_t* round_keys, int party,
                              pcgsq_exchange_fn exchange, void* user, pcgsq_blk* out_leaves,
                              int prg) {
  return guarded([&] {
    require(ctx && out_leaves, "dpf: null argument");
    require(party == 0 || party == 1, "dpf: party must be 0 or 1");
    CtxGuard g(ctx);
    Blk* d = dpf_single(ctx, B, n, w, roots, dltsft, round_keys, exchange, user, prg);
    hip_check(hipMemcpyAsync(out_leaves, d, sizeof(Blk) * ((size_t)B << n), hipMemcpyDeviceToHost,
                             ctx->stream), "D2H leaves");
    hip_check(hipStreamSynchronize(ctx->stream), "D2H leaves");
  });
}

int pcgsq_dpf_batch_full_eval_device(pcgsq_ctx* ctx, int B, int n, int w, const pcgsq_blk* roots,
                                     const pcgsq_blk* dltsft, const uint8_t* round_keys, int party,
                                     pcgsq_exchange_fn exchange, void* user,
                                     const pcgsq_blk** d_leaves, int prg) {
  return guarded([&] {
    require(ctx && d_leaves, "dpf: null argument");
    require(party == 0 || party == 1, "dpf: party must be 0 or 1");
    CtxGuard g(ctx);
    Blk* d = dpf_single(ctx, B, n, w, roots, dltsft, round_keys, exchange, user, prg);
    hip_check(hipStreamSynchronize(ctx->stream), "dpf");
    *d_leaves = reinterpret_cast<const pcgsq_blk*>(d);
  });
}

int pcgsq_dpf_batch_full_eval_both(pcgsq_ctx* ctx, int B, int n, int w, const pcgsq_blk* roots0,
                                   const pcgsq_blk* roots1, const pcgsq_blk* dltsft0,
                                   const pcgsq_blk* dltsft1, const uint8_t* round_keys,
                                   pcgsq_blk* leaves0, pcgsq_blk* leaves1, pcgsq_blk* cw_out,
                                   int prg) {
  return guarded([&] {
    require(ctx && roots0 && roots1 && dltsft0 && dltsft1 && leaves0 && leaves1,
            "dpf_both: null argument");
    require(B > 0, "dpf_both: B must be positive");
    require(w >= 1 && w <= 4 && n >= w && n % w == 0, "dpf_both: w must be in [1,4] and divide n");
    CtxGuard g(ctx);
    const int levels = n / w, nk = (1 << w) - 1;
    hipStream_t st = ctx->stream;
    const size_t nd = (size_t)B * levels * nk;
    ctx->in_a.ensure(sizeof(Blk) * 2 * B, "dpf roots");
    ctx->in_b.ensure(sizeof(Blk) * 2 * nd, "dpf dltsft");
    hip_check(hipMemcpyAsync(ctx->in_a.as<Blk>(), roots0, sizeof(Blk) * B, hipMemcpyHostToDevice, st), "H2D");
    hip_check(hipMemcpyAsync(ctx->in_a.as<Blk>() + B, roots1, sizeof(Blk) * B, hipMemcpyHostToDevice, st), "H2D");
    hip_check(hipMemcpyAsync(ctx->in_b.as<Blk>(), dltsft0, sizeof(Blk) * nd, hipMemcpyHostToDevice, st), "H2D");
    hip_check(hipMemcpyAsync(ctx->in_b.as<Blk>() + nd, dltsft1, sizeof(Blk) * nd, hipMemcpyHostToDevice, st), "H2D");
    const size_t nleaves = (size_t)B << n;
    ctx->leaves.ensure(sizeof(Blk) * 2 * nleaves, "dpf leaves");
    ctx->cwbuf.ensure(sizeof(Blk) * 2 * (size_t)levels * B * nk, "dpf cw");
    DpfRun r;
    r.ntrees = 2 * B;
    r.n = n;
    r.w = w;
    r.prg = prg;
    r.roots = ctx->in_a.as<Blk>();
    r.dltsft = ctx->in_b.as<Blk>();
    stage_aes(ctx, round_keys, nk, r);
    r.mode = kLeafStore;
    r.leaves = ctx->leaves.as<Blk>();
    r.cw_out = ctx->cwbuf.as<Blk>();
    dpf_expand(st, r, ctx->dpf);
    hip_check(hipMemcpyAsync(leaves0, r.leaves, sizeof(Blk) * nleaves, hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipMemcpyAsync(leaves1, r.leaves + nleaves, sizeof(Blk) * nleaves, hipMemcpyDeviceToHost, st), "D2H");
    std::vector<Blk> shares;
    if (cw_out) {
      shares.resize(2 * (size_t)levels * B * nk);
      hip_check(hipMemcpyAsync(shares.data(), r.cw_out, sizeof(Blk) * shares.size(),
                               hipMemcpyDeviceToHost, st), "D2H cw");
    }
    hip_check(hipStreamSynchronize(st), "dpf_both");
    if (cw_out) {
      // shares are [levels][2B][nk]; CW = share(party 0) ^ share(party 1)
      Blk* o = reinterpret_cast<Blk*>(cw_out);
      for (int i = 0; i < levels; ++i)
        for (int b = 0; b < B; ++b)
          for (int k = 0; k < nk; ++k) {
            const Blk a0 = shares[((size_t)i * 2 * B + b) * nk + k];
            const Blk a1 = shares[((size_t)i * 2 * B + B + b) * nk + k];
            o[((size_t)i * B + b) * nk + k] = Blk{a0.lo ^ a1.lo, a0.hi ^ a1.hi};
          }
    }
  });
}

int pcgsq_dpf_out_sums(pcgsq_ctx* ctx, const pcgsq_blk* d_leaves, int B, size_t D, uint64_t prime,
                       uint64_t* sumC, uint64_t* sumT) {
  return guarded([&] {
    require(ctx && d_leaves && sumC && sumT, "out_sums: null argument");
    require(prime == kPrime, "out_sums: only P = 2^62 - 3*2^25 + 1 is supported", kErrUnsupported);
    require(B > 0, "out_sums: B must be positive");
    const int D_log = ilog2_exact(D);
    require(D_log >= 1 && D_log <= 30, "out_sums: D must be a power of two");
    CtxGuard g(ctx);
    const int nparts = leaf_sums_nparts(D_log);
    ctx->part.ensure(sizeof(uint64_t) * 2 * (size_t)B * nparts, "out partials");
    ctx->sums.ensure(sizeof(uint64_t) * 2 * (size_t)B, "out sums");
    uint64_t* s = ctx->sums.as<uint64_t>();
    launch_leaf_sums(ctx->stream, reinterpret_cast<const Blk*>(d_leaves), B, D_log,
                     ctx->part.as<uint64_t>(), nparts, s, s + B);
    hip_check(hipMemcpyAsync(sumC, s, 8 * (size_t)B, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipMemcpyAsync(sumT, s + B, 8 * (size_t)B, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "out_sums");
  });
}

int pcgsq_dpf_out_scatter_g(pcgsq_ctx* ctx, const pcgsq_blk* d_leaves, int B, size_t D, int t,
                            int party, uint64_t prime, const uint64_t* CW, uint64_t* g_out, int N) {
  return guarded([&] {
    require(ctx && d_leaves && CW && g_out, "out_scatter_g: null argument");
    require(prime == kPrime, "out_scatter_g: only P = 2^62 - 3*2^25 + 1 is supported",
            kErrUnsupported);
    require(party == 0 || party == 1, "out_scatter_g: party must be 0 or 1");
    require(B > 0 && t > 0 && N > 0, "out_scatter_g: bad shape");
    const int D_log = ilog2_exact(D);
    require(D_log >= 1 && D_log <= 30, "out_scatter_g: D must be a power of two");
    require((size_t)N == (size_t)t * (D / 2), "out_scatter_g: N must equal t * D/2");
    CtxGuard g(ctx);
    upload(ctx->cwbuf, CW, 8 * (size_t)B, ctx->stream, "H2D cw");
    ctx->gbuf.ensure(8 * (size_t)N, "out g");
    GatherArgs ga;
    ga.leaves = reinterpret_cast<const Blk*>(d_leaves);
    ga.set_src_stride = 0;
    ga.CW = ctx->cwbuf.as<uint64_t>();
    ga.g = ctx->gbuf.as<uint64_t>();
    ga.nsets = 1;
    ga.sets_per_party = 1;
    ga.first_party = party;
    ga.B = B;
    ga.t = t;
    ga.N = N;
    ga.D_log = D_log;
    launch_gather_g(ctx->stream, ga);
    hip_check(hipMemcpyAsync(g_out, ga.g, 8 * (size_t)N, hipMemcpyDeviceToHost, ctx->stream), "D2H g");
    hip_check(hipStreamSynchronize(ctx->stream), "out_scatter_g");
  });
}

int pcgsq_dpf_out_sums_noalloc(pcgsq_ctx* ctx, const pcgsq_blk* d_leaves, int B, size_t D,
                               uint64_t prime, uint64_t* d_sumC, uint64_t* d_sumT) {
  return guarded([&] {
    require(ctx && d_leaves && d_sumC && d_sumT, "out_sums_noalloc: null argument");
    require(prime == kPrime, "out_sums_noalloc: only P = 2^62 - 3*2^25 + 1 is supported",
            kErrUnsupported);
    require(B > 0, "out_sums_noalloc: B must be positive");
    const int D_log = ilog2_exact(D);
    require(D_log >= 1 && D_log <= 30, "out_sums_noalloc: D must be a power of two");
    const int nparts = leaf_sums_nparts(D_log);
    const size_t need = sizeof(uint64_t) * 2 * (size_t)B * (size_t)nparts;
    require(need <= kNoAllocScratchBytes,
            "out_sums_noalloc: B * partials per tree needs " + std::to_string(need) +
                " scratch bytes, above the fixed " + std::to_string(kNoAllocScratchBytes),
            kErrInvalid);
    CtxGuard g(ctx);
    require(ctx->noalloc_scratch != nullptr, "out_sums_noalloc: context has no scratch",
            kErrInternal);
    // every partial of the B trees is written by exactly one workgroup of the
    // sums kernel before the fold reads it: nothing must start at zero
    launch_leaf_sums(ctx->stream, reinterpret_cast<const Blk*>(d_leaves), B, D_log,
                     ctx->noalloc_scratch, nparts, d_sumC, d_sumT);
  });
}

int pcgsq_poly_mul_supported(uint64_t prime, int N, int backend, char* reason, size_t reason_len) {
  std::string why;
  bool ok = prime == kPrime;
  if (!ok) why = "only P = 2^62 - 3*2^25 + 1 is supported";
  if (ok) ok = NttEngine::supported(N, backend, &why);
  if (!ok && reason && reason_len) {
    strncpy(reason, why.c_str(), reason_len - 1);
    reason[reason_len - 1] = 0;
  }
  return ok ? 1 : 0;
}

int pcgsq_poly_mul_u64(pcgsq_ctx* ctx, uint64_t* out, const uint64_t* a, const uint64_t* b,
                       int batch, int N, uint64_t prime, int backend, double* device_ms) {
  return guarded([&] {
    require(ctx && out && a && b, "poly_mul: null argument");
    require(prime == kPrime, "poly_mul: only P = 2^62 - 3*2^25 + 1 is supported", kErrUnsupported);
    std::string why;
    require(NttEngine::supported(N, backend, &why), "poly_mul: " + why, kErrUnsupported);
    require(batch > 0, "poly_mul: batch must be positive");
    CtxGuard g(ctx);
    const size_t bytes = 8 * (size_t)batch * N;
    hipStream_t st = ctx->stream;
    upload(ctx->in_a, a, bytes, st, "H2D a");
    upload(ctx->in_b, b, bytes, st, "H2D b");
    ctx->in_c.ensure(bytes, "poly_mul out");
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, st), "event");
    ctx->ntt.poly_mul(st, ctx->in_c.as<uint64_t>(), ctx->in_a.as<uint64_t>(),
                      ctx->in_b.as<uint64_t>(), batch, N, backend);
    hip_check(hipEventRecord(e1, st), "event");
    hip_check(hipMemcpyAsync(out, ctx->in_c.ptr, bytes, hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "poly_mul");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (device_ms) *device_ms = ms;
  });
}

int pcgsq_poly_mul_u64_dev(pcgsq_ctx* ctx, uint64_t* dev_out, const uint64_t* dev_a,
                           const uint64_t* dev_b, int batch, int N, int backend) {
  return guarded([&] {
    require(ctx && dev_out && dev_a && dev_b, "poly_mul_dev: null argument");
    CtxGuard g(ctx);
    ctx->ntt.poly_mul(ctx->stream, dev_out, dev_a, dev_b, batch, N, backend);
  });
}

static int ntt_host(pcgsq_ctx* ctx, uint64_t* a, int batch, int N, int backend, bool inv) {
  return guarded([&] {
    require(ctx && a, "ntt: null argument");
    require(batch > 0, "ntt: batch must be positive");
    std::string why;
    require(NttEngine::supported(N, backend, &why), "ntt: " + why, kErrUnsupported);
    CtxGuard g(ctx);
    const size_t bytes = 8 * (size_t)batch * N;
    upload(ctx->in_a, a, bytes, ctx->stream, "H2D ntt");
    if (inv) ctx->ntt.inverse(ctx->stream, ctx->in_a.as<uint64_t>(), batch, N, backend);
    else ctx->ntt.forward(ctx->stream, ctx->in_a.as<uint64_t>(), batch, N, backend);
    hip_check(hipMemcpyAsync(a, ctx->in_a.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream), "D2H");
    hip_check(hipStreamSynchronize(ctx->stream), "ntt");
  });
}

int pcgsq_ntt_forward(pcgsq_ctx* ctx, uint64_t* a, int batch, int N, int backend) {
  return ntt_host(ctx, a, batch, N, backend, false);
}
int pcgsq_ntt_inverse(pcgsq_ctx* ctx, uint64_t* a, int batch, int N, int backend) {
  return ntt_host(ctx, a, batch, N, backend, true);
}
int pcgsq_ntt_forward_dev(pcgsq_ctx* ctx, uint64_t* dev_a, int batch, int N, int backend) {
  return guarded([&] {
    require(ctx && dev_a, "ntt: null argument");
    CtxGuard g(ctx);
    ctx->ntt.forward(ctx->stream, dev_a, batch, N, backend);
  });
}
int pcgsq_ntt_inverse_dev(pcgsq_ctx* ctx, uint64_t* dev_a, int batch, int N, int backend) {
  return guarded([&] {
    require(ctx && dev_a, "ntt: null argument");
    CtxGuard g(ctx);
    ctx->ntt.inverse(ctx->stream, dev_a, batch, N, backend);
  });
}

int pcgsq_ntt_tune(pcgsq_ctx* ctx, int N, int n1_log, int variant, int cols1, int cols2) {
  return guarded([&] {
    require(ctx, "ntt_tune: null ctx");
    CtxGuard g(ctx);
    ctx->ntt.tune(N, n1_log, variant, cols1, cols2);
  });
}

int pcgsq_aes128_expand_key(const uint8_t key[16], uint8_t round_keys[176]) {
  return guarded([&] {
    require(key && round_keys, "aes: null argument");
    aes128_expand_key_host(key, round_keys);
  });
}

int pcgsq_aes_tree_round_keys(const uint8_t master[16], int nk, uint8_t* rks) {
  return guarded([&] {
    require(master && rks && nk >= 1 && nk <= 15, "aes tree keys: bad argument");
    aes_tree_round_keys_host(master, nk, rks);
  });
}

int pcgsq_aes128_ecb_encrypt(pcgsq_ctx* ctx, const uint8_t* keys, const uint8_t* in, uint8_t* out,
                             int count) {
  return guarded([&] {
    require(ctx && keys && in && out && count > 0, "aes ecb: bad argument");
    CtxGuard g(ctx);
    hipStream_t st = ctx->stream;
    uint32_t t0[256];
    aes_ttable_host(t0);
    upload(ctx->t0, t0, sizeof(t0), st, "aes t-table");
    upload(ctx->in_a, keys, 16 * (size_t)count, st, "aes keys");
    upload(ctx->in_b, in, 16 * (size_t)count, st, "aes in");
    ctx->in_c.ensure(16 * (size_t)count, "aes out");
    hipLaunchKernelGGL(aes_ecb_kernel, dim3((count + 127) / 128), dim3(128), 0, st,
                       ctx->in_a.as<uint8_t>(), ctx->in_b.as<uint8_t>(), ctx->in_c.as<uint8_t>(),
                       count, ctx->t0.as<uint32_t>());
    hip_check(hipGetLastError(), "launch aes_ecb");
    hip_check(hipMemcpyAsync(out, ctx->in_c.ptr, 16 * (size_t)count, hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "aes ecb");
  });
}

}  // extern "C"
