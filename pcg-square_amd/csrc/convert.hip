// Output-layer kernels (see convert.hpp).
#include "convert.hpp"
#include "internal.hpp"

namespace pcgsq {

namespace {

constexpr int kTpb = 256;

__global__ __launch_bounds__(kTpb) void fold_sums_kernel(const uint64_t* __restrict__ part,
                                                         int ntrees, int nparts,
                                                         uint64_t* __restrict__ sumC,
                                                         uint64_t* __restrict__ sumT) {
  const int T = blockIdx.x * kTpb + threadIdx.x;
  if (T >= ntrees) return;
  uint64_t c = 0, t = 0;
  const uint64_t* p = part + (size_t)T * nparts * 2;
  for (int i = 0; i < nparts; ++i) {
    c = fadd(c, p[2 * i]);
    t += p[2 * i + 1];
  }
  sumC[T] = c;
  sumT[T] = t;
}

// One thread per instance: both parties' A/B shares, the two Beaver opens and
// the reconstructed CW (pcg_ole_impl.h:222-302 with the network rounds
// replaced by reading the other party's values).
__global__ __launch_bounds__(kTpb) void beaver_cw_kernel(const uint64_t* __restrict__ part,
                                                         int ninst, int nparts,
                                                         const uint64_t* __restrict__ beta,
                                                         const uint64_t* __restrict__ r,
                                                         const uint64_t* __restrict__ s,
                                                         const uint64_t* __restrict__ rs,
                                                         size_t ps,
                                                         uint64_t* __restrict__ CW) {
  const int b = blockIdx.x * kTpb + threadIdx.x;
  if (b >= ninst) return;
  uint64_t d[2], e[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const uint64_t* pp = part + ((size_t)p * ninst + b) * nparts * 2;
    uint64_t c = 0, t = 0;
    for (int i = 0; i < nparts; ++i) {
      c = fadd(c, pp[2 * i]);
      t += pp[2 * i + 1];
    }
    uint64_t A = freduce64(t);
    if (p == 1) A = fneg(A);
    uint64_t Bv = c;
    if (p == 0) Bv = fneg(Bv);
    Bv = fadd(Bv, beta[p * ps + b]);
    d[p] = fsub(A, r[p * ps + b]);
    e[p] = fsub(Bv, s[p * ps + b]);
  }
  const uint64_t dv = fadd(d[0], d[1]), ev = fadd(e[0], e[1]);
  uint64_t cw = 0;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    uint64_t v = fadd(fadd(fmul(dv, s[p * ps + b]), fmul(ev, r[p * ps + b])), rs[p * ps + b]);
    if (p == 0) v = fadd(v, fmul(dv, ev));
    cw = fadd(cw, v);
  }
  CW[b] = cw;
}

// g[q] = sum over raw positions pos in {q, q+N} (sign - for q+N) of the
// leaves (inst, d) with (kk+ll)*bin + d == pos.  Consecutive threads take
// consecutive q, hence consecutive d: every leaf load is coalesced and every
// leaf is read exactly once.
template <bool FROM_LEAVES>
__global__ __launch_bounds__(kTpb) void gather_g_kernel(GatherArgs a) {
  const int bb = blockIdx.y;                    // block within the party
  const int party = a.first_party + blockIdx.z;
  const int set = blockIdx.z * a.sets_per_party + bb;
  const size_t q = (size_t)blockIdx.x * kTpb + threadIdx.x;
  if (q >= (size_t)a.N) return;
  const size_t D = (size_t)1 << a.D_log;
  const int bin_log = a.D_log - 1;
  const size_t bin = D >> 1;
  const uint64_t* ct = FROM_LEAVES ? nullptr : a.ct + a.set_src_stride * set;
  const Blk* lv = FROM_LEAVES ? a.leaves + a.set_src_stride * set : nullptr;
  const uint64_t* CW = a.CW + a.cw_stride * bb;
  const int t = a.t, B = a.B;
  uint64_t acc = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const size_t pos = q + (size_t)h * a.N;
    const int sb = (int)(pos >> bin_log);
    const size_t rr = pos & (bin - 1);
    uint64_t part = 0;
#pragma unroll
    for (int ds = 0; ds < 2; ++ds) {
      const int sd = sb - ds;             // diagonal kk + ll
      if (sd < 0 || sd > 2 * t - 2) continue;
      const size_t d = rr + (size_t)ds * bin;
      const int kk_lo = sd - t + 1 > 0 ? sd - t + 1 : 0;
      const int kk_hi = sd < t - 1 ? sd : t - 1;
      if constexpr (FROM_LEAVES) {
        for (int kk = kk_lo; kk <= kk_hi; ++kk) {
          const int inst = kk * t + (sd - kk);
          if (inst >= B) break;
          const Blk leaf = lv[(size_t)inst * D + d];
          const uint64_t C = freduce64(out_hash_lo64(leaf));
          part = fadd(part, (leaf.lo & 1ULL) ? fadd(C, CW[inst]) : C);
        }
      } else {
        // kRows rows' loads in flight per thread (the instance index grows
        // with kk, so rows past B end the range)
        constexpr int kRows = 4;
        for (int kk = kk_lo; kk <= kk_hi; kk += kRows) {
          uint64_t v[kRows], cw[kRows];
          bool ok[kRows];
#pragma unroll
          for (int j = 0; j < kRows; ++j) {
            const int k2 = kk + j;
            const int inst = k2 * t + (sd - k2);
            ok[j] = k2 <= kk_hi && inst < B;
            v[j] = ok[j] ? ct[(size_t)inst * D + d] : 0;
            cw[j] = ok[j] ? CW[inst] : 0;
          }
#pragma unroll
          for (int j = 0; j < kRows; ++j) {
            const uint64_t C = v[j] & ((1ULL << 63) - 1);
            part = fadd(part, (v[j] >> 63) ? fadd(C, cw[j]) : C);
          }
        }
      }
    }
    acc = h ? fsub(acc, part) : fadd(acc, part);
  }
  if (party == 1) acc = fneg(acc);
  a.g[(size_t)(party - a.first_party) * a.g_party_stride + (size_t)bb * a.N + q] = acc;
}

// hi * 2^64 + lo: sums of field elements without reduction (hi < 2^24).
struct Wide {
  uint64_t lo = 0;
  uint32_t hi = 0;
  __device__ __forceinline__ void add(uint64_t x) {
    lo += x;
    hi += lo < x;
  }
  // mod P, with 2^64 = 4 * (3 * 2^25 - 1) mod P
  __device__ __forceinline__ uint64_t mod() const {
    return freduce64(freduce64(lo) + (uint64_t)hi * (4 * ((3ULL << 25) - 1)));
  }
};

// Two output coefficients per thread; the landing slots are the same for the
// workgroup's 512 coefficients (one bin).  Thread i finds slot i in closed
// form and loads its signed CW; then, per 64-slot tau word, the workgroup
// tabulates for each nibble of the word the sum of the signed CWs of every
// subset of its 4 slots (16 tables of 16 entries in LDS), so a coefficient
// adds 16 table entries per word instead of testing 64 bits.  The
// coefficients' own loads are issued first and land meanwhile.
constexpr int kFixMaxWords = kGatherMaxSlots / 64;

__global__ __launch_bounds__(kTpb) void gather_fix_kernel(GatherFixArgs a) {
  static_assert(kGatherMaxSlots <= kTpb, "one thread per landing slot");
  __shared__ uint64_t sc[kGatherMaxSlots];   // signed CW per slot, 0 past the last
  __shared__ uint64_t tab[kFixMaxWords][16][16];
  const int bb = blockIdx.y, pz = blockIdx.z, party = a.first_party + pz;
  const uint32_t qt = blockIdx.x * (2 * kTpb);
  const uint32_t q0 = qt + 2 * threadIdx.x;
  const size_t set_off = ((size_t)pz * a.sets + bb) << a.N_log;
  const size_t word_stride = ((size_t)a.parties * a.sets) << a.N_log;
  Blk* gq = reinterpret_cast<Blk*>(a.g + (size_t)pz * a.g_party_stride + ((size_t)bb << a.N_log) +
                                   q0);
  const Blk g = *gq;
  Blk m[kFixMaxWords];
#pragma unroll
  for (int w = 0; w < kFixMaxWords; ++w)
    m[w] = w < a.mask_words
               ? *reinterpret_cast<const Blk*>(a.mask + (size_t)w * word_stride + set_off + q0)
               : Blk{0, 0};
  {
    int h = 0, inst = 0;
    uint64_t v = 0;
    if (threadIdx.x < kGatherMaxSlots &&
        landing_slot(qt, a.N_log, a.bin_log, a.t, a.B, (int)threadIdx.x, h, inst)) {
      const uint64_t cw = a.CW[(size_t)bb * a.B + inst];
      v = h ? kP - cw : cw;  // kP - 0 = kP adds nothing mod P
    }
    if (threadIdx.x < kGatherMaxSlots) sc[threadIdx.x] = v;
  }
  __syncthreads();
  {
    const int j = threadIdx.x >> 4, v = threadIdx.x & 15;
    for (int w = 0; w < a.mask_words; ++w) {
      uint64_t x = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if ((v >> b) & 1) x = fadd(x, sc[64 * w + 4 * j + b]);  // sc <= P: fadd stays below P
      tab[w][j][v] = x;
    }
  }
  __syncthreads();
  Wide s0, s1;
#pragma unroll
  for (int w = 0; w < kFixMaxWords; ++w) {
    if (w >= a.mask_words) break;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      s0.add(tab[w][j][(m[w].lo >> (4 * j)) & 15]);
      s1.add(tab[w][j][(m[w].hi >> (4 * j)) & 15]);
    }
  }
  uint64_t r0 = fadd(g.lo, s0.mod()), r1 = fadd(g.hi, s1.mod());
  if (party == 1) {
    r0 = fneg(r0);
    r1 = fneg(r1);
  }
  *gq = Blk{r0, r1};
}

__global__ __launch_bounds__(kTpb) void leaf_sums_kernel(const Blk* __restrict__ leaves, int B,
                                                         int D_log, int iters, int chunk_log,
                                                         int nparts,
                                                         uint64_t* __restrict__ part) {
  __shared__ uint64_t red[16];
  const size_t total = (size_t)B << D_log;
  const size_t base = (size_t)blockIdx.x * iters * kTpb;
  uint64_t v[2] = {0, 0};
  for (int it = 0; it < iters; ++it) {
    const size_t f = base + (size_t)it * kTpb + threadIdx.x;
    if (f >= total) break;
    const Blk leaf = leaves[f];
    v[0] = fadd(v[0], freduce64(out_hash_lo64(leaf)));
    v[1] += leaf.lo & 1ULL;
  }
  int first = 0;
  if (segment_reduce<2>(v, D_log,
                        [](int i, uint64_t x, uint64_t y) { return i == 0 ? fadd(x, y) : x + y; },
                        red, &first)) {
    const size_t f = base + (size_t)first;
    const size_t T = f >> D_log;
    const size_t pidx = (D_log >= 8) ? ((f & (((size_t)1 << D_log) - 1)) >> chunk_log) : 0;
    if (T < (size_t)B) {
      part[(T * nparts + pidx) * 2] = v[0];
      part[(T * nparts + pidx) * 2 + 1] = v[1];
    }
  }
}

int leaf_iters(int D_log) {
  const size_t D = (size_t)1 << D_log;
  if (D < (size_t)kTpb) return 1;
  return (int)std::min<size_t>(D / kTpb, 8);
}

}  // namespace

int leaf_sums_nparts(int D_log) {
  const size_t D = (size_t)1 << D_log;
  if (D < (size_t)kTpb) return 1;
  return (int)(D / ((size_t)leaf_iters(D_log) * kTpb));
}

void launch_fold_sums(hipStream_t st, const uint64_t* partials, int ntrees, int nparts,
                      uint64_t* sumC, uint64_t* sumT) {
  const int grid = (ntrees + kTpb - 1) / kTpb;
  hipLaunchKernelGGL(fold_sums_kernel, dim3(grid), dim3(kTpb), 0, st, partials, ntrees, nparts,
                     sumC, sumT);
  hip_check(hipGetLastError(), "launch fold_sums_kernel");
}

void launch_beaver_cw(hipStream_t st, const uint64_t* partials, int ninst, int nparts,
                      const uint64_t* beta, const uint64_t* r, const uint64_t* s,
                      const uint64_t* rs, size_t pstride, uint64_t* CW) {
  const int grid = (ninst + kTpb - 1) / kTpb;
  hipLaunchKernelGGL(beaver_cw_kernel, dim3(grid), dim3(kTpb), 0, st, partials, ninst, nparts,
                     beta, r, s, rs, pstride, CW);
  hip_check(hipGetLastError(), "launch beaver_cw_kernel");
}

void launch_gather_g(hipStream_t st, const GatherArgs& a) {
  require(a.N > 0 && a.t > 0 && a.B > 0 && a.nsets > 0 && a.sets_per_party > 0,
          "gather: bad shape");
  require(a.sets_per_party <= 65535 && a.nsets % a.sets_per_party == 0 &&
              a.nsets / a.sets_per_party <= 2,
          "gather: sets must be whole parties of at most 65535 blocks");
  ProfZone pz(a.prof, st, "out.gather");
  const int grid = (a.N + kTpb - 1) / kTpb;
  const dim3 g3(grid, a.sets_per_party, a.nsets / a.sets_per_party);
  if (a.leaves)
    hipLaunchKernelGGL(gather_g_kernel<true>, g3, dim3(kTpb), 0, st, a);
  else
    hipLaunchKernelGGL(gather_g_kernel<false>, g3, dim3(kTpb), 0, st, a);
  hip_check(hipGetLastError(), "launch gather_g_kernel");
}

void launch_gather_fix(hipStream_t st, const GatherFixArgs& a) {
  require(a.sets >= 1 && a.sets <= 65535 && a.B > 0 && a.t > 0 && a.mask_words >= 1 &&
              a.bin_log >= 9 && a.N_log >= a.bin_log && 4 * a.t <= 64 * a.mask_words &&
              a.mask_words <= kFixMaxWords && a.parties >= 1 &&
              a.first_party + a.parties <= 2,
          "gather fix: bad shape");
  ProfZone pz(a.prof, st, "out.gather_fix");
  const dim3 g3((unsigned)(((size_t)1 << a.N_log) / (2 * kTpb)), (unsigned)a.sets,
                (unsigned)a.parties);
  hipLaunchKernelGGL(gather_fix_kernel, g3, dim3(kTpb), 0, st, a);
  hip_check(hipGetLastError(), "launch gather_fix_kernel");
}

void launch_leaf_sums(hipStream_t st, const Blk* leaves, int B, int D_log, uint64_t* partials,
                      int nparts, uint64_t* sumC, uint64_t* sumT) {
  const int iters = leaf_iters(D_log);
  const size_t total = (size_t)B << D_log;
  const size_t grid = (total + (size_t)iters * kTpb - 1) / ((size_t)iters * kTpb);
  require(grid < (1ULL << 31), "leaf sums: launch too large");
  const int chunk_log = ilog2_exact((uint64_t)iters) + ilog2_exact((uint64_t)kTpb);
  hipLaunchKernelGGL(leaf_sums_kernel, dim3((unsigned)grid), dim3(kTpb), 0, st, leaves, B, D_log,
                     iters, chunk_log, nparts, partials);
  hip_check(hipGetLastError(), "launch leaf_sums_kernel");
  launch_fold_sums(st, partials, B, nparts, sumC, sumT);
}

}  // namespace pcgsq
