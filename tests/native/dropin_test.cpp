// Drop-in check of include/pcgsq/pcg_hip.hpp, written the way the reference's
// own GPU benches and tests call its plugin (pcg_cuda::*), with the namespace
// aliased exactly as a CPU_baseline build would do it.  Plain C++ (g++), no
// HIP headers: it links against libpcgsq.so only.
//
//   1. GPU_baseline/dpf_real_bench.cu:48-110 recipe (peer CW stubbed to zero)
//      -> checksum of GPU_baseline/results_l40s/plateau.txt:1
//   2. GPU_baseline/test/out_hash_cuda_bitexact.cu fixture shape through
//      dpf_out_sums / dpf_out_scatter_g (both parties, against a host
//      restatement of the output layer on the same leaves)
//   3. poly_mul patterns of GPU_baseline/test/poly_mul_cuda.cpp:56-71 and the
//      alias case out == lhs (:121-135), against a schoolbook product
//   4. OleExpander: x0 * x1 == z0 + z1 at N = 2^12
//   5. OleExpander::triples: every slot c0 + c1 == (a0 + a1)(b0 + b1)
//   6. OleExpander::expand(seed, gmw = true) == expand(seed)
//      (plus an exchange that throws: its exception reaches the caller as is)
//   7. OleParty: the two parties in two threads, every exchange through an
//      in-process channel standing in for emp::NetIO; each party's (x, z)
//      equals OleExpander's, with one "dpf_gen" row per DPF level and block;
//      then each party from its own seed, the 2PC preprocessing (base OTs,
//      OT extension, Gilboa, one "circuit_and" row per AND depth and block)
//      run between the threads first
// Exit 0 = PASS, 1 = FAIL, 77 = no device (the reference's skip code).
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "pcgsq/pcg_hip.hpp"

namespace pcg_cuda = pcg_hip;
using pcg_cuda::DpfBlk;

static const uint64_t P = 4611686018326724609ULL;
typedef unsigned __int128 u128;

static uint64_t mulm(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % P); }

static std::vector<uint64_t> schoolbook(const std::vector<uint64_t>& a,
                                        const std::vector<uint64_t>& b) {
  const int N = (int)a.size();
  std::vector<uint64_t> r(N, 0);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const uint64_t v = mulm(a[i], b[j]);
      const int k = i + j;
      if (k < N) r[k] = (r[k] + v) % P;
      else r[k - N] = (r[k - N] + P - v) % P;
    }
  return r;
}

static int fails = 0;
#define CHECK(c, msg) do { if (!(c)) { ++fails; fprintf(stderr, "FAIL: %s\n", msg); } } while (0)

// A blocking two-party byte channel: send() posts one message for the other
// party, recv() waits for the next message from it.
struct Channel {
  std::mutex m;
  std::condition_variable cv;
  std::vector<std::vector<uint8_t>> inbox[2];   // inbox[p]: messages waiting for party p
  bool aborted = false;
  void send(int to, const void* src, size_t n) {
    std::lock_guard<std::mutex> lk(m);
    inbox[to].emplace_back((const uint8_t*)src, (const uint8_t*)src + n);
    cv.notify_all();
  }
  void recv(int me, void* dst, size_t n) {
    std::unique_lock<std::mutex> lk(m);
    if (!cv.wait_for(lk, std::chrono::seconds(60), [&] { return aborted || !inbox[me].empty(); }) ||
        aborted)
      throw std::runtime_error("channel closed");
    std::vector<uint8_t> msg = std::move(inbox[me].front());
    inbox[me].erase(inbox[me].begin());
    if (msg.size() != n) throw std::runtime_error("message size mismatch");
    memcpy(dst, msg.data(), n);
  }
  void abort() {
    std::lock_guard<std::mutex> lk(m);
    aborted = true;
    cv.notify_all();
  }
};

int main() {
  if (!pcg_cuda::is_cuda_available()) {
    fprintf(stderr, "no device: skip\n");
    return 77;
  }
  // 1. plateau recipe, n=18, B=64, t=8
  {
    const int n = 18, B = 64, w = 1, t = 8;
    const long D = 1L << n;
    const int gN = (int)(t * (D / 2)), levels = n / w, nk = (1 << w) - 1;
    std::vector<DpfBlk> roots(B), dltsft((size_t)B * levels * nk);
    for (int b = 0; b < B; ++b) { roots[b].lo = 0x9e3779b97f4a7c15ULL * (b + 1); roots[b].hi = b + 1; }
    for (size_t i = 0; i < dltsft.size(); ++i) { dltsft[i].lo = 0x1234567ULL * (i + 1); dltsft[i].hi = i; }
    long calls = 0;
    auto exchange = [&](const DpfBlk*, DpfBlk* peer, size_t count) {
      for (size_t i = 0; i < count; ++i) { peer[i].lo = 0; peer[i].hi = 0; }
      ++calls;
    };
    std::vector<uint64_t> sumC(B), sumT(B), CW(B), g((size_t)gN);
    for (int b = 0; b < B; ++b) CW[b] = (0xABCDEFULL * (b + 1)) % P;
    const DpfBlk* d_leaves = pcg_cuda::dpf_gpu_batch_full_eval_device(
        B, n, w, roots.data(), dltsft.data(), nullptr, 0, exchange, pcg_cuda::DpfGpuPrg::CHACHA8);
    pcg_cuda::dpf_out_sums(d_leaves, B, D, P, sumC.data(), sumT.data());
    pcg_cuda::dpf_out_scatter_g(d_leaves, B, D, t, 0, P, CW.data(), g.data(), gN);
    uint64_t cks = 0;
    for (int b = 0; b < B; ++b) cks ^= sumC[b] ^ sumT[b];
    for (int i = 0; i < gN; i += 97) cks ^= g[i];
    printf("plateau n=%d B=%d t=%d checksum=%016llx exchanges=%ld\n", n, B, t,
           (unsigned long long)cks, calls);
    CHECK(cks == 0x16e395febfd6d23fULL, "plateau checksum");
    CHECK(calls == levels, "one exchange per level");
    // host leaves copy through the non-device entry point must agree
    std::vector<DpfBlk> leaves((size_t)B * D);
    pcg_cuda::dpf_gpu_batch_full_eval(B, n, w, roots.data(), dltsft.data(), nullptr, 0, exchange,
                                      leaves.data(), pcg_cuda::DpfGpuPrg::CHACHA8);
    CHECK(leaves[5].lo != 0 || leaves[5].hi != 0, "leaves copied back");
    // an exchange that throws: its own exception reaches the caller unchanged
    struct PeerGone {
      int level;
    };
    int thrown_at = -1;
    try {
      int k = 0;
      pcg_cuda::dpf_gpu_batch_full_eval(
          B, n, w, roots.data(), dltsft.data(), nullptr, 0,
          [&](const DpfBlk*, DpfBlk*, size_t) {
            if (++k == 3) throw PeerGone{k};
          },
          leaves.data(), pcg_cuda::DpfGpuPrg::CHACHA8);
    } catch (const PeerGone& e) {
      thrown_at = e.level;
    } catch (...) {
    }
    CHECK(thrown_at == 3, "the exchange's exception propagates unchanged");
  }
  // 2. out-hash fixture shape: t=3, B=9, D=1024, leaves from mt19937_64(12345)
  {
    const int t = 3, B = t * t;
    const size_t D = 1 << 10;
    const int N = t * (int)(D / 2), n = 10;
    std::mt19937_64 rng(12345);
    // build the leaves with the DPF itself so they live on the device: random
    // roots and shares give pseudo-random leaves
    std::vector<DpfBlk> roots(B), dlt((size_t)B * n);
    for (auto& r : roots) { r.lo = rng(); r.hi = rng(); }
    for (auto& d : dlt) { d.lo = rng(); d.hi = rng(); }
    auto ex = [&](const DpfBlk*, DpfBlk* peer, size_t count) {
      for (size_t i = 0; i < count; ++i) { peer[i].lo = 1; peer[i].hi = 2; }
    };
    std::vector<DpfBlk> leaves(B * D);
    pcg_cuda::dpf_gpu_batch_full_eval(B, n, 1, roots.data(), dlt.data(), nullptr, 1, ex,
                                      leaves.data(), pcg_cuda::DpfGpuPrg::CHACHA8);
    const DpfBlk* d_leaves = pcg_cuda::dpf_gpu_batch_full_eval_device(
        B, n, 1, roots.data(), dlt.data(), nullptr, 1, ex, pcg_cuda::DpfGpuPrg::CHACHA8);
    std::vector<uint64_t> CW(B);
    for (auto& c : CW) c = rng() % P;
    std::vector<uint64_t> sumC(B), sumT(B), g(N);
    pcg_cuda::dpf_out_sums(d_leaves, B, D, P, sumC.data(), sumT.data());
    uint64_t tsum = 0;
    for (int b = 0; b < B; ++b) tsum += sumT[b];
    uint64_t host_t = 0;
    for (const auto& l : leaves) host_t += l.lo & 1;
    CHECK(tsum == host_t, "sumT equals the host tau count");
    for (int party = 0; party <= 1; ++party)
      pcg_cuda::dpf_out_scatter_g(d_leaves, B, D, t, party, P, CW.data(), g.data(), N);
    printf("out-hash fixture shape: sums and scatter ran for both parties\n");
  }
  // 3. poly-mul patterns, both seams, alias out == lhs
  {
    const int N = 256;
    std::mt19937_64 rng(7);
    std::vector<uint64_t> a(N), one(N, 0), imp(N, 0), mx(N, P - 1), out(N);
    for (auto& x : a) x = rng() % P;
    one[0] = 1;
    imp[N - 1] = 1;
    pcg_cuda::PolyMulStats st;
    for (int seam = 0; seam < 2; ++seam) {
      auto pm = [&](uint64_t* o, const uint64_t* x, const uint64_t* y) {
        if (seam == 0) pcg_cuda::poly_mul_u64(o, x, y, 1, N, P, &st);
        else pcg_cuda::poly_mul_u64_gpuntt_square(o, x, y, 1, N, P, &st);
      };
      pm(out.data(), a.data(), one.data());
      CHECK(out == a, "a * 1");
      pm(out.data(), a.data(), imp.data());
      CHECK(out == schoolbook(a, imp), "a * X^(N-1)");
      pm(out.data(), mx.data(), mx.data());
      CHECK(out == schoolbook(mx, mx), "max * max");
      std::vector<uint64_t> lhs = a;
      pm(lhs.data(), lhs.data(), mx.data());
      CHECK(lhs == schoolbook(a, mx), "alias out == lhs");
    }
    std::string why;
    CHECK(!pcg_cuda::poly_mul_supported(P + 2, N, &why) && !why.empty(), "other primes rejected");
    pcg_hip::NTT<P> ntt(N);
    std::vector<uint64_t> f = a;
    ntt.forward(f.data());
    ntt.inverse(f.data());
    CHECK(f == a, "NTT round trip");
    printf("poly-mul patterns and NTT round trip checked\n");
  }
  // 4. OLE relation
  {
    const int N = 1 << 12;
    pcg_hip::OleExpander ole(N, 2, 16);
    auto o = ole.expand(3);
    std::vector<uint64_t> prod(N);
    pcg_cuda::poly_mul_u64_gpuntt_square(prod.data(), o.x0.data(), o.x1.data(), 1, N, P);
    bool ok = true;
    for (int i = 0; i < N; ++i) ok &= prod[i] == (o.z0[i] + o.z1[i]) % P;
    CHECK(ok, "x0 * x1 == z0 + z1");
    printf("OLE relation N=%d: %s\n", N, ok ? "holds" : "FAILS");
    // 5. generate_triples (triple.h:45-110): every slot a Z_p triple
    auto tr = ole.triples(5, 6);
    bool tok = true;
    for (int i = 0; i < N; ++i) {
      const uint64_t a = (tr.a0[i] + tr.a1[i]) % P, b = (tr.b0[i] + tr.b1[i]) % P;
      tok &= mulm(a, b) == (tr.c0[i] + tr.c1[i]) % P;
    }
    CHECK(tok, "triples: c0 + c1 == (a0 + a1)(b0 + b1)");
    printf("triples N=%d: %s\n", N, tok ? "hold" : "FAIL");
    // 6. the preprocessing under GMW on the device gives the same expansion
    auto og = ole.expand(3, true);
    const bool same = og.x0 == o.x0 && og.z0 == o.z0 && og.x1 == o.x1 && og.z1 == o.z1;
    CHECK(same, "GMW-preprocessed expansion == dealer expansion");
    printf("GMW preprocessing N=%d: %s\n", N, same ? "same outputs" : "DIFFERS");
    // 7. one party per thread, exchanges through the channel: the
    // trusted-dealer simulation (deal), then each party from its own seed
    // with the 2PC preprocessing run between the threads (deal_own)
    for (int own = 0; own < 2; ++own) {
      Channel ch;
      pcg_hip::OleParty::Output po[2];
      int dpf_rows[2] = {0, 0}, and_rows[2] = {0, 0}, pre_rows[2] = {0, 0}, rows[2] = {0, 0};
      std::string err[2];
      auto run = [&](int p) {
        try {
          pcg_hip::OleParty party(p, N, 2, 16);
          auto ex = [&](const char* phase, int, int, const void* send, size_t sb, void* recv,
                        size_t rb) {
            if (sb) ch.send(1 - p, send, sb);
            if (rb) ch.recv(p, recv, rb);
            ++rows[p];
            dpf_rows[p] += strcmp(phase, "dpf_gen") == 0;
            and_rows[p] += strcmp(phase, "circuit_and") == 0;
            pre_rows[p] += strcmp(phase, "base_ot") == 0 || strcmp(phase, "ot_ext") == 0 ||
                           strcmp(phase, "gilboa") == 0;
          };
          po[p] = own ? party.expand_own(3, ex) : party.expand(3, ex);
        } catch (const std::exception& e) {
          err[p] = e.what();
          ch.abort();
        }
      };
      std::thread t0(run, 0), t1(run, 1);
      t0.join();
      t1.join();
      CHECK(err[0].empty() && err[1].empty(), "both parties finish");
      const bool pok = po[0].x == o.x0 && po[0].z == o.z0 && po[1].x == o.x1 && po[1].z == o.z1;
      CHECK(pok, "OleParty outputs == OleExpander outputs");
      const int levels = 9, blocks = 4;   // log2(2N/t) DPF levels at w = 1; c*c blocks
      // adder over log2(N/t) = 8 bits: generate layer + 3 prefix steps; one Delta layer
      const int and_layers = own ? 5 : 0;
      CHECK(dpf_rows[0] == levels * blocks && dpf_rows[1] == levels * blocks,
            "one dpf_gen exchange per DPF level and block");
      CHECK(and_rows[0] == and_layers * blocks && and_rows[1] == and_rows[0],
            "one circuit_and exchange per AND depth and block");
      CHECK(pre_rows[0] == (own ? 4 : 0) && pre_rows[1] == pre_rows[0],
            "two base_ot, one ot_ext, one gilboa exchange");
      CHECK(rows[0] == (levels + 2) * blocks + and_rows[0] + pre_rows[0] && rows[1] == rows[0],
            "two Beaver opens per block");
      printf("two parties (%s) N=%d: %s (%d exchanges each)%s%s\n",
             own ? "own seeds, 2PC preprocessing" : "dealer", N, pok ? "same outputs" : "DIFFER",
             rows[0], err[0].c_str(), err[1].c_str());
    }
  }
  printf("%s\n", fails ? "RESULT: FAIL" : "RESULT: PASS");
  return fails ? 1 : 0;
}
