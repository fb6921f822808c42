// C entry points around the reference's own header-only primitives, compiled
// from the reference tree where it lies (oracle/Makefile target `ref`).  Used
// only by tests to check the oracle's restatement; the library built from it
// lands in oracle/_ref/ and is never part of the product.
//
// Wrapped reference code:
//   CPU_baseline/ntt.h      NTT<P>::forward / inverse / poly_mul
//   CPU_baseline/chacha8.h  ChaCha8PRG (set_nonce, set_counter, fill_blocks)
//   GPU_baseline/common/ntt_roots.h  negacyclic_psi
#include <cstdint>
#include <cstring>
#include <vector>

#include "ntt.h"        // CPU_baseline/ntt.h (includes common/ffp.h)
#include "chacha8.h"    // CPU_baseline/chacha8.h
#include "common/ntt_roots.h"

namespace {
constexpr uint64_t kP = 4611686018326724609ULL;
using F = FFp<kP>;
}

extern "C" {

int ref_ntt_forward(uint64_t* a, int N) {
    NTT<kP> ntt(N);
    std::vector<F> v(N);
    for (int i = 0; i < N; ++i) v[i] = F(a[i]);
    ntt.forward(v.data());
    for (int i = 0; i < N; ++i) a[i] = v[i].val();
    return 0;
}

int ref_ntt_inverse(uint64_t* a, int N) {
    NTT<kP> ntt(N);
    std::vector<F> v(N);
    for (int i = 0; i < N; ++i) v[i] = F(a[i]);
    ntt.inverse(v.data());
    for (int i = 0; i < N; ++i) a[i] = v[i].val();
    return 0;
}

int ref_poly_mul(uint64_t* c, const uint64_t* a, const uint64_t* b, int N) {
    NTT<kP> ntt(N);
    std::vector<F> va(N), vb(N), vc(N);
    for (int i = 0; i < N; ++i) { va[i] = F(a[i]); vb[i] = F(b[i]); }
    ntt.poly_mul(vc.data(), va.data(), vb.data());
    for (int i = 0; i < N; ++i) c[i] = vc[i].val();
    return 0;
}

uint64_t ref_negacyclic_psi(int N) { return pcg_cuda::negacyclic_psi(kP, N); }

void ref_chacha8_blocks(const uint8_t key[32], uint64_t nonce, uint64_t ctr,
                        uint8_t* out, size_t nblocks) {
    ChaCha8PRG prg(key);
    prg.set_nonce(nonce);
    prg.set_counter(ctr);
    prg.fill_blocks(out, nblocks);
}

}  // extern "C"
