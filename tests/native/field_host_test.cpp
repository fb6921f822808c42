// Host-side check of the device field arithmetic (the same inline functions
// the kernels use, compiled for the host): every result is compared with
// exact unsigned __int128 arithmetic on edge values and random inputs.
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>

#include "device_common.hpp"

using namespace pcgsq;
typedef unsigned __int128 u128;

static uint64_t rng_state = 0x1234567887654321ULL;
static uint64_t rnd() {
  uint64_t z = (rng_state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

int main(int argc, char** argv) {
  long iters = argc > 1 ? atol(argv[1]) : 2000000;
  long fails = 0;
  const uint64_t edges[] = {0, 1, 2, kP - 1, kP - 2, kP / 2, (1ULL << 62) - 1, 1ULL << 61,
                            kC62, kC62 + 1, 0xffffffffULL, 0x100000000ULL};
  const int ne = sizeof(edges) / sizeof(edges[0]);
  for (long it = 0; it < iters + ne * ne; ++it) {
    uint64_t a, b;
    if (it < ne * ne) { a = edges[it / ne] % kP; b = edges[it % ne] % kP; }
    else { a = rnd() % kP; b = rnd() % kP; }
    const uint64_t want = (uint64_t)(((u128)a * b) % kP);
    if (fmul(a, b) != want) { if (fails++ < 5) printf("fmul %llu %llu\n", (unsigned long long)a, (unsigned long long)b); }
    const uint64_t wq = (uint64_t)(((u128)b << 64) / kP);
    const uint64_t x = (it & 1) ? rnd() : a;            // Shoup takes any 64-bit x
    const uint64_t s = shoup(x, b, wq);
    if (s != (uint64_t)(((u128)x * b) % kP)) { if (fails++ < 5) printf("shoup\n"); }
    const uint64_t sl = shoup_lazy(x, b, wq);
    if (sl >= 2 * kP || sl % kP != s) { if (fails++ < 5) printf("shoup_lazy range\n"); }
    // the NTT first passes' constant formed from w alone: wq or wq - 1, and a
    // product with it in [0, 3P)
    const uint64_t wa = shoup_w_approx(b);
    if (wa != wq && wa + 1 != wq) { if (fails++ < 5) printf("shoup_w_approx\n"); }
    const uint64_t sa = shoup_lazy(x, b, wa);
    if (sa >= 3 * kP || sa % kP != s) { if (fails++ < 5) printf("shoup approx range\n"); }
    const uint64_t y = rnd();
    if (freduce64(y) != y % kP) { if (fails++ < 5) printf("freduce64\n"); }
    if (mulhi64(y, x) != (uint64_t)(((u128)y * x) >> 64)) { if (fails++ < 5) printf("mulhi64\n"); }
    if (mulP(y) != y * kP) { if (fails++ < 5) printf("mulP\n"); }
    if (fadd(a, b) != (uint64_t)(((u128)a + b) % kP)) { if (fails++ < 5) printf("fadd\n"); }
    if (fsub(a, b) != (uint64_t)(((u128)a + kP - b) % kP)) { if (fails++ < 5) printf("fsub\n"); }
  }
  printf("field_host_test: %ld iterations, %ld failures\n", iters, fails);
  return fails ? 1 : 0;
}
