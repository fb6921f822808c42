// Host-side internals shared by the translation units of libpcgsq.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace pcgsq {

constexpr uint64_t kPrime = 4611686018326724609ULL;

// Status codes of the C ABI (mirrors include/pcgsq.h).
enum Status : int {
  kOk = 0,
  kErrInvalid = 1,      // bad argument
  kErrUnsupported = 2,  // shape/prime not supported (the reference's "exit 77" case)
  kErrDevice = 3,       // HIP runtime error
  kErrNoMemory = 4,     // device allocation failed / over budget
  kErrCallback = 5,     // an exchange callback reported failure
  kErrInternal = 6,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory) ? kErrNoMemory : kErrDevice;
  throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

inline void require(bool ok, const std::string& msg, int code = kErrInvalid) {
  if (!ok) throw Error(code, msg);
}

void set_last_error(const std::string& msg);

// Run `f`, translating exceptions into a status code + last error string so
// no C++ exception crosses the C ABI.
template <class F>
int guarded(F&& f) {
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return kErrNoMemory;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return kErrInternal;
  }
}

inline int ilog2_exact(uint64_t x) {
  int r = 0;
  while ((1ULL << r) < x) ++r;
  return ((1ULL << r) == x) ? r : -1;
}

// Host field helpers (u128 reduction; setup paths only).
inline uint64_t hmul(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) % kPrime);
}
inline uint64_t hpow(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  b %= kPrime;
  while (e) {
    if (e & 1) r = hmul(r, b);
    b = hmul(b, b);
    e >>= 1;
  }
  return r;
}
inline uint64_t hinv(uint64_t a) { return hpow(a, kPrime - 2); }
// floor(w * 2^64 / P): the Shoup companion of a constant multiplier w < P.
inline uint64_t shoup_companion(uint64_t w) {
  return (uint64_t)(((unsigned __int128)w << 64) / kPrime);
}

// Device buffer that grows on demand and is released with its owner.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void ensure(size_t need, const char* what) {
    if (need <= bytes) return;
    release();
    if (need == 0) return;
    hip_check(hipMalloc(&ptr, need), what);
    bytes = need;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Page-locked host buffer that grows on demand (staging of exchanged bytes).
struct PinnedBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void ensure(size_t need, const char* what) {
    if (need <= bytes) return;
    release();
    if (need == 0) return;
    hip_check(hipHostMalloc(&ptr, need, hipHostMallocDefault), what);
    bytes = need;
  }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
  ~PinnedBuf() { release(); }
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
};

}  // namespace pcgsq
