// common.hpp -- shared host/device helpers of librtm_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <string>

#define RTM_EXPORT extern "C" __attribute__((visibility("default")))

namespace erm {

// 10th-order staggered first-derivative weights (W/rtmlib-mod.f90:15): the
// quotient of two exactly representable singles, rounded once, as the
// Fortran default-real literals are.
constexpr float C1 = 35.0f / 294912.0f;
constexpr float C2 = -405.0f / 229376.0f;
constexpr float C3 = 567.0f / 40960.0f;
constexpr float C4 = -735.0f / 8192.0f;
constexpr float C5 = 19845.0f / 16384.0f;

enum Status { OK = 0, E_ARG = -1, E_NOMEM = -2, E_HIP = -3, E_NONFINITE = -4, E_CFL = -5 };

// Thread-local message of the last failure (rtm_hip_last_error).
void set_error(const char* fmt, ...);
const char* last_error();

// Returns E_HIP with a message when a HIP call fails.
#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::erm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                      \
      return ::erm::E_HIP;                                                             \
    }                                                                                  \
  } while (0)

#define STCHK(expr)            \
  do {                         \
    int s_ = (expr);           \
    if (s_ != 0) return s_;    \
  } while (0)

// Owning device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t n, const char* what) {
    release();
    if (n == 0) return OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc(%zu bytes) for %s failed: %s", n, what, hipGetErrorString(e));
      return E_NOMEM;
    }
    bytes = n;
    return OK;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// Round up to a multiple of m.
inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Column stride of device fields: 32 floats (128 B) so every column starts
// on a cache-line boundary.
inline int device_ldz(int nz) { return round_up(nz, 32); }

}  // namespace erm
