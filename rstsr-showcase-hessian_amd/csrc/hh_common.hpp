// Common definitions for libhesship: status codes, error capture, device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hesship.h"

namespace hh {

// Exception carrying an ABI status code; converted to a return code at the C boundary.
struct Error : std::runtime_error {
  int32_t code;
  Error(int32_t c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string& msg);

#define HH_HIP(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      throw ::hh::Error(HH_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_) \
                                        + " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define HH_REQUIRE(cond, msg)                                                       \
  do {                                                                              \
    if (!(cond)) throw ::hh::Error(HH_ERR_INVALID, std::string("invalid argument: ") + (msg)); \
  } while (0)

// Owning device allocation (grown on demand, never shrunk).
// Counts frees of device buffers: a captured graph that refers to library buffers is valid only
// while the count is unchanged.
inline std::atomic<uint64_t>& devbuf_generation() {
  static std::atomic<uint64_t> g{0};
  return g;
}

// Counts changes of component state that a captured response graph bakes in: re-prepared
// orbitals (and with them the reduced-precision copies made on first use), batch sizes, the fused
// grid path, profiling events.  A graph captured under one count is not replayed under another.
inline std::atomic<uint64_t>& response_state_epoch() {
  static std::atomic<uint64_t> g{0};
  return g;
}
inline void bump_response_state() { ++response_state_epoch(); }

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) {
      (void)hipFree(ptr);
      ++devbuf_generation();
    }
    ptr = nullptr;
    bytes = 0;
  }
  void* reserve(size_t n) {
    if (n <= bytes) return ptr;
    release();
    if (n == 0) return nullptr;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
      ptr = nullptr;
      throw Error(HH_ERR_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e));
    }
    bytes = n;
    return ptr;
  }
  template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// compute units of the current device (256 if the query fails)
int cu_count();

}  // namespace hh

// Wrap a C-ABI body: translate exceptions into status codes.
#define HH_ABI_BEGIN try {
#define HH_ABI_END                                   \
  return HH_OK;                                      \
  }                                                  \
  catch (const ::hh::Error& e) {                     \
    ::hh::set_last_error(e.what());                  \
    return e.code;                                   \
  }                                                  \
  catch (const std::exception& e) {                  \
    ::hh::set_last_error(e.what());                  \
    return HH_ERR_INTERNAL;                          \
  }
