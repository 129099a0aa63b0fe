// Shared host/device helpers of libfrpa: status handling, device buffers, launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "frpa.h"

namespace frpa {

void set_error(const char *fmt, ...);
void clear_error();

struct Error {
  int code;
};

#define FRPA_REQUIRE(cond, code, ...)                                                                          \
  do {                                                                                                         \
    if (!(cond)) {                                                                                             \
      ::frpa::set_error(__VA_ARGS__);                                                                          \
      throw ::frpa::Error{code};                                                                               \
    }                                                                                                          \
  } while (0)

#define FRPA_HIP(call)                                                                                         \
  do {                                                                                                         \
    hipError_t e_ = (call);                                                                                    \
    if (e_ != hipSuccess) {                                                                                    \
      ::frpa::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_));           \
      throw ::frpa::Error{FRPA_ERR_HIP};                                                                       \
    }                                                                                                          \
  } while (0)

// Propagates the status of a nested C-ABI call (its message is already set).
#define FRPA_CHECK_OK(call)                                                                                    \
  do {                                                                                                         \
    const int rc_ = (call);                                                                                    \
    if (rc_ != FRPA_OK)                                                                                        \
      throw ::frpa::Error{rc_};                                                                                \
  } while (0)

// Wraps a C-ABI body: translates exceptions into status codes.
template <class F>
int guarded(F &&f) {
  try {
    clear_error();
    f();
    return FRPA_OK;
  } catch (const Error &e) {
    return e.code;
  } catch (const std::exception &e) {
    set_error("exception: %s", e.what());
    return FRPA_ERR_INVALID;
  }
}

// Owning device buffer (hipMalloc / hipFree).
template <class T>
struct DevBuf {
  T *ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr)
      (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && ptr)
      return;
    release();
    if (count == 0)
      return;
    FRPA_HIP(hipMalloc((void **)&ptr, count * sizeof(T)));
    n = count;
  }
  // Per-call scratch: reallocated only when a call needs more than it holds (n is then the
  // capacity), so that calls with smaller sizes never free memory a captured graph still uses.
  void grow(size_t count) {
    if (count > n || !ptr)
      alloc(std::max<size_t>(count, 1));
  }
  void upload(const T *host, size_t count) {
    alloc(count);
    if (count)
      FRPA_HIP(hipMemcpy(ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
};

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("launch of %s failed: %s", what, hipGetErrorString(e));
    throw Error{FRPA_ERR_HIP};
  }
}

// Fill kernel (used instead of hipMemsetAsync so that the calls can be captured in a graph).
void fill_f64(double *p, size_t n, double v, hipStream_t s);

} // namespace frpa
