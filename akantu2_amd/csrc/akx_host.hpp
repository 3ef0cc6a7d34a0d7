// akx_host.hpp -- host-side utilities shared by the translation units of the
// library: status-code errors, HIP error checks, device buffers.
#pragma once

#include "../../include/akantu_mi355x.h"
#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

namespace akx {

// ---------------------------------------------------------------------------
// Errors: C++ exceptions inside, status codes across the ABI.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

extern thread_local std::string g_last_error;

#define HIPCHECK(x)                                                            \
  do {                                                                         \
    hipError_t err_ = (x);                                                     \
    if (err_ != hipSuccess)                                                    \
      throw Error(err_ == hipErrorOutOfMemory ? AKX_ERR_MEMORY : AKX_ERR_HIP,  \
                  std::string(#x) + ": " + hipGetErrorString(err_));           \
  } while (0)

inline void require(bool c, const std::string &msg) {
  if (!c)
    throw Error(AKX_ERR_INVALID, msg);
}

// ---------------------------------------------------------------------------
// Device buffer
// ---------------------------------------------------------------------------
template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p)
      return;
    release();
    if (count == 0)
      return;
    HIPCHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  // grow only (keeps the buffer, and its address, when it is big enough)
  void reserve(size_t count) {
    if (count > n || !p)
      alloc(count > 0 ? count : 1);
  }
  size_t bytes() const { return n * sizeof(T); }
};


// Run f and turn any exception into a status code (nothing throws across
// the C ABI).
template <class F> inline int guard(F &&f) {
  try {
    f();
    return AKX_OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return AKX_ERR_MEMORY;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return AKX_ERR_INVALID;
  } catch (...) {
    g_last_error = "unknown error";
    return AKX_ERR_INVALID;
  }
}


} // namespace akx
