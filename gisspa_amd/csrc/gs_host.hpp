// gs_host.hpp -- host-side helpers shared by the C ABI sources: an owning
// device buffer and the status-returning error checks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/gisspa.h"

namespace gs {

// Device allocation that grows on demand and frees itself.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int ensure(size_t n) {
    if (n <= bytes) return GS_OK;
    release();
    if (n == 0) return GS_OK;
    if (hipMalloc(&p, n) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return GS_ERR_NOMEM;
    }
    bytes = n;
    return GS_OK;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

inline int hip_ok(hipError_t e) {
  if (e == hipSuccess) return GS_OK;
  (void)hipGetLastError();
  return GS_ERR_DEVICE;
}

}  // namespace gs

#define HIPCHK(x)                       \
  do {                                  \
    int _r = gs::hip_ok(x);             \
    if (_r != GS_OK) return _r;         \
  } while (0)
#define CHK(x)                          \
  do {                                  \
    int _r = (x);                       \
    if (_r != GS_OK) return _r;         \
  } while (0)
