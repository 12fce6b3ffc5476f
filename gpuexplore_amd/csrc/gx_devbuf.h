// gx_devbuf.h -- an owning device allocation (hipMalloc / hipFree), move-only.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace gx {

struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p) { o.p = nullptr; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  // releases what the buffer held, then allocates `bytes` (8 when 0)
  hipError_t alloc(size_t bytes) {
    reset();
    return hipMalloc(&p, bytes ? bytes : 8);
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

}  // namespace gx
