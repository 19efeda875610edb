// alloc.h - stream-ordered device memory (hipMallocAsync on the context's stream, with the
// device's default pool configured to keep freed blocks for reuse).
#pragma once

#include "common.h"

namespace smcm {

struct DevAlloc {
  hipStream_t stream = nullptr;
  size_t live_bytes = 0;
  size_t peak_bytes = 0;

  void* alloc(size_t bytes) {
    if (bytes == 0) bytes = 256;
    void* p = nullptr;
    hipError_t e = hipMallocAsync(&p, bytes, stream);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw HipError((int)e, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                 hipGetErrorString(e));
    }
    live_bytes += bytes;
    if (live_bytes > peak_bytes) peak_bytes = live_bytes;
    return p;
  }
  template <class T>
  T* alloc_n(int64_t n) {
    return (T*)alloc(sizeof(T) * (size_t)(n > 0 ? n : 1));
  }
  void free(void* p, size_t bytes = 0) {
    if (!p) return;
    SMCM_HIP(hipFreeAsync(p, stream));
    live_bytes = live_bytes > bytes ? live_bytes - bytes : 0;
  }
};

// RAII device buffer bound to an allocator
template <class T>
struct DevBuf {
  DevAlloc* a = nullptr;
  T* p = nullptr;
  int64_t n = 0;
  DevBuf() = default;
  DevBuf(DevAlloc& al, int64_t count) : a(&al), p(al.alloc_n<T>(count)), n(count) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : a(o.a), p(o.p), n(o.n) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    reset();
    a = o.a;
    p = o.p;
    n = o.n;
    o.p = nullptr;
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p && a) a->free(p, sizeof(T) * (size_t)(n > 0 ? n : 1));
    p = nullptr;
  }
  T* release() {
    T* q = p;
    p = nullptr;
    return q;
  }
};

}  // namespace smcm
