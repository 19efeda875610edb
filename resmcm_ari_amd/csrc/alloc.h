// alloc.h - stream-ordered device memory for one context.
//
// Small blocks go through hipMallocAsync / hipFreeAsync on the context's stream.  Large blocks
// (>= kCacheMin) are kept mapped after they are freed and handed out again (best fit, at most
// 1.5x the request), up to cache_cap bytes: remapping tens of GB of HBM for every call costs far
// more than the products themselves.  All work of a context runs on its one stream, so a block
// freed here can be reused by any later allocation in stream order.  When an allocation fails
// the oldest cached blocks, then the whole cache, are released, this stream synchronised and the
// device pool trimmed before each further try (no device-wide wait: other contexts' streams keep
// running).  An allocation also fails
// when it would leave less than `headroom` bytes of the device free for the runtime.
// Freed small blocks are kept too, up to small_cache_cap bytes (0 = none), so that repeated calls
// make no device-pool calls in their steady state: each costs microseconds (a third of a 1.2 ms
// chain of three products went to them), and the runtime serialises them, so that one mapping a
// new large block for another thread holds every other allocation, small ones included, for as
// long as the mapping takes (seconds for tens of GB).
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace smcm {

struct DevAlloc {
  static constexpr size_t kCacheMin = (size_t)64 << 20;
  hipStream_t stream = nullptr;
  // Arena mode (arena != nullptr): every block is carved from the caller's device region
  // [arena, arena + arena_cap) and nothing is ever allocated or freed on the device.  Blocks are
  // handed out in order; freeing the most recent block returns its bytes (scoped buffers are
  // freed in reverse order), other frees keep their bytes until the allocator is dropped.  An
  // allocation beyond arena_cap throws (EINVAL) before any kernel could use the block.
  unsigned char* arena = nullptr;
  size_t arena_cap = 0, arena_top = 0, arena_high = 0;
  std::vector<size_t> arena_stack;
  size_t live_bytes = 0;
  size_t peak_bytes = 0;
  size_t cached_bytes = 0;  // large blocks (>= kCacheMin) in the cache
  size_t cache_cap = SIZE_MAX;
  size_t cached_small = 0;  // small blocks in the cache
  size_t small_cache_cap = (size_t)256 << 20;
  std::unordered_map<void*, size_t> sizes;
  std::multimap<size_t, std::pair<void*, uint64_t>> cache;  // size -> (block, insertion order)
  uint64_t cache_seq = 0;
  // called when an allocation fails, before this allocator's own cache is released: frees what
  // idle allocators tied to this one (a context's batch workers) keep cached
  std::function<void()> on_oom;
  // A block kept all-zero between uses (the dense SpGEMM tier's accumulators): a user returns
  // every byte it changed to zero, and marks the block dirty while it may not have; a dirty block
  // is cleared before its next use.  Freed with the cache.
  // While a product holds pointers into the block (zero_pins > 0), release_cache() keeps it: an
  // allocation failing in the middle of that product releases the cache and trims the device
  // pool, which would otherwise unmap the block under the kernels still to run on it.
  void* zero_block = nullptr;
  size_t zero_bytes = 0;
  bool zero_dirty = false;
  int zero_pins = 0;
  // Failure injection for tests of the recovery path: when oom_inject > 0, the oom_inject-th
  // device allocation from now on takes the out-of-memory path once (as if hipMallocAsync had
  // failed) before its real attempt.
  int64_t oom_inject = 0;
  int64_t oom_retries = 0;  // allocations that failed once and were tried again after a release
  // Device memory kept free for the runtime: the scratch of kernels with a private segment is
  // allocated by the runtime when such a kernel is dispatched, and a failure there aborts the
  // queue (and the process).  An allocation that would grow the device pool into this reserve
  // takes the out-of-memory path instead (release, trim, retry, then an error the caller sees).
  size_t headroom = (size_t)2 << 30;

  // can `bytes` be allocated without leaving less than `headroom` free on the device?  Callers
  // hold pool_mutex().
  bool room_for(size_t bytes) const {
    if (headroom == 0) return true;
    int dev = 0;
    hipMemPool_t pool;
    uint64_t reserved = 0, used = 0;
    size_t freeb = 0, total = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess ||
        hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) != hipSuccess ||
        hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) != hipSuccess) {
      (void)hipGetLastError();
      return true;  // no pool figures: the allocation itself decides
    }
    // A small block fits in the pool's free part when used + bytes is within its reservation.  A
    // large one is checked as if it were mapped anew: the pool's free part may be held in
    // pieces smaller than the request (large freed blocks stay in this allocator's cache), and a
    // pool that grows past the reserve leaves the runtime nothing for kernel scratch.
    if (bytes < kCacheMin && used + bytes <= reserved) return true;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
      (void)hipGetLastError();
      return true;
    }
    const uint64_t grow = bytes >= kCacheMin ? bytes : used + bytes - reserved;
    return grow + headroom <= freeb;
  }
  // one check-and-allocate at a time in the process: concurrent allocators (batch workers,
  // server workers) cannot pass the check together and then share the reserve out
  static std::mutex& pool_mutex() {
    static std::mutex mu;
    return mu;
  }
  hipError_t try_alloc(void** p, size_t bytes) {
    std::lock_guard<std::mutex> lk(pool_mutex());
    if (!room_for(bytes)) return hipErrorOutOfMemory;
    return hipMallocAsync(p, bytes, stream);
  }

  void uncount(size_t sz) { (sz >= kCacheMin ? cached_bytes : cached_small) -= sz; }

  // frees the oldest cached blocks until at least `want` bytes are out; returns the bytes freed
  size_t evict_oldest(size_t want) {
    size_t got = 0;
    while (got < want && !cache.empty()) {
      auto old = cache.begin();
      for (auto it = cache.begin(); it != cache.end(); ++it)
        if (it->second.second < old->second.second) old = it;
      (void)hipFreeAsync(old->second.first, stream);
      got += old->first;
      uncount(old->first);
      cache.erase(old);
    }
    return got;
  }

  // hipMallocAsync on this stream.  On failure the oldest cached blocks (as many bytes as the
  // request) go back to the device first, then the whole cache; the pool is trimmed before
  // each further try.  Blocks kept for reuse are what makes repeated products fast: mapping
  // them again costs more than the products.
  // With SMCM_TRACE set, an allocation that took the out-of-memory path or 50 ms or more is
  // reported on stderr.
  void* raw_alloc(size_t bytes) {
    static const bool trace = std::getenv("SMCM_TRACE") != nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t retries0 = oom_retries;
    struct Report {
      const DevAlloc& al;
      size_t bytes;
      std::chrono::steady_clock::time_point t0;
      int64_t retries0;
      ~Report() {
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms >= 50.0 || al.oom_retries != retries0)
          fprintf(stderr, "[smcm] device allocation of %zu MiB: %.1f ms%s\n", bytes >> 20, ms,
                  al.oom_retries != retries0 ? " (out-of-memory path: cached blocks released, pool trimmed)" : "");
      }
    };
    std::unique_ptr<Report> report(trace ? new Report{*this, bytes, t0, retries0} : nullptr);
    void* p = nullptr;
    hipError_t e = (oom_inject > 0 && --oom_inject == 0) ? hipErrorOutOfMemory : try_alloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
      oom_retries++;
      (void)hipGetLastError();
      if (evict_oldest(bytes) > 0) {
        (void)hipStreamSynchronize(stream);
        trim_pool();
        e = try_alloc(&p, bytes);
      }
    }
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      if (on_oom) on_oom();
      release_cache();
      (void)hipStreamSynchronize(stream);
      trim_pool();
      e = try_alloc(&p, bytes);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw HipError((int)e, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                 hipGetErrorString(e) + " (" + std::to_string(headroom >> 20) +
                                 " MiB of the device are kept free for the runtime)");
    }
    return p;
  }
  void free_zero_block() {
    if (zero_block) (void)hipFreeAsync(zero_block, stream);
    zero_block = nullptr;
    zero_bytes = 0;
    zero_dirty = false;
  }

  void* alloc(size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    if (bytes == 0) bytes = 256;
    if (arena) {
      if (bytes > arena_cap - arena_top)
        throw HipError(-22, "scratch exhausted: " + std::to_string(arena_top + bytes) + " bytes needed, " +
                                std::to_string(arena_cap) + " available");
      arena_stack.push_back(arena_top);
      void* q = arena + arena_top;
      arena_top += bytes;
      if (arena_top > arena_high) arena_high = arena_top;
      return q;
    }
    void* p = nullptr;
    size_t got = bytes;
    if (bytes >= kCacheMin || small_cache_cap > 0) {
      auto it = cache.lower_bound(bytes);
      if (it != cache.end() && it->first <= bytes + bytes / 2) {
        p = it->second.first;
        got = it->first;
        uncount(got);
        cache.erase(it);
      }
    }
    if (!p) p = raw_alloc(bytes);
    sizes[p] = got;
    live_bytes += got;
    if (live_bytes > peak_bytes) peak_bytes = live_bytes;
    return p;
  }
  template <class T>
  T* alloc_n(int64_t n) {
    return (T*)alloc(sizeof(T) * (size_t)(n > 0 ? n : 1));
  }
  // never throws (it runs in destructors, also while an error unwinds)
  void free(void* p) {
    if (!p) return;
    if (arena) {
      if (!arena_stack.empty() && (unsigned char*)p == arena + arena_stack.back()) {
        arena_top = arena_stack.back();
        arena_stack.pop_back();
      }
      return;
    }
    auto it = sizes.find(p);
    const size_t sz = it != sizes.end() ? it->second : 0;
    if (it != sizes.end()) sizes.erase(it);
    live_bytes = live_bytes > sz ? live_bytes - sz : 0;
    if (sz >= kCacheMin ? cached_bytes + sz <= cache_cap : sz > 0 && cached_small + sz <= small_cache_cap) {
      cache.emplace(sz, std::make_pair(p, cache_seq++));
      (sz >= kCacheMin ? cached_bytes : cached_small) += sz;
      return;
    }
    if (hipFreeAsync(p, stream) != hipSuccess) (void)hipGetLastError();
  }
  static void trim_pool() {
    int dev = 0;
    hipMemPool_t pool;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
      (void)hipMemPoolTrimTo(pool, 0);
    (void)hipGetLastError();
  }
  void release_cache() {
    for (auto& kv : cache) (void)hipFreeAsync(kv.second.first, stream);
    cache.clear();
    cached_bytes = 0;
    cached_small = 0;
    if (zero_pins == 0) free_zero_block();
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
    if (p && a) a->free(p);
    p = nullptr;
  }
  T* release() {
    T* q = p;
    p = nullptr;
    return q;
  }
};

}  // namespace smcm
