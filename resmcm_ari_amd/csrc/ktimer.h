// ktimer.h - per-kernel timing with HIP events recorded on the stream the kernels run on.
// Disabled by default; when enabled every timed launch is bracketed by two events and the
// elapsed times are folded into per-name totals at the next flush (after a stream sync).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "common.h"

namespace smcm {

struct KernelTimer {
  struct Agg {
    double ms = 0;
    int64_t count = 0;
    double bytes = 0;  // algorithmic bytes of the timed launches
  };
  bool enabled = false;
  std::map<std::string, Agg> agg;

  struct Rec {
    const char* name;
    hipEvent_t a, b;
    double bytes;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;

  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    SMCM_HIP(hipEventCreate(&e));
    return e;
  }
  // returns a token for stop(); -1 when disabled
  int start(const char* name, hipStream_t st) {
    if (!enabled) return -1;
    Rec r{name, ev(), ev(), 0.0};
    SMCM_HIP(hipEventRecord(r.a, st));
    pending.push_back(r);
    return (int)pending.size() - 1;
  }
  void stop(int tok, hipStream_t st, double bytes) {
    if (tok < 0) return;
    pending[tok].bytes = bytes;
    SMCM_HIP(hipEventRecord(pending[tok].b, st));
  }
  // algorithmic bytes of a timed launch known only after it was enqueued (before the flush)
  void set_bytes(int tok, double bytes) {
    if (tok >= 0) pending[tok].bytes = bytes;
  }
  // call after the stream has been synchronised
  void flush() {
    for (auto& r : pending) {
      float ms = 0;
      SMCM_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      Agg& g = agg[r.name];
      g.ms += ms;
      g.count += 1;
      g.bytes += r.bytes;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
  void reset() {
    pending.clear();
    agg.clear();
  }
  ~KernelTimer() {
    for (auto& r : pending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace smcm
