// common.h - shared types and wave64 helpers for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace smcm {

constexpr int kWave = 64;

// Device CSR: int64 row pointers, int32 column ids (type-local), optional uint64 values
// (nullptr = every entry has value 1, used for base relations).
struct Csr {
  int32_t nrows = 0, ncols = 0;
  int64_t nnz = 0;
  int64_t* ptr = nullptr;
  int32_t* col = nullptr;
  uint64_t* val = nullptr;
  bool sorted = true;     // every row ascending (non-strict)
  bool distinct = false;  // no column repeated within a row (products; strictly sorted host input)
};

struct HipError : std::runtime_error {
  int code;
  HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SMCM_HIP(expr)                                                                      \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw ::smcm::HipError((int)_e, std::string(#expr) + ": " + hipGetErrorString(_e) +   \
                                          " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define SMCM_LAUNCH_CHECK() SMCM_HIP(hipGetLastError())

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Pinned host buffer of the calling thread for small device-to-host readbacks (counters, sizes).
// A copy into pageable memory blocks the host until the stream has drained; a copy into pinned
// memory is queued like a kernel, so several readbacks share one stream synchronisation.  Each
// thread runs one library call at a time and reads its values right after that call's
// synchronisation.  Allocated on first use and never freed (1 KB per host thread that runs a
// product): freeing it from a thread-exit destructor could run after the HIP runtime has shut down.
constexpr size_t kHostStageBytes = 1024;
inline unsigned char* host_stage() {
  thread_local unsigned char* p = nullptr;
  if (!p) SMCM_HIP(hipHostMalloc((void**)&p, kHostStageBytes, hipHostMallocDefault));
  return p;
}

#ifdef __HIP_DEVICE_COMPILE__
#define SMCM_DEV __device__ __forceinline__
#else
#define SMCM_DEV __device__ inline
#endif

SMCM_DEV int lane_id() { return (int)__lane_id(); }

// number of set bits of `mask` strictly below this lane
SMCM_DEV int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

SMCM_DEV uint64_t readlane_u64(uint64_t x, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

SMCM_DEV double readlane_f64(double x, int lane) {
  uint64_t b = __double_as_longlong(x);
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// A value the whole wave holds (read from LDS at a wave-uniform address, say) moved into scalar
// registers: the compiler cannot prove such a value uniform and would keep it in a VGPR per lane.
SMCM_DEV int32_t uniform_i32(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
SMCM_DEV int64_t uniform_i64(int64_t x) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(uint64_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
SMCM_DEV uint64_t uniform_u64(uint64_t x) { return (uint64_t)uniform_i64((int64_t)x); }

// The wave's index in its workgroup, in a scalar register: loops over wave-strided work then run
// on the scalar unit instead of as exec-masked vector loops.
SMCM_DEV int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave)); }

SMCM_DEV int64_t wave_sum_i64(int64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

SMCM_DEV int64_t wave_max_i64(int64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int64_t y = __shfl_xor(x, o);
    x = y > x ? y : x;
  }
  return x;
}

}  // namespace smcm
