// scan.hip - exclusive prefix sums over int64 (row pointers, product offsets).
//
// Three-phase tile scan: per-tile totals, a (recursive) scan of the totals, then a tile-local
// scan with the tile offset added.  Tiles of 2048 elements are staged through LDS so that
// global loads and stores are coalesced; thread-level sums run over 8 contiguous elements.
#include "kernels.h"

namespace smcm {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kTile = kScanThreads * kScanItems;

// block-wide exclusive scan of one value per thread; returns the exclusive prefix and the
// block total through *total
__device__ int64_t block_exclusive_scan(int64_t x, int64_t* total) {
  __shared__ int64_t wsum[kScanThreads / kWave];
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  int64_t incl = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int64_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  int64_t woff = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / kWave; ++w) {
    if (w < wid) woff += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return woff + incl - x;
}

__global__ void __launch_bounds__(kScanThreads) tile_sum_kernel(const int64_t* __restrict__ in, int64_t n,
                                                                int64_t* __restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * kTile;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t idx = base + (int64_t)i * kScanThreads + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  int64_t tot;
  (void)block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// out[i+1] = offset(tile) + inclusive prefix within tile; out[0] = 0 by tile 0
__global__ void __launch_bounds__(kScanThreads) tile_scan_kernel(const int64_t* __restrict__ in, int64_t n,
                                                                 const int64_t* __restrict__ offs,
                                                                 int64_t* __restrict__ out) {
  __shared__ int64_t buf[kTile];
  const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t idx = base + (int64_t)i * kScanThreads + threadIdx.x;
    buf[i * kScanThreads + threadIdx.x] = idx < n ? in[idx] : 0;
  }
  __syncthreads();
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    s += buf[threadIdx.x * kScanItems + i];
    v[i] = s;
  }
  int64_t tot;
  int64_t ex = block_exclusive_scan(s, &tot);
  const int64_t off = offs ? offs[blockIdx.x] : 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) buf[threadIdx.x * kScanItems + i] = v[i] + ex + off;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t idx = base + (int64_t)i * kScanThreads + threadIdx.x;
    if (idx < n) out[idx + 1] = buf[i * kScanThreads + threadIdx.x];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

__global__ void zero_first_kernel(int64_t* out) { out[0] = 0; }

}  // namespace

void scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, DevAlloc& al) {
  hipStream_t st = al.stream;
  if (n <= 0) {
    zero_first_kernel<<<1, 1, 0, st>>>(out);
    SMCM_LAUNCH_CHECK();
    return;
  }
  const int64_t ntiles = ceil_div(n, kTile);
  if (ntiles == 1) {
    tile_scan_kernel<<<1, kScanThreads, 0, st>>>(in, n, nullptr, out);
    SMCM_LAUNCH_CHECK();
    return;
  }
  DevBuf<int64_t> sums(al, ntiles), offs(al, ntiles + 1);
  tile_sum_kernel<<<(unsigned)ntiles, kScanThreads, 0, st>>>(in, n, sums.p);
  SMCM_LAUNCH_CHECK();
  scan_exclusive_i64(sums.p, offs.p, ntiles, al);
  tile_scan_kernel<<<(unsigned)ntiles, kScanThreads, 0, st>>>(in, n, offs.p, out);
  SMCM_LAUNCH_CHECK();
}

}  // namespace smcm
