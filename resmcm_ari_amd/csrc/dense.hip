// dense.hip - dense-accumulator tier for products with a small output space.
//
// Reference: the row kernel of SpGEMM (efficient/resmcm.cpp:586-615) visits every product of a
// row of A and marks its column; here, when rows(A) x cols(B) is small (a few million cells),
// the whole product is accumulated at once in an all-zero device array D[rows x cols] of uint64
// counts plus a presence bitmap (one bit per cell, rows padded to 64-bit words):
//   1. scatter: the entries of A are taken 64 at a time by a wave; each lane finds its entry's
//      row (cooperative search over A's row pointers) and B segment, a wave prefix sum lays the
//      chunk's products out flat, and lane l takes products l, l+64, ...: D[u, col] += a * b
//      (device atomics) and the presence bit of (u, col) is set.  Work is split by products, not
//      by rows, so a hub row of A (tens of thousands of entries) is spread over many waves
//      instead of serialising one workgroup;
//   2. count: a wave per (row, 16-word chunk of the bitmap) pops its words; one scan over all
//      chunks gives every chunk its output offset (and the row pointers of C);
//   3. compact: a wave per (row, chunk) writes its columns in order (lane = bit, 8 words of
//      loads in flight) with their counts, and returns every touched word of D and of the bitmap
//      to zero, so the array is all-zero again for the next product.
// Presence bits, not nonzero counts, define the structure: a count that wraps to 0 mod 2^64 is
// still an entry, as in the Boolean product.  One host synchronisation (the output size) per
// product, and no per-tier launches.
#include "kernels.h"

#include <algorithm>
#include <cstring>

namespace smcm {

namespace {

constexpr int kDThreads = 256;
constexpr int kDWaves = kDThreads / kWave;
constexpr int kDChunk = 16;  // presence words per count / compaction task

// row of entry e0: the u with ptr[u] <= e0 < ptr[u+1] (0 <= e0 < ptr[m]).  64 probes per step
// narrow [lo, hi) by 64x, so m rows take ceil(log64 m) dependent loads.
__device__ int32_t wave_row_of(const int64_t* __restrict__ ptr, int32_t m, int64_t e0) {
  const int lane = lane_id();
  int64_t lo = 0, hi = m;  // ptr[lo] <= e0 < ptr[hi]
  while (hi - lo > 1) {
    const int64_t step = (hi - lo + kWave - 1) / kWave;
    const int64_t idx = lo + (int64_t)lane * step;
    const bool le = idx < hi && ptr[idx] <= e0;  // a prefix of the lanes (ptr non-decreasing)
    const uint64_t mk = __ballot(le);            // bit 0 set: ptr[lo] <= e0
    const int j = 63 - __clzll((long long)mk);
    const int64_t nlo = lo + (int64_t)j * step;
    hi = nlo + step < hi ? nlo + step : hi;
    lo = nlo;
  }
  return (int32_t)lo;
}

// Sorts (key, value) pairs ascending by key across the wave (bitonic, xor shuffles), then gives
// the last lane of each run of equal keys the run's sum (mod 2^64).
__device__ void wave_sort_sum(int32_t& k, uint64_t& v) {
  const int lane = lane_id();
#pragma unroll
  for (int size = 2; size <= kWave; size <<= 1) {
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1) {
      const int32_t ok = __shfl_xor(k, j);
      const uint64_t ov = __shfl_xor(v, j);
      const bool asc = (lane & size) == 0;
      const bool lower = (lane & j) == 0;
      if ((lower == asc) ? (ok < k) : (ok > k)) {
        k = ok;
        v = ov;
      }
    }
  }
  // segmented inclusive sum: keys are sorted, so equal end points mean one run in between
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint64_t y = __shfl_up(v, o);
    const int32_t ky = __shfl_up(k, o);
    if (lane >= o && ky == k) v += y;
  }
}

// sum of one value per thread of the block (NWV waves), added to *out by one atomic: the same
// address updated by every wave would serialise
template <int NWV>
__device__ void block_add(unsigned long long x, unsigned long long* out) {
  __shared__ unsigned long long part[NWV];
  x = (unsigned long long)wave_sum_i64((int64_t)x);
  if (lane_id() == 0) part[wave_id()] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
#pragma unroll
    for (int w = 0; w < NWV; ++w) t += part[w];
    if (t) atomicAdd(out, t);
  }
}

constexpr int kLcombCols = 4096;  // products with at most this many columns combine hub rows in LDS

// LCOMB (products with at most kLcombCols columns): the THREADS / 64 waves of a block take
// consecutive 64-entry chunks together; when all of the block's products fall in one row (a hub
// row of A), they are summed in an LDS copy of that row and the block makes one device atomic per
// distinct column, instead of every wave adding to the row's top columns.
template <bool AV, bool BV, bool LCOMB, int THREADS>
__global__ void __launch_bounds__(THREADS) dense_scatter_kernel(
    const int64_t* __restrict__ aptr, const int32_t* __restrict__ acol, const uint64_t* __restrict__ aval, int32_t m,
    int64_t nnz_a, const int64_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
    const uint64_t* __restrict__ bval, int32_t ncols, int32_t nw, unsigned long long* __restrict__ D,
    unsigned long long* __restrict__ pres, unsigned long long* __restrict__ counters) {
  constexpr int NWV = THREADS / kWave;
  __shared__ unsigned long long lv[LCOMB ? kLcombCols : 1];
  __shared__ uint32_t lp[LCOMB ? kLcombCols / 32 : 1];
  __shared__ int32_t s_lo, s_hi;
  const int lane = lane_id(), wid = wave_id();
  if constexpr (LCOMB) {
    for (int i = threadIdx.x; i < kLcombCols; i += THREADS) lv[i] = 0ull;
    for (int i = threadIdx.x; i < kLcombCols / 32; i += THREADS) lp[i] = 0u;
    if (threadIdx.x == 0) {
      s_lo = INT32_MAX;
      s_hi = -1;
    }
    __syncthreads();
  }
  unsigned long long products = 0;
  // the block's waves take consecutive chunks; the loop is uniform over the block
  for (int64_t b0 = (int64_t)blockIdx.x * NWV * kWave; b0 < nnz_a; b0 += (int64_t)gridDim.x * NWV * kWave) {
    const int64_t e0 = b0 + (int64_t)wid * kWave;
    const int64_t e = e0 + lane;
    const bool valid = e < nnz_a;
    int32_t u = 0;
    int64_t bs = 0, n = 0;
    uint64_t a = 1;
    if (e0 < nnz_a) {  // wave-uniform
      // this lane's row: rows from the chunk's first row on, 64 row ends at a time
      const int32_t u0 = wave_row_of(aptr, m, e0);
      u = u0;
      bool open = valid;
      for (int64_t base = u0; __ballot(open); base += kWave) {
        const int64_t r = base + 1 + lane;
        const int64_t q = r <= m ? aptr[r] : INT64_MAX;  // end of row base + lane
        // number of row ends <= e among the 64 (sorted): the lane's row is base + that count
        int pos = 0;
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) {
          const int64_t qq = __shfl(q, pos + s - 1);
          if (open && qq <= e) pos += s;
        }
        const int64_t q63 = __shfl(q, kWave - 1);
        if (open && !(pos == kWave - 1 && q63 <= e)) {
          u = (int32_t)(base + pos);
          open = false;
        }
      }
      if (valid) {
        const int32_t k = acol[e];
        bs = bptr[k];
        n = bptr[k + 1] - bs;
        if (AV) a = aval[e];
      }
    }
    // exclusive prefix of the segment lengths over the wave
    int64_t incl = n;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int64_t y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    const int64_t off = incl - n;
    const int64_t total = __shfl(incl, kWave - 1);
    products += lane == 0 ? (unsigned long long)total : 0ull;
    bool in_lds = false;
    int32_t brow = 0;
    if constexpr (LCOMB) {
      // do all of the block's products lie in one row?
      const int64_t lo = wave_max_i64(valid && n > 0 ? -(int64_t)u : INT64_MIN);
      const int64_t hi = wave_max_i64(valid && n > 0 ? (int64_t)u : -1);
      if (lane == 0 && total > 0) {
        atomicMin(&s_lo, (int32_t)(-lo));
        atomicMax(&s_hi, (int32_t)hi);
      }
      __syncthreads();
      in_lds = s_lo == s_hi;
      brow = s_lo;
    }
    for (int64_t p0 = 0; p0 < total; p0 += kWave) {
      const int64_t p = p0 + lane;
      // owner: the last lane whose offset is <= p (zero-length lanes after it start beyond p)
      int j = 0;
#pragma unroll
      for (int s = 32; s >= 1; s >>= 1) {
        const int64_t oo = __shfl(off, j + s);
        if (oo <= p) j += s;
      }
      const int64_t oj = __shfl(off, j);
      const int64_t bj = __shfl(bs, j);
      const uint64_t aj = __shfl(a, j);
      const int32_t uj = __shfl(u, j);
      const bool act = p < total;  // lane 0 always is
      int32_t c = INT32_MAX;
      uint64_t v = 0;
      if (act) {
        const int64_t b = bj + (p - oj);
        c = bcol[b];
        v = BV ? aj * bval[b] : aj;
      }
      if (in_lds) {
        if (act) {
          atomicAdd(&lv[c], (unsigned long long)v);
          atomicOr(&lp[c >> 5], 1u << (c & 31));
        }
        continue;
      }
      // Products of one row (a hub row of A fills whole waves) often repeat a column: sorted
      // across the wave and summed per column, they take one atomic per distinct column.
      // Same-address device atomics serialise, and a hub row sends thousands to its top columns.
      bool emit = act;
      int32_t urow = uj;
      const int32_t u_first = __shfl(uj, 0);
      if (__ballot(act && uj != u_first) == 0) {
        wave_sort_sum(c, v);  // the lanes' pairs move: every one of them belongs to row u_first
        const int32_t cn = __shfl_down(c, 1);
        emit = c != INT32_MAX && (lane == kWave - 1 || cn != c);
        urow = u_first;
      }
      if (emit) {
        atomicAdd(&D[(int64_t)urow * ncols + c], (unsigned long long)v);
        // Bits are only ever set here, so a word read with the bit already set needs no atomic.
        // A stale read only costs a redundant atomic.
        unsigned long long* pw = &pres[(int64_t)urow * nw + (c >> 6)];
        const unsigned long long bit = 1ull << (c & 63);
        if (!(__hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit)) atomicOr(pw, bit);
      }
    }
    if constexpr (LCOMB) {
      __syncthreads();
      if (in_lds) {  // block-uniform: the block's sums of row brow leave, one atomic per column
        for (int32_t c = threadIdx.x; c < ncols; c += THREADS) {
          if ((lp[c >> 5] >> (c & 31)) & 1u) {
            atomicAdd(&D[(int64_t)brow * ncols + c], lv[c]);
            unsigned long long* pw = &pres[(int64_t)brow * nw + (c >> 6)];
            const unsigned long long bit = 1ull << (c & 63);
            if (!(__hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit)) atomicOr(pw, bit);
            lv[c] = 0ull;
          }
        }
        __syncthreads();  // every bit read before the words are cleared
        for (int32_t w = threadIdx.x; w < (ncols + 31) / 32; w += THREADS) lp[w] = 0u;
      }
      if (threadIdx.x == 0) {
        s_lo = INT32_MAX;
        s_hi = -1;
      }
      __syncthreads();
    }
  }
  block_add<NWV>(products, &counters[0]);
}

// popcount of every (row, kDChunk-word chunk) of the bitmap: cc[u * nch + c]
__global__ void __launch_bounds__(kDThreads) dense_count_kernel(int32_t m, int32_t nw, int32_t nch,
                                                                const unsigned long long* __restrict__ pres,
                                                                int64_t* __restrict__ cc) {
  const int lane = lane_id();
  const int64_t tasks = (int64_t)m * nch;
  for (int64_t t = (int64_t)blockIdx.x * kDWaves + wave_id(); t < tasks; t += (int64_t)gridDim.x * kDWaves) {
    const int64_t u = t / nch;
    const int32_t w = (int32_t)(t - u * nch) * kDChunk + lane;
    int64_t s = (lane < kDChunk && w < nw) ? __popcll(pres[u * nw + w]) : 0;
    s = wave_sum_i64(s);
    if (lane == 0) cc[t] = s;
  }
}

// Task t = (row u, chunk c of kDChunk presence words) writes its columns in order from off[t]
// (the exclusive prefix of the chunk counts): lane = bit, 8 words per step with their loads of D
// in flight together; D and the bitmap are cleared behind.  Chunk 0 of a row writes cptr[u].
// Short chunks keep a dense row's words spread over many waves.
__global__ void __launch_bounds__(kDThreads) dense_compact_kernel(
    int32_t m, int32_t ncols, int32_t nw, int32_t nch, unsigned long long* __restrict__ pres,
    unsigned long long* __restrict__ D, const int64_t* __restrict__ off, int64_t* __restrict__ cptr,
    int32_t* __restrict__ ccol, uint64_t* __restrict__ cval, unsigned long long* __restrict__ counters) {
  constexpr int kStep = 8;
  const int lane = lane_id();
  const int64_t tasks = (int64_t)m * nch;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counters[0] = 0;  // read back before this launch
    cptr[m] = off[tasks];
  }
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  for (int64_t t = (int64_t)blockIdx.x * kDWaves + wave_id(); t < tasks; t += (int64_t)gridDim.x * kDWaves) {
    const int64_t u = t / nch;
    const int32_t w0 = (int32_t)(t - u * nch) * kDChunk;
    const int64_t base = off[t];
    if (w0 == 0 && lane == 0) cptr[u] = base;
    const int32_t wn = nw - w0 < kDChunk ? nw - w0 : kDChunk;
    unsigned long long* prow = pres + u * nw + w0;
    unsigned long long* drow = D + u * ncols + (int64_t)w0 * kWave;
    const uint64_t x = lane < wn ? prow[lane] : 0ull;
    const int64_t cnt = __popcll(x);
    int64_t incl = cnt;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int64_t y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    const int64_t excl = incl - cnt;
    const uint64_t any = __ballot(x != 0);
    for (int i0 = 0; i0 < wn; i0 += kStep) {
      if (!((any >> i0) & 0xffull)) continue;
      uint64_t wd[kStep], v[kStep];
#pragma unroll
      for (int s = 0; s < kStep; ++s) {
        wd[s] = i0 + s < kWave ? readlane_u64(x, i0 + s) : 0ull;
        v[s] = ((wd[s] >> lane) & 1ull) ? drow[(int64_t)(i0 + s) * kWave + lane] : 0ull;
      }
#pragma unroll
      for (int s = 0; s < kStep; ++s) {
        if ((wd[s] >> lane) & 1ull) {
          const int64_t pos = base + (int64_t)readlane_u64((uint64_t)excl, i0 + s) + __popcll(wd[s] & below);
          ccol[pos] = w0 * kWave + (i0 + s) * kWave + lane;
          cval[pos] = v[s];
          drow[(int64_t)(i0 + s) * kWave + lane] = 0;
        }
      }
    }
    if (x) prow[lane] = 0;
  }
}

// The context's all-zero block (D, bitmap and two counters), grown on demand; a use that did not
// finish (an exception between scatter and compaction) leaves it marked dirty, and it is cleared
// again before the next use.
unsigned char* zero_block(DevAlloc& al, size_t bytes) {
  if (al.zero_block && al.zero_bytes < bytes) al.free_zero_block();
  if (!al.zero_block) {
    const size_t grow = std::max(bytes, (size_t)16 << 20);
    al.zero_block = al.raw_alloc(grow);
    al.zero_bytes = grow;
    al.zero_dirty = true;
  }
  if (al.zero_dirty) {
    fill_i64((int64_t*)al.zero_block, (int64_t)(al.zero_bytes / 8), 0, al.stream);
    al.zero_dirty = false;
  }
  return (unsigned char*)al.zero_block;
}

// holds the all-zero block for one product: an allocation that fails and releases the cache in
// the meantime leaves the block mapped
struct ZeroPin {
  DevAlloc& al;
  explicit ZeroPin(DevAlloc& a) : al(a) { al.zero_pins++; }
  ~ZeroPin() { al.zero_pins--; }
  ZeroPin(const ZeroPin&) = delete;
  ZeroPin& operator=(const ZeroPin&) = delete;
};

template <bool AV, bool BV>
void launch_scatter(const Csr& A, const Csr& B, int32_t nw, unsigned long long* D, unsigned long long* pres,
                    unsigned long long* counters, hipStream_t st) {
  const int64_t waves = ceil_div(A.nnz, kWave);
  if (B.ncols <= kLcombCols) {
    constexpr int T = 1024;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(waves, T / kWave), 512));
    dense_scatter_kernel<AV, BV, true, T><<<grid, T, 0, st>>>(A.ptr, A.col, A.val, A.nrows, A.nnz, B.ptr, B.col,
                                                              B.val, B.ncols, nw, D, pres, counters);
  } else {
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(waves, kDWaves), 2048));
    dense_scatter_kernel<AV, BV, false, kDThreads><<<grid, kDThreads, 0, st>>>(
        A.ptr, A.col, A.val, A.nrows, A.nnz, B.ptr, B.col, B.val, B.ncols, nw, D, pres, counters);
  }
  SMCM_LAUNCH_CHECK();
}

// products of A * B (the sum over A's entries of their B row lengths) into *out (zeroed by the
// caller; one atomic per block)
__global__ void __launch_bounds__(kDThreads) dense_products_kernel(const int32_t* __restrict__ acol, int64_t nnz_a,
                                                                   const int64_t* __restrict__ bptr,
                                                                   unsigned long long* __restrict__ out) {
  unsigned long long s = 0;
  for (int64_t e = (int64_t)blockIdx.x * kDThreads + threadIdx.x; e < nnz_a; e += (int64_t)gridDim.x * kDThreads) {
    const int32_t k = acol[e];
    s += (unsigned long long)(bptr[k + 1] - bptr[k]);
  }
  block_add<kDWaves>(s, out);
}

}  // namespace

bool spgemm_dense(const Csr& A, const Csr& B, Csr& C, const SpgemmOpts& o, SpgemmStats* stats, DevAlloc& al,
                  KernelTimer* tm) {
  const int64_t m = A.nrows, ncols = B.ncols;
  const int64_t cells = m * ncols;
  if (o.dense_cells <= 0 || cells > o.dense_cells || o.force_tier >= 0 || al.arena || m == 0 || ncols == 0)
    return false;
  hipStream_t st = al.stream;
  if (ncols > kLcombCols && o.dense_row_products > 0 && A.nnz > 0) {
    // no LDS combining over this many columns: hub rows would send their products to the same
    // columns by device atomics, which the row tiers' LDS windows do far faster
    DevBuf<unsigned long long> cnt(al, 1);
    fill_i64((int64_t*)cnt.p, 1, 0, st);
    dense_products_kernel<<<(unsigned)std::min<int64_t>(ceil_div(A.nnz, kDThreads), 1024), kDThreads, 0, st>>>(
        A.col, A.nnz, B.ptr, cnt.p);
    SMCM_LAUNCH_CHECK();
    unsigned char* hs = host_stage() + kHostStageDense;
    SMCM_HIP(hipMemcpyAsync(hs, cnt.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    SMCM_HIP(hipStreamSynchronize(st));
    unsigned long long products = 0;
    memcpy(&products, hs, sizeof(products));
    if ((double)products >= (double)o.dense_row_products * (double)m) return false;
  }
  const int32_t nw = (int32_t)ceil_div(ncols, kWave);
  const int32_t nch = (int32_t)ceil_div(nw, kDChunk);
  const int64_t tasks = m * nch;
  const size_t dbytes = (size_t)cells * 8, pbytes = (size_t)m * nw * 8;
  unsigned char* z = zero_block(al, dbytes + pbytes + 64);
  ZeroPin pin(al);  // D, pres and counters point into the block until the compaction is enqueued
  auto* D = (unsigned long long*)z;
  auto* pres = (unsigned long long*)(z + dbytes);
  auto* counters = (unsigned long long*)(z + dbytes + pbytes);
  DevBuf<int64_t> cc(al, tasks), off(al, tasks + 1);
  const unsigned tgrid = (unsigned)std::min<int64_t>(ceil_div(tasks, kDWaves), 65536);
  al.zero_dirty = true;
  int acc_tok = -1;
  {
    acc_tok = tm ? tm->start("spgemm.dense.accumulate", st) : -1;
    if (A.nnz > 0) {
      if (A.val && B.val) launch_scatter<true, true>(A, B, nw, D, pres, counters, st);
      else if (A.val) launch_scatter<true, false>(A, B, nw, D, pres, counters, st);
      else if (B.val) launch_scatter<false, true>(A, B, nw, D, pres, counters, st);
      else launch_scatter<false, false>(A, B, nw, D, pres, counters, st);
    }
    dense_count_kernel<<<tgrid, kDThreads, 0, st>>>((int32_t)m, nw, nch, pres, cc.p);
    SMCM_LAUNCH_CHECK();
    // algorithmic bytes are set below, once the product count is known
    if (tm) tm->stop(acc_tok, st, 0.0);
  }
  scan_exclusive_i64(cc.p, off.p, tasks, al);
  unsigned char* hs = host_stage() + kHostStageDense;
  SMCM_HIP(hipMemcpyAsync(hs, off.p + tasks, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SMCM_HIP(hipMemcpyAsync(hs + 8, counters, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  SMCM_HIP(hipStreamSynchronize(st));
  int64_t nnz = 0;
  unsigned long long products = 0;
  memcpy(&nnz, hs, sizeof(nnz));
  memcpy(&products, hs + 8, sizeof(products));
  if (tm)
    tm->set_bytes(acc_tok, 8.0 * (double)(m + 1) + (20.0 + (A.val ? 8 : 0)) * (double)A.nnz +
                               (4.0 + (B.val ? 8 : 0)) * (double)products);
  C.nnz = nnz;
  C.col = al.alloc_n<int32_t>(nnz);
  C.val = al.alloc_n<uint64_t>(nnz);
  {
    const int tok = tm ? tm->start("spgemm.dense.compact", st) : -1;
    dense_compact_kernel<<<tgrid, kDThreads, 0, st>>>((int32_t)m, (int32_t)ncols, nw, nch, pres, D, off.p, C.ptr,
                                                      C.col, C.val, counters);
    SMCM_LAUNCH_CHECK();
    if (tm) tm->stop(tok, st, 8.0 * (double)(m + 1) + 12.0 * (double)nnz);
  }
  al.zero_dirty = false;
  if (stats) {
    *stats = SpgemmStats{};
    stats->products = (int64_t)products;
    stats->nnz = nnz;
    stats->rows[3] = m;  // every row accumulated by the dense tier
    stats->windows = 1;
  }
  return true;
}

}  // namespace smcm
