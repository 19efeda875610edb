// spgemm.hip - row-wise Gustavson SpGEMM with uint64 path counts for gfx950.
//
// Reference: SpGEMM / row_wise_SpGEMM_based_on_bitmap (efficient/resmcm.cpp:586-647): for every
// row u of A, the union over k in row_u(A) of row_k(B).  Here every column also carries the
// number of paths, sum_k A[u,k] * B[k,v] mod 2^64, and rows are emitted sorted by column.
//
// Pipeline (all on one stream):
//   1. analysis: len[e] = |row_{A.col[e]}(B)| for every entry of A, S = exclusive scan(len);
//      ub(u) = S[A.ptr[u+1]] - S[A.ptr[u]] is the row's intermediate-product count and an
//      upper bound of its width.
//   2. binning: ub == 0 -> empty; ub <= s_max -> wave tier; ub <= m_max with a sparse estimate
//      -> sort tier; otherwise the bitmap tier (rows with ub > heavy_ub first).  The estimator's
//      predicted width r(u) * ncols is what sends a short-but-dense row to the bitmap tier.
//   3. symbolic pass per tier -> exact row widths; scan -> C.ptr
//   4. numeric pass per tier -> C.col (ascending), C.val
//
// Tiers:
//   wave   : one 64-lane wave per row; products expanded into lanes, bitonic sort across the
//            wave with xor shuffles, equal keys merged with a segmented scan.
//   sort   : one workgroup per row; products expanded into LDS, bitonic sort in LDS, merged.
//   bitmap : one workgroup per row (persistent grid pulling rows from a queue); a column bitmap
//            of the row lives in LDS (one window of up to NW columns at a time).  Symbolic =
//            popcount.  Numeric = bitmap, then a prefix count per 64-bit word gives every
//            column its output slot, so columns come out sorted without a sort; one wave writes
//            one 64-bit word (lane l <-> bit l) so output stores are coalesced.  Values are
//            accumulated in LDS when the row fits (VCAP slots), otherwise with global atomics
//            into the row's (pre-zeroed) output slots.
// Products of an A entry are visited by the thread owning the entry when its B row is short,
// or by a whole wave (coalesced B loads) when it is long.
#include <map>
#include <mutex>
#include <tuple>

#include "kernels.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <vector>

namespace smcm {

namespace {

constexpr int kThreads = 256;
constexpr int kShortRow = 16;  // B rows up to this length are walked by one thread
// byte offsets of the readbacks in the thread's pinned host stage
constexpr size_t kStageBin = 0, kStageTierNnz = 320, kStageNnz = 384, kStageSplit = 448;
static_assert(kStageSplit + 112 <= kHostStageDense, "host stage sections overlap");
// byte offsets in a product's device counter block
// (sums: per tier A entries and products, per tier output entries, per tier largest product
// count, per tier largest row of A: 5 * kNumBins values; then the largest row sum of A's values
// and the largest value of B)
constexpr size_t kMetaCounts = 0, kMetaQueue = 32, kMetaSums = 64, kMetaSplit = 320, kMetaBytes = 448;
static_assert(kMetaSums + (5 * 6 + 2) * 8 <= kMetaSplit, "tier sums overlap the split counters");
constexpr int64_t kMaxValueScan = (int64_t)1 << 25;  // operands scanned for their largest value

// ------------------------------------------------------------------------------------------
// analysis
__global__ void entry_len_kernel(const int32_t* __restrict__ acol, int64_t nnz, const int64_t* __restrict__ bptr,
                                 int64_t* __restrict__ len) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    int32_t k = acol[e];
    len[e] = bptr[k + 1] - bptr[k];
  }
}

enum Tier { kEmpty = 0, kWaveTier = 1, kSortTier = 2, kBitmapTier = 3, kHeavyTier = 4, kCopyTier = 5, kNumBins = 6 };


constexpr int kBinRowsPerThread = 4;  // a block bins kBinRowsPerThread * kThreads rows

// Tier of every row; rows are appended to their tier's list.  Counts and sums are aggregated per
// wave, then per block in LDS, so each block makes one global atomic per tier and counter: the
// same few counters updated by every wave would serialise the launch.
__global__ void __launch_bounds__(kThreads) bin_rows_kernel(const int64_t* __restrict__ aptr, int32_t r0, int32_t r1,
                                                            int32_t m, const int64_t* __restrict__ S,
                                                            const double* __restrict__ est_r, int est_stride,
                                                            double ncols, int s_max, int m_max, int dense_div,
                                                            int64_t heavy_ub, int force_tier, int wave_bitmap_max_ub,
                                                            int copy_ok, int32_t* __restrict__ bins,
                                                            int32_t* __restrict__ counts,
                                                            unsigned long long* __restrict__ sums) {
  __shared__ int32_t lcnt[kNumBins], lbase[kNumBins];
  __shared__ unsigned long long lsum[2 * kNumBins], lmax[2 * kNumBins];  // largest ub, largest d
  if (threadIdx.x < kNumBins) lcnt[threadIdx.x] = 0;
  if (threadIdx.x < 2 * kNumBins) lmax[threadIdx.x] = 0;
  if (threadIdx.x < 2 * kNumBins) lsum[threadIdx.x] = 0;
  __syncthreads();
  const int lane = lane_id();
  int tiers[kBinRowsPerThread], slot[kBinRowsPerThread];
#pragma unroll
  for (int i = 0; i < kBinRowsPerThread; ++i) {
    const int64_t uu = (int64_t)r0 + ((int64_t)blockIdx.x * kBinRowsPerThread + i) * kThreads + threadIdx.x;
    const int32_t u = (int32_t)(uu < r1 ? uu : r1);
    int tier = -1;
    int64_t ub = 0, d = 0;
    if (uu < r1) {
      ub = S[aptr[u + 1]] - S[aptr[u]];
      d = aptr[u + 1] - aptr[u];
      if (ub == 0) {
        tier = kEmpty;
      } else if (copy_ok && d == 1 && ub > s_max) {
        tier = kCopyTier;  // one entry: the row is a scaled copy of one row of B
      } else if (force_tier >= 0) {
        tier = force_tier == 0 ? kWaveTier : force_tier == 1 ? kSortTier : kBitmapTier;
        if (tier == kWaveTier && ub > 64) tier = kSortTier;
        if (tier == kSortTier && ub > (wave_bitmap_max_ub ? wave_bitmap_max_ub : m_max)) tier = kBitmapTier;
      } else if (ub <= s_max) {
        tier = kWaveTier;
      } else if (wave_bitmap_max_ub) {
        tier = ub <= wave_bitmap_max_ub ? kSortTier : kBitmapTier;
      } else {
        bool dense = false;
        if (est_r) dense = est_r[(size_t)u * est_stride] * ncols * (double)dense_div >= ncols;
        tier = (ub <= m_max && !dense) ? kSortTier : kBitmapTier;
      }
      if (tier == kBitmapTier && ub > heavy_ub) tier = kHeavyTier;
    }
    tiers[i] = tier;
    slot[i] = 0;
    // wave-aggregated: a slot in the block's part of the tier list; per-tier sums of A entries
    // and products feed the algorithmic-byte accounting
    for (int t = 0; t < kNumBins; ++t) {
      const uint64_t mask = __ballot(tier == t);
      if (!mask) continue;
      const int leader = __ffsll((long long)mask) - 1;
      int32_t base = 0;
      const int64_t sd = wave_sum_i64(tier == t ? d : 0);
      const int64_t su = wave_sum_i64(tier == t ? ub : 0);
      const int64_t mu = wave_max_i64(tier == t ? ub : 0);
      const int64_t md = wave_max_i64(tier == t ? d : 0);
      if (lane == leader) {
        base = atomicAdd(&lcnt[t], (int32_t)__popcll(mask));
        atomicAdd(&lsum[2 * t], (unsigned long long)sd);
        atomicAdd(&lsum[2 * t + 1], (unsigned long long)su);
        atomicMax(&lmax[t], (unsigned long long)mu);
        atomicMax(&lmax[kNumBins + t], (unsigned long long)md);
      }
      base = __shfl(base, leader);
      if (tier == t) slot[i] = base + lanes_below(mask);
    }
  }
  __syncthreads();
  if (threadIdx.x < kNumBins) {
    const int32_t c = lcnt[threadIdx.x];
    lbase[threadIdx.x] = c ? atomicAdd(&counts[threadIdx.x], c) : 0;
  }
  if (threadIdx.x < 2 * kNumBins && lsum[threadIdx.x]) atomicAdd(&sums[threadIdx.x], lsum[threadIdx.x]);
  if (threadIdx.x < 2 * kNumBins && lmax[threadIdx.x])
    atomicMax(&sums[3 * kNumBins + threadIdx.x], lmax[threadIdx.x]);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kBinRowsPerThread; ++i) {
    const int t = tiers[i];
    if (t > kEmpty) {  // empty rows are only counted (t is -1 past r1)
      const int64_t u = (int64_t)r0 + ((int64_t)blockIdx.x * kBinRowsPerThread + i) * kThreads + threadIdx.x;
      bins[(size_t)t * m + lbase[t] + slot[i]] = (int32_t)u;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Symbolic pass over a right operand kept as a bitmap (B with few rows: its rows x column words
// fit a few MB).  Row u's width is the popcount of the OR of the bitmap rows of its A entries;
// a wave takes a row, lanes take 64-bit words (coalesced loads of each B row's words).  Work is
// nnz(A row) x words per row instead of the row's products, which the caller checks is cheaper.
__global__ void __launch_bounds__(kThreads) bbits_kernel(const int64_t* __restrict__ bptr,
                                                         const int32_t* __restrict__ bcol, int32_t nrows, int32_t nw,
                                                         unsigned long long* __restrict__ bits) {
  // bits must be zero; one wave per row of B, a lane per entry.  The row is sorted, so lanes
  // sharing a word are neighbours: their bits are ORed across the run first and only the run's
  // last lane makes the atomic (a run cut by the end of a step is finished by the next step's
  // atomic on the same word).
  const int lane = lane_id();
  for (int64_t r = (int64_t)blockIdx.x * (kThreads / kWave) + wave_id(); r < nrows;
       r += (int64_t)gridDim.x * (kThreads / kWave)) {
    const int64_t s0 = bptr[r], s1 = bptr[r + 1];
    for (int64_t g0 = s0; g0 < s1; g0 += 8 * kWave) {
      int32_t cs[8];  // eight steps' columns in flight per lane
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t e = g0 + q * kWave + lane;
        cs[q] = e < s1 ? bcol[e] : -1;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (g0 + q * kWave >= s1) break;  // wave-uniform
        const bool valid = cs[q] >= 0;
        const int32_t wd = valid ? cs[q] >> 6 : INT32_MAX;
        unsigned long long b = valid ? 1ull << (cs[q] & 63) : 0ull;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
          const unsigned long long y = __shfl_up(b, o);
          const int32_t wy = __shfl_up(wd, o);
          if (lane >= o && wy == wd) b |= y;
        }
        const int32_t wn = __shfl_down(wd, 1);
        if (valid && (lane == kWave - 1 || wn != wd)) atomicOr(&bits[r * nw + wd], b);
      }
    }
  }
}

constexpr int kBbitsLight = 8;  // rows of A with at most this many entries: a wave per row
constexpr int64_t kBbitsMaxEntries = 2048;  // longest row of A the bitmap symbolic pass takes

// adds a lane's word count c (word w of the row) to the row's per-window counts: windows of
// 2^wsh words; steps of 64 words starting at w0 (wcnt zeroed by the caller, atomics: several
// waves may add to one window)
__device__ __forceinline__ void bbits_window_add(int32_t* __restrict__ wcnt, int64_t u, int32_t nwin, int wsh,
                                                 int32_t w0, int32_t nw, int32_t w, int32_t c, int32_t sc) {
  const int lane = lane_id();
  if (wsh >= 6) {  // the step's 64 words lie in one window
    if (lane == 0 && sc) atomicAdd(&wcnt[u * nwin + (w0 >> wsh)], sc);
  } else {  // windows of fewer than 64 words lie inside the step
    const int32_t first = w0 >> wsh, last = min(w0 + kWave - 1, nw - 1) >> wsh;
    for (int32_t win = first; win <= last; ++win) {
      const int32_t wc = (int32_t)wave_sum_i64((w < nw && (w >> wsh) == win) ? c : 0);
      if (lane == 0 && wc) atomicAdd(&wcnt[u * nwin + win], wc);
    }
  }
}

// Light rows (at most kBbitsLight entries): a wave per row, the row's bitmap rows read eight at
// a time, widths into row_nnz and per-window counts into wcnt (windows of 2^log2w columns,
// log2w >= 6) when set.  Heavier rows are appended to heavy[] (count in *nheavy, zeroed).
// the non-empty rows of a product: the binning's tier lists, concatenated by index
struct RowLists {
  const int32_t* rows[kNumBins];
  int32_t start[kNumBins + 1];  // exclusive prefix of the list lengths
};

__global__ void __launch_bounds__(kThreads) bbits_symbolic_kernel(
    const int64_t* __restrict__ aptr, const int32_t* __restrict__ acol, const unsigned long long* __restrict__ bits,
    int32_t nw, RowLists lists, int log2w, int32_t nwin, int64_t* __restrict__ row_nnz, int32_t* __restrict__ wcnt,
    int32_t* __restrict__ heavy, int32_t* __restrict__ nheavy) {
  const int lane = lane_id();
  const int wsh = log2w - 6;  // log2 of the words per window
  const int32_t n = lists.start[kNumBins];
  for (int64_t i = (int64_t)blockIdx.x * (kThreads / kWave) + wave_id(); i < n;
       i += (int64_t)gridDim.x * (kThreads / kWave)) {
    int t = 0;
    while (i >= lists.start[t + 1]) ++t;  // wave-uniform, at most kNumBins steps
    const int64_t u = lists.rows[t][i - lists.start[t]];
    const int64_t rs = uniform_i64(aptr[u]), re = uniform_i64(aptr[u + 1]);
    if (re - rs > kBbitsLight) {
      if (lane == 0) heavy[atomicAdd(nheavy, 1)] = (int32_t)u;
      continue;
    }
    int32_t k[kBbitsLight];
#pragma unroll
    for (int j = 0; j < kBbitsLight; ++j) k[j] = rs + j < re ? acol[rs + j] : -1;
    int64_t total = 0;
    int32_t cur_win = -1, cur_cnt = 0;  // windows of 64+ words: the open window's count (lane 0)
    // four 64-word steps at a time: every load of their words in flight together
    for (int32_t g0 = 0; g0 < nw; g0 += 4 * kWave) {
      unsigned long long x[4] = {0ull, 0ull, 0ull, 0ull};
#pragma unroll
      for (int j = 0; j < kBbitsLight; ++j) {
        if (k[j] < 0) continue;  // wave-uniform
        const unsigned long long* brow = bits + (int64_t)k[j] * nw;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int32_t w = g0 + q * kWave + lane;
          if (w < nw) x[q] |= brow[w];
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int32_t w0 = g0 + q * kWave;
        if (w0 >= nw) break;  // wave-uniform
        const int32_t w = w0 + lane;
        const int32_t c = __popcll(x[q]);
        const int32_t sc = (int32_t)wave_sum_i64(c);
        total += sc;
        if (wcnt) {
          if (wsh >= 6) {
            // this wave owns the row's counts: plain stores of whole windows, no atomics
            // (their completion would hold up the next loads' waits)
            const int32_t win = w0 >> wsh;
            if (win != cur_win) {
              if (lane == 0 && cur_win >= 0 && cur_cnt) wcnt[u * nwin + cur_win] = cur_cnt;
              cur_win = win;
              cur_cnt = 0;
            }
            cur_cnt += sc;
          } else {
            bbits_window_add(wcnt, u, nwin, wsh, w0, nw, w, c, sc);
          }
        }
      }
    }
    if (lane == 0) {
      if (wcnt && cur_win >= 0 && cur_cnt) wcnt[u * nwin + cur_win] = cur_cnt;
      row_nnz[u] = total;
    }
  }
}

// Heavy rows: a workgroup per (row, 64-word chunk); its waves split the row's entries (eight
// bitmap rows in flight per lane), OR their words in LDS, and wave 0 counts the chunk into the
// row's width and window counts (row_nnz zeroed by the caller: chunks add).
__global__ void __launch_bounds__(kThreads) bbits_heavy_kernel(
    const int64_t* __restrict__ aptr, const int32_t* __restrict__ acol, const unsigned long long* __restrict__ bits,
    int32_t nw, int log2w, int32_t nwin, const int32_t* __restrict__ heavy, const int32_t* __restrict__ nheavy,
    int64_t* __restrict__ row_nnz, int32_t* __restrict__ wcnt) {
  constexpr int NWV = kThreads / kWave;
  __shared__ unsigned long long acc[kWave];
  const int lane = lane_id(), wid = wave_id();
  const int32_t nch = (nw + kWave - 1) / kWave;
  const int64_t tasks = (int64_t)(*nheavy) * nch;
  for (int64_t t = blockIdx.x; t < tasks; t += gridDim.x) {
    const int64_t u = heavy[t / nch];
    const int32_t w0 = (int32_t)(t % nch) * kWave;
    const int32_t w = w0 + lane;
    const int64_t rs = aptr[u], re = aptr[u + 1];
    if (wid == 0) acc[lane] = 0ull;
    __syncthreads();
    unsigned long long x = 0;
    if (w < nw) {
      int64_t e = rs + wid;
      for (; e + 7 * NWV < re; e += 8 * NWV) {
        unsigned long long y[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) y[j] = bits[(int64_t)acol[e + j * NWV] * nw + w];
#pragma unroll
        for (int j = 0; j < 8; ++j) x |= y[j];
      }
      for (; e < re; e += NWV) x |= bits[(int64_t)acol[e] * nw + w];
    }
    if (x) atomicOr(&acc[lane], x);
    __syncthreads();
    if (wid == 0) {
      const int32_t c = __popcll(acc[lane]);
      const int32_t sc = (int32_t)wave_sum_i64(c);
      if (lane == 0 && sc) atomicAdd((unsigned long long*)&row_nnz[u], (unsigned long long)sc);
      if (wcnt) bbits_window_add(wcnt, u, nwin, log2w - 6, w0, nw, w, c, sc);
    }
    __syncthreads();
  }
}

// largest of n uint64 values into *out (atomicMax; *out zeroed by the caller)
__global__ void __launch_bounds__(kThreads) max_u64_kernel(const uint64_t* __restrict__ v, int64_t n,
                                                           unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[kThreads / kWave];
  unsigned long long mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    mx = v[i] > mx ? v[i] : mx;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(mx, o);
    mx = y > mx ? y : mx;
  }
  if (lane_id() == 0) part[threadIdx.x / kWave] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 1; w < kThreads / kWave; ++w) mx = part[w] > mx ? part[w] : mx;
    if (mx) atomicMax(out, mx);
  }
}

// largest row sum of A's values into *out (a wave per row; a sum that overflows saturates)
__global__ void __launch_bounds__(kThreads) row_sum_max_kernel(const int64_t* __restrict__ aptr,
                                                               const uint64_t* __restrict__ aval, int32_t m,
                                                               unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[kThreads / kWave];
  const int lane = lane_id();
  unsigned long long mx = 0;
  for (int64_t u = (int64_t)blockIdx.x * (kThreads / kWave) + wave_id(); u < m;
       u += (int64_t)gridDim.x * (kThreads / kWave)) {
    const int64_t rs = aptr[u], re = aptr[u + 1];
    unsigned long long sum = 0;
    bool sat = false;
    for (int64_t e = rs + lane; e < re; e += kWave) {
      const unsigned long long y = sum + aval[e];
      sat |= y < sum;
      sum = y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long y = __shfl_xor(sum, o);
      const unsigned long long z = sum + y;
      sat |= z < sum;
      sum = z;
    }
    sat = __ballot(sat) != 0;
    const unsigned long long r = sat ? ~0ull : sum;
    mx = r > mx ? r : mx;
  }
  if (lane == 0) part[wave_id()] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / kWave; ++w) mx = part[w] > mx ? part[w] : mx;
    if (mx) atomicMax(out, mx);
  }
}

// largest value of every row of B into rmax[k] (1 for every row when B has no values; 0 for an
// empty row): a wave per row
__global__ void __launch_bounds__(kThreads) row_max_kernel(const int64_t* __restrict__ bptr,
                                                           const uint64_t* __restrict__ bval, int32_t nrows,
                                                           unsigned long long* __restrict__ rmax) {
  const int lane = lane_id();
  for (int64_t k = (int64_t)blockIdx.x * (kThreads / kWave) + wave_id(); k < nrows;
       k += (int64_t)gridDim.x * (kThreads / kWave)) {
    const int64_t rs = bptr[k], re = bptr[k + 1];
    unsigned long long mx = 0;
    if (!bval) {
      mx = re > rs ? 1ull : 0ull;
    } else {
      for (int64_t e = rs + lane; e < re; e += kWave) mx = bval[e] > mx ? bval[e] : mx;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(mx, o);
        mx = y > mx ? y : mx;
      }
    }
    if (lane == 0) rmax[k] = mx;
  }
}

// Bound of every count in row u of A * B: sum over the row's entries of A[u,e] (1 without values)
// times the largest value of B's row k_e (saturating).  big[u] = 1 when the bound reaches 2^32.
// stats[0] counts such rows, stats[1] is the largest product count among them and stats[2] their
// product total (one atomic each per block).  A wave per row.
__global__ void __launch_bounds__(kThreads) row_bound_kernel(const int64_t* __restrict__ aptr,
                                                             const int32_t* __restrict__ acol,
                                                             const uint64_t* __restrict__ aval, int32_t m,
                                                             const int64_t* __restrict__ bptr,
                                                             const unsigned long long* __restrict__ rmax,
                                                             uint8_t* __restrict__ big,
                                                             unsigned long long* __restrict__ stats) {
  __shared__ unsigned long long part[3][kThreads / kWave];
  const int lane = lane_id();
  unsigned long long mine = 0, pmax = 0, psum = 0;
  for (int64_t u = (int64_t)blockIdx.x * (kThreads / kWave) + wave_id(); u < m;
       u += (int64_t)gridDim.x * (kThreads / kWave)) {
    const int64_t rs = aptr[u], re = aptr[u + 1];
    unsigned long long sum = 0, prods = 0;
    bool sat = false;
    for (int64_t e = rs + lane; e < re; e += kWave) {
      const int32_t k = acol[e];
      const unsigned long long a = aval ? aval[e] : 1ull, r = rmax[k];
      const unsigned long long t = a * r;
      sat |= __umul64hi(a, r) != 0;
      const unsigned long long y = sum + t;
      sat |= y < sum;
      sum = y;
      prods += (unsigned long long)(bptr[k + 1] - bptr[k]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long y = __shfl_xor(sum, o);
      const unsigned long long z = sum + y;
      sat |= z < sum;
      sum = z;
      prods += __shfl_xor(prods, o);
    }
    const bool b = __ballot(sat) != 0 || sum >= (1ull << 32);
    if (lane == 0) {
      big[u] = b ? 1 : 0;
      if (b) {
        mine += 1;
        pmax = prods > pmax ? prods : pmax;
        psum += prods;
      }
    }
  }
  if (lane == 0) {
    part[0][wave_id()] = mine;
    part[1][wave_id()] = pmax;
    part[2][wave_id()] = psum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0, mx = 0, ps = 0;
    for (int w = 0; w < kThreads / kWave; ++w) {
      t += part[0][w];
      mx = part[1][w] > mx ? part[1][w] : mx;
      ps += part[2][w];
    }
    if (t) {
      atomicAdd(&stats[0], t);
      atomicMax(&stats[1], mx);
      atomicAdd(&stats[2], ps);
    }
  }
}

// splits a row list by the big[] flags: rows with small bounds to small_out, the others to
// big_out (counts[0], counts[1], zeroed by the caller; order within each list not kept)
__global__ void __launch_bounds__(kThreads) split_big_kernel(const int32_t* __restrict__ rows, int32_t n,
                                                             const uint8_t* __restrict__ big,
                                                             int32_t* __restrict__ small_out,
                                                             int32_t* __restrict__ big_out,
                                                             int32_t* __restrict__ counts) {
  const int lane = lane_id();
  for (int64_t i0 = ((int64_t)blockIdx.x * (kThreads / kWave) + wave_id()) * kWave; i0 < n;
       i0 += (int64_t)gridDim.x * kThreads) {
    const int64_t i = i0 + lane;
    const bool valid = i < n;
    const int32_t u = valid ? rows[i] : 0;
    const bool b = valid && big[u];
    const uint64_t mb = __ballot(b), ms = __ballot(valid && !b);
    int32_t bb = 0, bs = 0;
    if (lane == 0) {
      if (mb) bb = atomicAdd(&counts[1], (int32_t)__popcll(mb));
      if (ms) bs = atomicAdd(&counts[0], (int32_t)__popcll(ms));
    }
    bb = __shfl(bb, 0);
    bs = __shfl(bs, 0);
    if (b) big_out[bb + lanes_below(mb)] = u;
    else if (valid) small_out[bs + lanes_below(ms)] = u;
  }
}

// sum of symbolic row widths per tier; the empty tier has no row list (its rows are only
// counted), so blockIdx.y covers the tiers after it
__global__ void __launch_bounds__(kThreads) tier_nnz_kernel(const int32_t* __restrict__ bins, int32_t m,
                                                            const int32_t* __restrict__ counts,
                                                            const int64_t* __restrict__ row_nnz,
                                                            unsigned long long* __restrict__ out) {
  const int t = kEmpty + 1 + (int)blockIdx.y;
  const int32_t n = counts[t];
  int64_t s = 0;
  for (int32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    s += row_nnz[bins[(size_t)t * m + i]];
  s = wave_sum_i64(s);
  if (lane_id() == 0 && s) atomicAdd(&out[t], (unsigned long long)s);
}

// rowptr of a block relative to its first row
__global__ void rebase_ptr_kernel(const int64_t* __restrict__ ptr, int32_t n, int64_t* __restrict__ out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) out[i] = ptr[i] - ptr[0];
}

// ------------------------------------------------------------------------------------------
// shared product traversal for workgroup tiers: f(col, val, pos) for every product of row
// [rs, re) of A; pos = product index within the row (used by the sort tier).
template <bool AV, bool BV, bool NEED_POS, class F>
__device__ __forceinline__ void for_each_product(const int64_t rs, const int64_t re, const int32_t* __restrict__ acol,
                                                 const uint64_t* __restrict__ aval,
                                                 const int64_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
                                                 const uint64_t* __restrict__ bval, const int64_t* __restrict__ S,
                                                 int32_t* q, int* qn, F&& f) {
  const int nthreads = blockDim.x;
  const int nwaves = nthreads / kWave;
  const int wid = threadIdx.x / kWave, lane = lane_id();
  const int64_t S0 = NEED_POS ? S[rs] : 0;
  for (int64_t base = rs; base < re; base += nthreads) {
    const int64_t e = base + threadIdx.x;
    if (e < re) {
      const int32_t k = acol[e];
      const int64_t b0 = bptr[k], b1 = bptr[k + 1];
      if (b1 - b0 > kShortRow) {
        int slot = atomicAdd(qn, 1);
        q[slot] = (int32_t)(e - base);
      } else if (b1 > b0) {
        const uint64_t a = AV ? aval[e] : 1ull;
        const int64_t p0 = NEED_POS ? S[e] - S0 : 0;
        for (int64_t b = b0; b < b1; ++b) f(bcol[b], BV ? a * bval[b] : a, p0 + (b - b0));
      }
    }
    __syncthreads();
    const int n = *qn;
    for (int qi = wid; qi < n; qi += nwaves) {
      const int64_t ee = base + q[qi];
      const int32_t k = acol[ee];
      const int64_t b0 = bptr[k], b1 = bptr[k + 1];
      const uint64_t a = AV ? aval[ee] : 1ull;
      const int64_t p0 = NEED_POS ? S[ee] - S0 : 0;
      // four loads in flight per lane before their products are applied
      int64_t b = b0 + lane;
      for (; b + 3 * kWave < b1; b += 4 * kWave) {
        int32_t c[4];
        uint64_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          c[u] = bcol[b + u * kWave];
          v[u] = BV ? bval[b + u * kWave] : 1ull;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) f(c[u], BV ? a * v[u] : a, p0 + (b + u * kWave - b0));
      }
      for (; b < b1; b += kWave) f(bcol[b], BV ? a * bval[b] : a, p0 + (b - b0));
    }
    __syncthreads();
    if (threadIdx.x == 0) *qn = 0;
    __syncthreads();
  }
}

__device__ __forceinline__ int64_t block_sum_i64(int64_t x, int64_t* red) {
  x = wave_sum_i64(x);
  const int wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  __syncthreads();
  if (lane_id() == 0) red[wid] = x;
  __syncthreads();
  int64_t s = 0;
  for (int w = 0; w < nw; ++w) s += red[w];
  __syncthreads();
  return s;
}

// block exclusive scan of one int32 per thread
__device__ __forceinline__ int32_t block_exscan_i32(int32_t x, int32_t* red, int32_t* total) {
  const int lane = lane_id(), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  int32_t incl = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  __syncthreads();
  if (lane == kWave - 1) red[wid] = incl;
  __syncthreads();
  int32_t off = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    if (w < wid) off += red[w];
    tot += red[w];
  }
  __syncthreads();
  *total = tot;
  return off + incl - x;
}

// The traversal's scan: one barrier.  Its caller guarantees that every wave passed a barrier
// after the previous reads of red and passes another before the next write to it (the traversal
// round's own barriers do both).
__device__ __forceinline__ int64_t block_exscan_i64_once(int64_t x, int64_t* red, int64_t* total) {
  const int lane = lane_id(), wid = wave_id(), nw = blockDim.x / kWave;
  int64_t incl = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int64_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == kWave - 1) red[wid] = incl;
  __syncthreads();
  int64_t off = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    const int64_t r = uniform_i64(red[w]);
    if (w < wid) off += r;
    tot += r;
  }
  *total = tot;
  return off + incl - x;
}


// ------------------------------------------------------------------------------------------
// Balanced product traversal for the workgroup-per-row tiers.  For one row of A, every entry e
// (B row k) contributes a segment [off, off + cnt) of row k of B, given by seg().  Segments are
// classed by length and compacted (one packed block scan):
//   cnt <= kTiny        walked by the thread that owns the entry;
//   cnt <= kQuarter     a 16-lane quarter of a wave per segment, four segments per wave step;
//   longer              cut into pieces of 64·E products taken round-robin by the waves, STEP
//                       pieces per wave step, lanes reading consecutive elements (E·STEP loads
//                       in flight per lane).  The numeric window kernel uses E = 1, STEP = 8
//                       (short pieces fill the lanes) and four quarter steps in flight; the
//                       bitmap tier E = 4, STEP = 2 (fewer pieces to look up for its cheaper
//                       per-product work).  More steps in flight spill registers.
// All lanes of a step read consecutive elements of one B row (coalesced), a single very long B
// row (a hub) is shared by the whole workgroup, and there is no per-product search: a piece's
// segment is found once per piece (wave-uniform LDS reads).
constexpr int kChunk = 512;  // A entries per traversal round
constexpr int64_t kSegMaxBytes = (int64_t)1 << 30;  // largest segment table of B per product
constexpr int kTiny = 4;  // segments of up to this many elements are walked by their own thread
constexpr int kCompactBatch = 4;  // 64-column words per step of the window kernel's compaction
constexpr int kQuarter = 16;

constexpr int64_t kStaticLds = 16 * 1024;  // static LDS of the bitmap and window kernels (upper bound)
constexpr int64_t kWindowCountBytes = (int64_t)1 << 30;  // cap of the per-(row, window) count table

// Resident workgroups per CU of a kernel at a block size and dynamic LDS size (cached: the
// runtime query is made once per combination).
int occupancy(const void* kernel, int threads, size_t lds) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, int, size_t>, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_tuple(kernel, threads, lds);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int nb = 0;
  SMCM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, lds));
  nb = std::max(nb, 1);
  cache.emplace(key, nb);
  return nb;
}

// persistent grid of the window kernels: every CU filled, or, when the options reserve part of
// the device for other streams, one workgroup per CU of the share (the dispatcher spreads
// workgroups over all CUs, so several per CU would reach into the reserved part)
unsigned persistent_grid(const SpgemmOpts& o, int per_cu) {
  const bool shared = o.device_cus > 0 && o.cus < o.device_cus;
  return (unsigned)std::max<int64_t>(1, (int64_t)o.cus * (shared ? 1 : per_cu));
}


struct TravLds {
  int64_t start[kChunk];  // first B index of a compacted segment (quarter segments, then long)
  uint64_t av[kChunk];    // A value of its entry
  int32_t len[kChunk];    // its length
  int32_t Q[kChunk + 1];  // exclusive prefix of piece counts over the long segments
  int64_t red[1024 / kWave];
};

// One traversal round: every participating thread brings at most one segment (st, cnt) of B
// with the A value a of its entry (cnt = 0: none).  Ends with a barrier.
// COOP: f(valid, column, value) is called by every lane of the wave together (valid = the lane
// holds a product), so that it can combine lanes (the symbolic pass merges bits of one word);
// otherwise f(column, value) is called by the lanes that hold a product.
template <bool BV, int E = 1, int STEP = 8, int QH = 2, bool COOP = false, class F>
__device__ __forceinline__ void trav_round(int64_t st, int64_t cnt, uint64_t a, const int32_t* __restrict__ bcol,
                                           const uint64_t* __restrict__ bval, TravLds& L, F&& f) {
  auto call = [&](bool valid, int32_t c, uint64_t v) {
    if constexpr (COOP) f(valid, c, v);
    else if (valid) f(c, v);
  };
  constexpr int kPieceLen = E * kWave;
  const int nwaves = blockDim.x / kWave;
  const int wid = threadIdx.x / kWave, lane = lane_id();
  int64_t packed = 0;  // pieces << 32 | long << 16 | quarter
  // tiny segments: the owning thread's loads are issued before the scan's barriers and
  // accumulated after them
  int32_t tc[kTiny];
  uint64_t tv[kTiny];
  int tn = 0;
  if (cnt > 0) {
    if (cnt <= kTiny) {
      tn = (int)cnt;
#pragma unroll
      for (int x = 0; x < kTiny; ++x)
        if (x < tn) {
          tc[x] = bcol[st + x];
          tv[x] = BV ? bval[st + x] : 1ull;
        }
    } else if (cnt <= kQuarter) {
      packed = 1;
    } else {
      packed = ((int64_t)((cnt + kPieceLen - 1) / kPieceLen) << 32) | (1 << 16);
    }
  }
  int64_t total;
  const int64_t ex = block_exscan_i64_once(packed, L.red, &total);
#pragma unroll
  for (int x = 0; x < kTiny; ++x) call(x < tn, tc[x], BV ? a * tv[x] : a);
  const int32_t nquarter = (int32_t)(total & 0xffff), nlong = (int32_t)((total >> 16) & 0xffff);
  const int32_t npieces = (int32_t)(total >> 32);
  if (packed & 0xffff) {
    const int32_t slot = (int32_t)(ex & 0xffff);
    L.start[slot] = st;
    L.av[slot] = a;
    L.len[slot] = (int32_t)cnt;
  } else if (packed) {
    const int32_t li = (int32_t)((ex >> 16) & 0xffff);
    L.start[nquarter + li] = st;
    L.av[nquarter + li] = a;
    L.len[nquarter + li] = (int32_t)cnt;
    L.Q[li] = (int32_t)(ex >> 32);
  }
  if (threadIdx.x == 0) L.Q[nlong] = npieces;
  __syncthreads();
  // quarter segments, four per wave step (lane / 16 selects the segment), QH steps of loads in
  // flight per lane
  const int q = lane >> 4, x = lane & 15;
  for (int32_t g = wid; 4 * g < nquarter; g += QH * nwaves) {
    int32_t c[QH];
    uint64_t v[QH];
    bool ok[QH];
#pragma unroll
    for (int h = 0; h < QH; ++h) {
      const int32_t sidx = 4 * (g + h * nwaves) + q;
      ok[h] = sidx < nquarter && x < L.len[sidx];
      if (ok[h]) {
        const int64_t b = L.start[sidx] + x;
        c[h] = bcol[b];
        v[h] = BV ? bval[b] : 1ull;
      }
    }
#pragma unroll
    for (int h = 0; h < QH; ++h) {
      uint64_t av2 = 0;
      if (ok[h]) av2 = L.av[4 * (g + h * nwaves) + q];
      call(ok[h], c[h], BV ? av2 * v[h] : av2);
    }
  }
  // pieces of the long segments (E elements per lane), STEP pieces per wave step: E * STEP loads
  // in flight per lane.  The segment of piece ii is the number of long segments whose pieces end
  // at or before ii: with at most 64 long segments one ballot over their ends (read once per
  // round, a lane per segment) gives it, so the STEP lookups of a step do not wait on each other.
  const int32_t qend = lane < nlong ? L.Q[lane + 1] : INT32_MAX;
  int t = 0;
  for (int32_t i = wid; i < npieces; i += STEP * nwaves) {
    int32_t n[STEP];
    int64_t b[STEP];
    uint64_t av1[STEP];
#pragma unroll
    for (int h = 0; h < STEP; ++h) {
      const int32_t ii = i + h * nwaves;
      n[h] = 0;
      if (ii < npieces) {
        if (nlong <= kWave) {
          t = (int)__popcll(__ballot(qend <= ii));
        } else {
          while (uniform_i32(L.Q[t + 1]) <= ii) ++t;  // wave-uniform; segments are visited in order
        }
        const int32_t j = ii - uniform_i32(L.Q[t]);
        const int32_t sidx = nquarter + t;
        n[h] = min(kPieceLen, uniform_i32(L.len[sidx]) - j * kPieceLen);
        b[h] = uniform_i64(L.start[sidx]) + (int64_t)j * kPieceLen;
        av1[h] = uniform_u64(L.av[sidx]);
      }
    }
    int32_t c[STEP][E];
    uint64_t v[STEP][E];
#pragma unroll
    for (int h = 0; h < STEP; ++h) {
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const int xx = lane + u * kWave;
        if (xx < n[h]) {
          c[h][u] = bcol[b[h] + xx];
          v[h][u] = BV ? bval[b[h] + xx] : 1ull;
        }
      }
    }
#pragma unroll
    for (int h = 0; h < STEP; ++h) {
#pragma unroll
      for (int u = 0; u < E; ++u) call(lane + u * kWave < n[h], c[h][u], BV ? av1[h] * v[h][u] : av1[h]);
    }
  }
  __syncthreads();
}

template <bool AV, bool BV, int E = 1, int STEP = 8, class SegF, class F>
__device__ __forceinline__ void traverse(const int64_t rs, const int64_t re, const int32_t* __restrict__ acol,
                                         const uint64_t* __restrict__ aval, const int64_t* __restrict__ bptr,
                                         const int32_t* __restrict__ bcol, const uint64_t* __restrict__ bval,
                                         TravLds& L, SegF&& seg, F&& f) {
  const int chunk = min(kChunk, (int)blockDim.x);  // entries handled per round (one per thread)
  const bool mine = (int)threadIdx.x < chunk;
  // Two-stage prefetch of the entries: while round r expands, the B row extent of round r+1 (its
  // column arrived during round r-1) and the column of round r+2 are loaded, so a round with
  // short segments does not wait for acol, then bptr, then bcol.
  int32_t k0 = -1, k1 = -1;  // columns of rounds r and r+1 (-1: no entry)
  int64_t lo0 = 0, hi0 = 0;  // B row extent of round r's entry
  uint64_t a0 = 1ull, a1 = 1ull;
  {
    const int64_t e = rs + threadIdx.x;
    if (mine && e < re) {
      k0 = acol[e];
      a0 = AV ? aval[e] : 1ull;
    }
    if (k0 >= 0) {
      lo0 = bptr[k0];
      hi0 = bptr[k0 + 1];
    }
    const int64_t e1 = e + chunk;
    if (mine && e1 < re) {
      k1 = acol[e1];
      a1 = AV ? aval[e1] : 1ull;
    }
  }
  for (int64_t base = rs; base < re; base += chunk) {
    int64_t lo1 = 0, hi1 = 0;
    if (k1 >= 0) {
      lo1 = bptr[k1];
      hi1 = bptr[k1 + 1];
    }
    int32_t k2 = -1;
    uint64_t a2 = 1ull;
    const int64_t e2 = base + 2 * (int64_t)chunk + threadIdx.x;
    if (mine && e2 < re) {
      k2 = acol[e2];
      a2 = AV ? aval[e2] : 1ull;
    }
    int64_t st = 0, cnt = 0;
    uint64_t a = 1ull;
    if (k0 >= 0) {
      int64_t off = 0;
      seg(base + threadIdx.x, k0, lo0, hi0 - lo0, off, cnt);
      a = a0;
      st = lo0 + off;
    }
    trav_round<BV, E, STEP>(st, cnt, a, bcol, bval, L, f);
    k0 = k1;
    lo0 = lo1;
    hi0 = hi1;
    a0 = a1;
    k1 = k2;
    a1 = a2;
  }
}

// ------------------------------------------------------------------------------------------
// wave tier: ub <= 64.  One wave per row, 4 rows per block.
template <bool NUMERIC, bool AV, bool BV>
__global__ void __launch_bounds__(kThreads) wave_tier_kernel(
    const int64_t* __restrict__ aptr, const int32_t* __restrict__ acol, const uint64_t* __restrict__ aval,
    const int64_t* __restrict__ bptr, const int32_t* __restrict__ bcol, const uint64_t* __restrict__ bval,
    const int64_t* __restrict__ S, const int32_t* __restrict__ rows, int32_t nrows, int64_t* __restrict__ row_nnz,
    const int64_t* __restrict__ cptr, int32_t* __restrict__ ccol, uint64_t* __restrict__ cval) {
  __shared__ int32_t skey[kThreads];
  __shared__ uint64_t sval[NUMERIC ? kThreads : 1];
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  const int32_t ri = blockIdx.x * (kThreads / kWave) + wid;
  if (ri >= nrows) return;  // whole wave exits together
  const int32_t u = rows[ri];
  const int64_t rs = aptr[u], re = aptr[u + 1];
  const int64_t S0 = S[rs];
  const int ub = (int)(S[re] - S0);
  int32_t* key = skey + wid * kWave;
  uint64_t* val = sval + (NUMERIC ? wid * kWave : 0);
  // expand: each lane writes the products of one entry at their row positions
  for (int64_t cb = rs; cb < re; cb += kWave) {
    const int64_t e = cb + lane;
    if (e < re) {
      const int32_t k = acol[e];
      const int64_t b0 = bptr[k], b1 = bptr[k + 1];
      const int off = (int)(S[e] - S0);
      const uint64_t a = AV ? aval[e] : 1ull;
      for (int64_t b = b0; b < b1; ++b) {
        key[off + (b - b0)] = bcol[b];
        if constexpr (NUMERIC) val[off + (b - b0)] = BV ? a * bval[b] : a;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int32_t k = lane < ub ? key[lane] : INT32_MAX;
  uint64_t v = 0;
  if constexpr (NUMERIC) v = lane < ub ? val[lane] : 0ull;
  // bitonic sort across the wave (ascending by key)
#pragma unroll
  for (int size = 2; size <= kWave; size <<= 1) {
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1) {
      const int32_t ok = __shfl_xor(k, j);
      uint64_t ov = 0;
      if constexpr (NUMERIC) ov = __shfl_xor(v, j);
      const bool asc = (lane & size) == 0;
      const bool lower = (lane & j) == 0;
      const bool take = (lower == asc) ? (ok < k) : (ok > k);
      if (take) {
        k = ok;
        if constexpr (NUMERIC) v = ov;
      }
    }
  }
  const int32_t kprev = __shfl_up(k, 1);
  const bool valid = k != INT32_MAX;
  const bool head = valid && (lane == 0 || kprev != k);
  const uint64_t hmask = __ballot(head);
  if constexpr (!NUMERIC) {
    if (lane == 0) row_nnz[u] = __popcll(hmask);
  } else {
    // segmented inclusive scan of values over runs of equal keys
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint64_t y = __shfl_up(v, o);
      const int32_t ky = __shfl_up(k, o);
      if (lane >= o && ky == k) v += y;
    }
    const int32_t knext = __shfl_down(k, 1);
    const bool last = valid && (lane == kWave - 1 || knext != k);
    if (last) {
      const int64_t pos = cptr[u] + (__popcll(hmask & ((lane == 63) ? ~0ull : ((2ull << lane) - 1ull))) - 1);
      ccol[pos] = k;
      cval[pos] = v;
    }
  }
}

// ------------------------------------------------------------------------------------------
// mid tier for products with few columns (wave bitmap; rows with 64 < ub <= wave_bitmap_max_ub):
// one wave per row, four rows per block.
// The wave's own LDS region holds the row's column bitmap (nw 32-bit words), the exclusive
// popcount prefix of every word (16-bit, widths <= 65535) and cap value / column slots.  Symbolic:
// mark the columns, count the bits.  Numeric: mark, rank every word, accumulate each product at
// the rank of its column, list every column at its rank, then write the row in column order
// with consecutive lanes writing consecutive entries; a row wider than cap accumulates its values
// with device atomics at their final positions instead.  No
// workgroup barrier: the four waves of a block work on four rows independently.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct WaveBitmapLayout {
  int32_t nw = 0;        // bitmap words (ncols / 32, rounded up)
  int32_t cap = 0;       // value (and column) slots, >= every routed row's ub
  size_t rank_off = 0, val_off = 0, col_off = 0, per_wave = 0;
  static size_t bytes(int64_t nw, int64_t cap, bool numeric) {
    if (!numeric) return (size_t)nw * 4;
    return ((size_t)nw * 6 + 7) / 8 * 8 + (size_t)cap * 12;
  }
  static WaveBitmapLayout make(int32_t ncols, int32_t cap, bool numeric) {
    WaveBitmapLayout l;
    l.nw = (ncols + 31) / 32;
    l.cap = numeric ? cap : 0;
    l.rank_off = (size_t)l.nw * 4;
    l.val_off = (l.rank_off + (numeric ? (size_t)l.nw * 2 : 0) + 7) / 8 * 8;
    l.col_off = l.val_off + (size_t)l.cap * 8;
    l.per_wave = bytes(l.nw, cap, numeric);
    return l;
  }
};

// the mid tier (64 < ub <= wave_bitmap_max_ub) of a product with at most wave_bitmap_cols columns runs one
// wave per row over an LDS column bitmap; four waves' regions must fit the LDS of a workgroup
inline bool wave_bitmap_for(const SpgemmOpts& o, int32_t ncols) {
  if (o.wave_bitmap_cols <= 0 || ncols > o.wave_bitmap_cols || o.wave_bitmap_cap < 64 || o.wave_bitmap_cap > 65535 ||
      o.wave_bitmap_max_ub < 64 || o.wave_bitmap_max_ub > 65535)
    return false;
  const int64_t nw = (ncols + 31) / 32;
  return (int64_t)WaveBitmapLayout::bytes(nw, o.wave_bitmap_cap, true) * (kThreads / kWave) <= o.lds_bytes;
}

// f(col, val) for every product of row [rs, re) of A, one wave: lanes load 64 entries at a time;
// a B segment of at most 8 products is walked by its lane (all its loads issued first), longer
// ones by the whole wave with four loads in flight per lane
template <bool AV, bool BV, class F>
__device__ __forceinline__ void wave_products(int64_t rs, int64_t re, const int32_t* __restrict__ acol,
                                              const uint64_t* __restrict__ aval, const int64_t* __restrict__ bptr,
                                              const int32_t* __restrict__ bcol, const uint64_t* __restrict__ bval,
                                              F&& f) {
  constexpr int kLane = 8;
  const int lane = lane_id();
  for (int64_t cb = rs; cb < re; cb += kWave) {
    const int64_t e = cb + lane;
    int64_t b0 = 0, b1 = 0;
    uint64_t a = 1ull;
    if (e < re) {
      const int32_t k = acol[e];
      b0 = bptr[k];
      b1 = bptr[k + 1];
      if constexpr (AV) a = aval[e];
    }
    const int64_t len = b1 - b0;
    const bool longseg = len > kLane;
    if (!longseg && len > 0) {
      int32_t c[kLane];
      uint64_t v[kLane];
#pragma unroll
      for (int x = 0; x < kLane; ++x) {
        if (x < len) {
          c[x] = bcol[b0 + x];
          v[x] = BV ? bval[b0 + x] : 1ull;
        }
      }
#pragma unroll
      for (int x = 0; x < kLane; ++x)
        if (x < len) f(c[x], BV ? a * v[x] : a);
    }
    uint64_t mask = __ballot(longseg);
    while (mask) {
      const int src = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const int64_t s0 = __shfl(b0, src), s1 = __shfl(b1, src);
      const uint64_t sa = __shfl(a, src);
      for (int64_t b = s0 + lane; b < s1; b += 4 * kWave) {
        int32_t c[4];
        uint64_t v[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          if (b + x * kWave < s1) {
            c[x] = bcol[b + x * kWave];
            v[x] = BV ? bval[b + x * kWave] : 1ull;
          }
        }
#pragma unroll
        for (int x = 0; x < 4; ++x)
          if (b + x * kWave < s1) f(c[x], BV ? sa * v[x] : sa);
      }
    }
  }
}

template <bool NUMERIC, bool AV, bool BV>
__global__ void __launch_bounds__(kThreads) wave_bitmap_kernel(
    const int64_t* __restrict__ aptr, const int32_t* __restrict__ acol, const uint64_t* __restrict__ aval,
    const int64_t* __restrict__ bptr, const int32_t* __restrict__ bcol, const uint64_t* __restrict__ bval,
    const int32_t* __restrict__ rows, int32_t nrows, WaveBitmapLayout lay, int64_t* __restrict__ row_nnz,
    const int64_t* __restrict__ cptr, int32_t* __restrict__ ccol, uint64_t* __restrict__ cval) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  const int32_t ri = blockIdx.x * (kThreads / kWave) + wid;
  if (ri >= nrows) return;  // whole wave exits together
  unsigned char* base = lds + (size_t)wid * lay.per_wave;
  uint32_t* bits = (uint32_t*)base;
  uint16_t* rank = (uint16_t*)(base + lay.rank_off);
  uint64_t* vals = (uint64_t*)(base + lay.val_off);
  int32_t* cols = (int32_t*)(base + lay.col_off);
  const int32_t nw = lay.nw;
  const int32_t u = rows[ri];
  const int64_t rs = aptr[u], re = aptr[u + 1];
  for (int32_t w = lane; w < nw; w += kWave) bits[w] = 0u;
  wave_lds_sync();
  wave_products<false, false>(rs, re, acol, aval, bptr, bcol, bval,
                              [&](int32_t c, uint64_t) { atomicOr(&bits[c >> 5], 1u << (c & 31)); });
  wave_lds_sync();
  const int32_t wpl = (nw + kWave - 1) / kWave;
  const int32_t w0 = lane * wpl, w1 = min(w0 + wpl, nw);
  int32_t cnt = 0;
  for (int32_t w = w0; w < w1; ++w) cnt += __popc(bits[w]);
  int32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  const int32_t total = __shfl(incl, kWave - 1);
  if constexpr (!NUMERIC) {
    if (lane == 0) row_nnz[u] = total;
  } else {
    int32_t run = incl - cnt;
    for (int32_t w = w0; w < w1; ++w) {
      rank[w] = (uint16_t)run;
      run += __popc(bits[w]);
    }
    const int64_t o = cptr[u];
    if (total <= lay.cap) {
      for (int32_t i = lane; i < total; i += kWave) vals[i] = 0ull;
      wave_lds_sync();
      wave_products<AV, BV>(rs, re, acol, aval, bptr, bcol, bval, [&](int32_t c, uint64_t v) {
        const int32_t w = c >> 5;
        const int32_t r = rank[w] + __popc(bits[w] & ((1u << (c & 31)) - 1u));
        atomicAdd((unsigned long long*)&vals[r], (unsigned long long)v);
      });
      // every lane lists the columns of its own words at their ranks, then the row leaves in
      // column order with consecutive lanes writing consecutive entries
      for (int32_t w = w0; w < w1; ++w) {
        uint32_t x = bits[w];
        int32_t r = rank[w];
        while (x) {
          cols[r++] = (w << 5) | (__ffs(x) - 1);
          x &= x - 1u;
        }
      }
      wave_lds_sync();
      for (int32_t p = lane; p < total; p += kWave) {
        ccol[o + p] = cols[p];
        cval[o + p] = vals[p];
      }
    } else {
      // wider than the LDS value slots: columns written from the bitmap, values accumulated
      // with device atomics at their final positions (zeroed first)
      wave_lds_sync();
      for (int32_t w = w0; w < w1; ++w) {
        uint32_t x = bits[w];
        int64_t r = o + rank[w];
        while (x) {
          ccol[r++] = (w << 5) | (__ffs(x) - 1);
          x &= x - 1u;
        }
      }
      for (int32_t i = lane; i < total; i += kWave) cval[o + i] = 0ull;
      __threadfence();
      wave_products<AV, BV>(rs, re, acol, aval, bptr, bcol, bval, [&](int32_t c, uint64_t v) {
        const int32_t w = c >> 5;
        const int32_t r = rank[w] + __popc(bits[w] & ((1u << (c & 31)) - 1u));
        atomicAdd((unsigned long long*)&cval[o + r], (unsigned long long)v);
      });
    }
  }
}

// ------------------------------------------------------------------------------------------
// sort tier: one workgroup per row, products sorted in LDS.  Dynamic LDS: key[P] (+ val[P]).
template <bool NUMERIC, bool AV, bool BV>
__global__ void __launch_bounds__(kThreads) sort_tier_kernel(
    const int64_t* __restrict__ aptr, const int32_t* __restrict__ acol, const uint64_t* __restrict__ aval,
    const int64_t* __restrict__ bptr, const int32_t* __restrict__ bcol, const uint64_t* __restrict__ bval,
    const int64_t* __restrict__ S, const int32_t* __restrict__ rows, int32_t nrows, int cap,
    int64_t* __restrict__ row_nnz, const int64_t* __restrict__ cptr, int32_t* __restrict__ ccol,
    uint64_t* __restrict__ cval) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* val = (uint64_t*)smem;                                  // [cap] (numeric)
  int32_t* key = (int32_t*)(smem + (NUMERIC ? (size_t)cap * 8 : 0));  // [cap]
  __shared__ int32_t q[kThreads];
  __shared__ int qn;
  __shared__ int64_t red64[kThreads / kWave];
  __shared__ int32_t red32[kThreads / kWave];
  if (threadIdx.x == 0) qn = 0;
  for (int32_t ri = blockIdx.x; ri < nrows; ri += gridDim.x) {
    const int32_t u = rows[ri];
    const int64_t rs = aptr[u], re = aptr[u + 1];
    const int ub = (int)(S[re] - S[rs]);
    int P = 64;
    while (P < ub) P <<= 1;
    for (int i = ub + threadIdx.x; i < P; i += kThreads) key[i] = INT32_MAX;
    __syncthreads();
    for_each_product<AV, BV, true>(rs, re, acol, aval, bptr, bcol, bval, S, q, &qn,
                                   [&](int32_t c, uint64_t v, int64_t pos) {
                                     key[pos] = c;
                                     if constexpr (NUMERIC) val[pos] = v;
                                   });
    // bitonic sort of P keys (values follow)
    for (int size = 2; size <= P; size <<= 1) {
      for (int j = size >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < (P >> 1); i += kThreads) {
          const int lo = ((i / j) * 2 * j) + (i % j);
          const int hi = lo + j;
          const bool asc = (lo & size) == 0;
          const int32_t a = key[lo], b = key[hi];
          if ((a > b) == asc) {
            key[lo] = b;
            key[hi] = a;
            if constexpr (NUMERIC) {
              uint64_t t = val[lo];
              val[lo] = val[hi];
              val[hi] = t;
            }
          }
        }
        __syncthreads();
      }
    }
    // heads of runs; each thread owns a contiguous range of positions
    const int per = (P + kThreads - 1) / kThreads;
    const int p0 = threadIdx.x * per;
    const int p1 = min(p0 + per, ub);
    int32_t nh = 0;
    for (int p = p0; p < p1; ++p) nh += (p == 0 || key[p] != key[p - 1]) ? 1 : 0;
    if constexpr (!NUMERIC) {
      int64_t tot = block_sum_i64(nh, red64);
      if (threadIdx.x == 0) row_nnz[u] = tot;
    } else {
      int32_t total;
      int32_t rank = block_exscan_i32(nh, red32, &total);
      const int64_t obase = cptr[u];
      for (int p = p0; p < p1; ++p) {
        if (p == 0 || key[p] != key[p - 1]) {
          const int32_t kk = key[p];
          uint64_t s = val[p];
          for (int x = p + 1; x < ub && key[x] == kk; ++x) s += val[x];
          ccol[obase + rank] = kk;
          cval[obase + rank] = s;
          ++rank;
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// bitmap tier.  Dynamic LDS: bm[NW/32] u32 | sum[ceil(NW/2048)] u32 (8-byte padded) |
// rank[NW/64] u32 | val[VCAP] u64 (numeric).  Bit i of the summary marks a nonzero 64-bit word i
// of the bitmap, so scans, ranks, column writes and clears visit occupied words only.
__host__ __device__ inline int64_t bm_sum_words(int64_t nw) { return (nw / 64 + 31) / 32; }
__host__ __device__ inline int64_t bm_sum_bytes(int64_t nw) { return (bm_sum_words(nw) * 4 + 7) / 8 * 8; }
struct BitmapLayout {
  int64_t nw = 0;    // window width in columns (multiple of 128)
  int64_t vcap = 0;  // LDS value slots
  size_t bytes_sym() const { return (size_t)(nw / 8 + bm_sum_bytes(nw)); }
  size_t bytes_num() const { return bytes_sym() + (size_t)nw / 16 + (size_t)vcap * 8; }
};
// offsets inside the dynamic LDS block (nw multiple of 128 keeps every section 8-byte aligned)
__host__ __device__ inline int64_t bm_sum_off(int64_t nw) { return nw / 8; }
__host__ __device__ inline int64_t bm_rank_off(int64_t nw) { return nw / 8 + bm_sum_bytes(nw); }
__host__ __device__ inline int64_t bm_val_off(int64_t nw) { return bm_rank_off(nw) + nw / 16; }

template <bool NUMERIC, bool AV, bool BV>
__global__ void bitmap_tier_kernel(const int64_t* __restrict__ aptr, const int32_t* __restrict__ acol,
                                   const uint64_t* __restrict__ aval, const int64_t* __restrict__ bptr,
                                   const int32_t* __restrict__ bcol, const uint64_t* __restrict__ bval,
                                   const int32_t* __restrict__ rows0, int32_t n0, const int32_t* __restrict__ rows1,
                                   int32_t n1, int32_t ncols, int64_t nw, int64_t vcap, int32_t* __restrict__ queue,
                                   int64_t* __restrict__ row_nnz, const int64_t* __restrict__ cptr,
                                   int32_t* __restrict__ ccol, uint64_t* __restrict__ cval, int log2w = 0,
                                   int32_t nwin = 0, int32_t* __restrict__ wcnt = nullptr) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* bm = (uint32_t*)smem;
  uint64_t* bm64 = (uint64_t*)smem;
  uint32_t* sm = (uint32_t*)(smem + bm_sum_off(nw));
  uint32_t* rank = (uint32_t*)(smem + bm_rank_off(nw));
  uint64_t* sval = (uint64_t*)(smem + bm_val_off(nw));
  __shared__ TravLds L;
  __shared__ int32_t cur;
  __shared__ int64_t red64[1024 / kWave];
  __shared__ int32_t red32[1024 / kWave];
  const int nthreads = blockDim.x;
  const int nwaves = nthreads / kWave;
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  auto whole_row = [](int64_t, int32_t, int64_t, int64_t len, int64_t& off, int64_t& cnt) {
    off = 0;
    cnt = len;
  };
  const int32_t total_rows = n0 + n1;
  if constexpr (NUMERIC) {
    // bitmap and summary start all zero; each window of the numeric pass clears only the words
    // it set
    for (int64_t i = threadIdx.x; i < nw / 32; i += nthreads) bm[i] = 0u;
    for (int64_t i = threadIdx.x; i < bm_sum_words(nw); i += nthreads) sm[i] = 0u;
    __syncthreads();
  }
  while (true) {
    if (threadIdx.x == 0) cur = atomicAdd(queue, 1);
    __syncthreads();
    const int32_t ri = cur;
    __syncthreads();
    if (ri >= total_rows) break;
    const int32_t u = ri < n0 ? rows0[ri] : rows1[ri - n0];
    const int64_t rs = aptr[u], re = aptr[u + 1];
    int64_t width = 0;
    int64_t rowbase = 0;
    bool lds_vals = false;
    if constexpr (NUMERIC) {
      rowbase = cptr[u];
      width = cptr[u + 1] - rowbase;
      lds_vals = width <= vcap;
    }
    int64_t done = 0;  // output slots produced by earlier windows
    int64_t cnt = 0;
    for (int64_t w0 = 0; w0 < ncols; w0 += nw) {
      const int64_t wlen = min((int64_t)ncols - w0, nw);
      const int64_t nwords32 = (wlen + 31) / 32;
      const int64_t nwords64 = (wlen + 63) / 64;
      const int64_t nsum = (nwords64 + 31) / 32;
      if constexpr (!NUMERIC) {
        // the symbolic pass scans every word anyway: plain clear, no summary
        for (int64_t i = threadIdx.x; i < nwords64 * 2; i += nthreads) bm[i] = 0u;
        __syncthreads();
      }
      traverse<false, false, 4, 2>(rs, re, acol, nullptr, bptr, bcol, nullptr, L, whole_row,
                             [&](int32_t c, uint64_t) {
                               const int64_t x = (int64_t)c - w0;
                               if (x >= 0 && x < wlen) {
                                 if constexpr (NUMERIC) {
                                   // the first bit of a 32-bit word marks its 64-bit word
                                   if (atomicOr(&bm[x >> 5], 1u << (x & 31)) == 0u)
                                     atomicOr(&sm[x >> 11], 1u << ((x >> 6) & 31));
                                 } else {
                                   atomicOr(&bm[x >> 5], 1u << (x & 31));
                                 }
                               }
                             });
      if constexpr (!NUMERIC) {
        int64_t local = 0;
        for (int64_t i = threadIdx.x; i < nwords32; i += nthreads) local += __popc(bm[i]);
        cnt += block_sum_i64(local, red64);
        if (wcnt) {
          // columns per window of the window tier (2^log2w columns, aligned; w0 is a multiple
          // of 128 and 2^log2w >= 64, so no 64-bit word straddles two windows): a wave per window
          const int64_t xa = w0 >> log2w, xb = (w0 + wlen - 1) >> log2w;
          for (int64_t x = xa + wid; x <= xb; x += nwaves) {
            const int64_t ja = max((x << log2w) - w0, (int64_t)0) >> 6;
            const int64_t jb = min(((x + 1) << log2w) - w0, wlen + 63) >> 6;
            int64_t c = 0;
            for (int64_t j = ja + lane; j < jb; j += kWave) c += __popcll(bm64[j]);
            c = wave_sum_i64(c);
            if (lane == 0 && c) atomicAdd(&wcnt[(size_t)u * nwin + x], (int32_t)c);
          }
          __syncthreads();  // the next column window clears the bitmap
        }
      } else {
        // prefix count per 64-bit word: each thread scans a contiguous range of words
        const int64_t per = (nwords64 + nthreads - 1) / nthreads;
        const int64_t wa = threadIdx.x * per, wb = min(wa + per, nwords64);
        int32_t s = 0;
        for (int64_t i = wa; i < wb; ++i) s += __popcll(bm64[i]);
        int32_t wtot;
        int32_t off = block_exscan_i32(s, red32, &wtot);
        for (int64_t i = wa; i < wb; ++i) {
          rank[i] = (uint32_t)off;
          off += __popcll(bm64[i]);
        }
        __syncthreads();
        // columns (and zeroed values when accumulating globally): a wave per summary word, its
        // occupied 64-bit words one at a time, lane = bit
        for (int64_t q = wid; q < nsum; q += nwaves) {
          for (uint32_t bits = sm[q]; bits; bits &= bits - 1) {
            const int64_t i = q * 32 + __ffs(bits) - 1;
            const uint64_t word = bm64[i];
            if ((word >> lane) & 1ull) {
              const int64_t pos = rowbase + done + rank[i] + __popcll(word & ((1ull << lane) - 1ull));
              ccol[pos] = (int32_t)(w0 + i * 64 + lane);
              if (!lds_vals) cval[pos] = 0ull;
            }
          }
        }
        if (lds_vals) {
          for (int64_t i = threadIdx.x; i < wtot; i += nthreads) sval[i] = 0ull;
        } else {
          __threadfence();
        }
        __syncthreads();
        traverse<AV, BV, 4, 2>(rs, re, acol, aval, bptr, bcol, bval, L, whole_row, [&](int32_t c, uint64_t v) {
                                          const int64_t x = (int64_t)c - w0;
                                          if (x >= 0 && x < wlen) {
                                            const int64_t wd = x >> 6;
                                            const int32_t slot =
                                                (int32_t)rank[wd] + __popcll(bm64[wd] & ((1ull << (x & 63)) - 1ull));
                                            if (lds_vals)
                                              atomicAdd((unsigned long long*)&sval[slot], (unsigned long long)v);
                                            else
                                              atomicAdd((unsigned long long*)&cval[rowbase + done + slot],
                                                        (unsigned long long)v);
                                          }
                                        });
        if (lds_vals) {
          for (int64_t i = threadIdx.x; i < wtot; i += nthreads) cval[rowbase + done + i] = sval[i];
        }
        done += wtot;
        __syncthreads();
      }
      if constexpr (NUMERIC) {
        // clear the words this window set, then their summary
        for (int64_t q = threadIdx.x; q < nsum; q += nthreads) {
          for (uint32_t bits = sm[q]; bits; bits &= bits - 1) bm64[q * 32 + __ffs(bits) - 1] = 0ull;
          sm[q] = 0u;
        }
        __syncthreads();
      }
    }
    if constexpr (!NUMERIC) {
      if (threadIdx.x == 0) row_nnz[u] = cnt;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// window tier (numeric pass of rows wider than the bitmap tier's LDS value slots).
// A dense accumulator over a window of W = 2^log2w columns lives in LDS (W uint64 values +
// presence bits).  Windows are aligned (window w covers columns [w*W, (w+1)*W)); a bound table
// built once per product gives, for every row k of B, the first element of k in each window, so
// the segment of an entry in a window is found in O(1) and every product is visited exactly
// once.  A finished window is compacted in column order (wave per 64-bit presence word, lane =
// bit) and the row continues at the next window that holds any of its products.
struct WindowLayout {
  int64_t w = 0;     // window width in columns (power of two >= 64)
  int log2w = 6;
  int64_t nwin = 0;  // windows covering the columns of B
  int vbytes = 8;    // LDS accumulator width: 8, or 4 when no count of the product reaches 2^32
  size_t bytes() const { return (size_t)w * vbytes + (size_t)w / 8 + (size_t)w / 16; }
};

// Window-major bound table: bnd[w * nrows + k] = offset (within row k of B) of the first element
// with column >= w*W; bnd[nwin * nrows + k] = length of row k.  Window-major so that the two rows
// of the table a window reads (w and w + 1) are contiguous.  One wave per row of B; each entry
// written once.
__global__ void __launch_bounds__(kThreads) window_bounds_kernel(const int64_t* __restrict__ bptr,
                                                                 const int32_t* __restrict__ bcol, int32_t nrows,
                                                                 int log2w, int32_t nwin, int32_t* __restrict__ bnd) {
  const int lane = lane_id();
  const int64_t wpb = kThreads / kWave;
  for (int64_t k = (int64_t)blockIdx.x * wpb + threadIdx.x / kWave; k < nrows; k += (int64_t)gridDim.x * wpb) {
    const int64_t b0 = bptr[k];
    const int32_t len = (int32_t)(bptr[k + 1] - b0);
    int32_t* out = bnd + k;
    for (int32_t x = lane; x < len; x += kWave) {
      const int32_t w = bcol[b0 + x] >> log2w;
      const int32_t wp = x == 0 ? -1 : (bcol[b0 + x - 1] >> log2w);
      for (int32_t ww = wp + 1; ww <= w; ++ww) out[(size_t)ww * nrows] = x;
    }
    const int32_t wl = len > 0 ? (bcol[b0 + len - 1] >> log2w) : -1;
    for (int32_t ww = wl + 1 + lane; ww <= nwin; ww += kWave) out[(size_t)ww * nrows] = len;
  }
}

// Segment table of B per group of G numeric windows: seg[g * nrows + k] = absolute [first, end)
// of the elements of row k of B with columns in windows [g*G, min((g+1)*G, nwin)).  A task then
// gets its segment with one 16-byte load; from the bound table it takes three (two offsets in
// different rows of the table, then the row start).
__global__ void __launch_bounds__(kThreads) window_segs_kernel(const int64_t* __restrict__ bptr,
                                                               const int32_t* __restrict__ bnd, int32_t nrows,
                                                               int32_t nwin, int32_t G, int32_t ngroups,
                                                               longlong2* __restrict__ seg) {
  const int64_t total = (int64_t)ngroups * nrows;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kThreads) {
    const int32_t g = (int32_t)(i / nrows), k = (int32_t)(i - (int64_t)g * nrows);
    const int32_t wa = g * G, wb = min(wa + G, nwin);
    const int64_t b = bptr[k];
    seg[i] = make_longlong2(b + bnd[(size_t)wa * nrows + k], b + bnd[(size_t)wb * nrows + k]);
  }
}

template <bool AV, bool BV>
__global__ void window_tier_kernel(const int64_t* __restrict__ aptr, const int32_t* __restrict__ acol,
                                   const uint64_t* __restrict__ aval, const int64_t* __restrict__ bptr,
                                   const int32_t* __restrict__ bcol, const uint64_t* __restrict__ bval,
                                   const int32_t* __restrict__ rows, int32_t nrows, int32_t ncols, int log2w,
                                   int32_t nwin, const int32_t* __restrict__ bnd, int32_t nbrows,
                                   int32_t* __restrict__ queue, const int64_t* __restrict__ cptr,
                                   int32_t* __restrict__ ccol, uint64_t* __restrict__ cval) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t W = 1 << log2w;
  uint64_t* vals = (uint64_t*)smem;
  uint32_t* pres = (uint32_t*)(smem + (size_t)W * 8);
  uint64_t* pres64 = (uint64_t*)pres;
  __shared__ TravLds L;
  __shared__ int32_t cur, wmin, wmax;
  __shared__ int32_t wcnt[1024 / kWave];
  const int nthreads = blockDim.x, nwaves = nthreads / kWave;
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  const int chunk = min(kChunk, nthreads);
  const uint64_t below = (1ull << lane) - 1ull;
  for (int i = threadIdx.x; i < W; i += nthreads) vals[i] = 0ull;
  for (int i = threadIdx.x; i < W / 32; i += nthreads) pres[i] = 0u;
  __syncthreads();
  while (true) {
    if (threadIdx.x == 0) {
      cur = atomicAdd(queue, 1);
      wmin = INT32_MAX;
      wmax = -1;
    }
    __syncthreads();
    const int32_t ri = cur;
    if (ri >= nrows) break;
    const int32_t u = rows[ri];
    const int64_t rs = aptr[u], re = aptr[u + 1];
    const int64_t rowbase = cptr[u];
    const int64_t width = cptr[u + 1] - rowbase;
    // rows with at most `chunk` entries keep their entry (B row, A value, position) in registers
    const bool fast = re - rs <= chunk;
    int32_t my_k = 0;
    int64_t my_b0 = 0;
    uint64_t my_a = 1ull;
    // window range of the row: first / last window holding one of its products
    for (int64_t e = rs + threadIdx.x; e < re; e += nthreads) {
      const int32_t k = acol[e];
      const int64_t b0 = bptr[k];
      const int32_t len = (int32_t)(bptr[k + 1] - b0);
      if (len > 0) {
        atomicMin(&wmin, bcol[b0] >> log2w);
        atomicMax(&wmax, bcol[b0 + len - 1] >> log2w);
      }
      if (fast) {
        my_k = k;
        my_b0 = b0;
        my_a = AV ? aval[e] : 1ull;
      }
    }
    __syncthreads();
    const int32_t wfirst = wmin, wlast = wmax;
    // pass 0 accumulates values only and takes the nonzero slots as the row's columns; if that
    // finds fewer columns than the symbolic width (a count summed to 0 mod 2^64), pass 1 redoes
    // the row with explicit presence bits
    for (int pass = 0; pass < 2; ++pass) {
      const bool with_pres = pass == 1;
      int64_t done = 0;
      int32_t my_s = 0;  // fast path: position of the entry's next element in its B row
      for (int32_t w = wfirst; w <= wlast; ++w) {
        const int64_t w0 = (int64_t)w << log2w;
        auto acc = [&](int32_t c, uint64_t v) {
          const int32_t x = (int32_t)(c - w0);
          atomicAdd((unsigned long long*)&vals[x], (unsigned long long)v);
          if (with_pres) atomicOr(&pres[x >> 5], 1u << (x & 31));
        };
        if (fast) {
          int64_t st = 0, cnt = 0;
          if (rs + (int64_t)threadIdx.x < re) {
            const int32_t s1 = bnd[(size_t)(w + 1) * nbrows + my_k];
            st = my_b0 + my_s;
            cnt = s1 - my_s;
            my_s = s1;
          }
          trav_round<BV>(st, cnt, my_a, bcol, bval, L, acc);
        } else {
          const int32_t wcur = w;
          traverse<AV, BV>(
              rs, re, acol, aval, bptr, bcol, bval, L,
              [&](int64_t, int32_t k, int64_t, int64_t, int64_t& off, int64_t& cnt) {
                const int32_t* bk = bnd + (size_t)wcur * nbrows + k;
                off = bk[0];
                cnt = bk[nbrows] - off;
              },
              acc);
        }
        // compaction in column order: each wave owns a contiguous run of 64-column words
        const int64_t wend = min(w0 + (int64_t)W, (int64_t)ncols);
        const int32_t nwords = (int32_t)((wend - w0 + 63) / 64);
        const int32_t per = (nwords + nwaves - 1) / nwaves;
        const int32_t ja = min(wid * per, nwords), jb = min(ja + per, nwords);
        int32_t mine = 0;
        for (int32_t j = ja; j < jb; ++j) {
          const uint64_t m = with_pres ? pres64[j] : __ballot(vals[j * 64 + lane] != 0ull);
          mine += __popcll(m);
        }
        if (lane == 0) wcnt[wid] = mine;
        __syncthreads();
        int32_t woff = 0, wtot = 0;
        for (int v2 = 0; v2 < nwaves; ++v2) {
          const int32_t cw = wcnt[v2];
          woff += v2 < wid ? cw : 0;
          wtot += cw;
        }
        int64_t pos0 = done + woff;
        for (int32_t j = ja; j < jb; ++j) {
          const int32_t xi = j * 64 + lane;
          const uint64_t v = vals[xi];
          const uint64_t m = with_pres ? pres64[j] : __ballot(v != 0ull);
          if ((m >> lane) & 1ull) {
            const int64_t slot = pos0 + __popcll(m & below);
            if (slot < width) {
              ccol[rowbase + slot] = (int32_t)(w0 + xi);
              cval[rowbase + slot] = v;
            }
          }
          if (m) vals[xi] = 0ull;
          if (with_pres && lane == 0) pres64[j] = 0ull;
          pos0 += __popcll(m);
        }
        done += wtot;
        // the next window's first accumulations (inline tiny segments) precede its first barrier:
        // every wave must be done with this window's slots before any wave moves on
        __syncthreads();
      }
      __syncthreads();
      if (done == width) break;
    }
  }
}

// Window-major window tier.  Task t = w * nrows + ri is the part of the ri-th wide row that falls
// in window w; tasks are claimed in that order, so all workgroups work on the same window (or two
// neighbours) at a time and the B elements they gather - the window's slice of every B row - stay
// in L2 instead of being fetched again for every row.  The symbolic pass counted each row's
// columns per window, which gives every task its output offset and size (empty tasks end at
// once).  Pass 0 accumulates values only and takes the nonzero slots as the window's columns;
// when that finds fewer columns than counted (a sum that wrapped to 0 mod 2^64) the window is
// redone with presence bits before anything is written.
struct WinRow {
  int64_t rs, re;  // extent of the row in A
  int64_t base;    // absolute output offset of the row
};

// Task metadata of one launch: info[ri] for the ri-th wide row and the exclusive prefix of its
// per-window counts, pre[ri * (nwin + 1) + w] = columns in windows < w (pre[.. + nwin] = width).
// One wave per row.
__global__ void __launch_bounds__(kThreads) window_prep_kernel(const int32_t* __restrict__ rows, int32_t nrows,
                                                               const int64_t* __restrict__ aptr,
                                                               const int64_t* __restrict__ cptr,
                                                               const int32_t* __restrict__ wcnt, int32_t nwin,
                                                               WinRow* __restrict__ info, int32_t* __restrict__ pre) {
  const int lane = lane_id();
  const int64_t wpb = kThreads / kWave;
  for (int64_t ri = (int64_t)blockIdx.x * wpb + threadIdx.x / kWave; ri < nrows; ri += (int64_t)gridDim.x * wpb) {
    const int32_t u = rows[ri];
    if (lane == 0) info[ri] = WinRow{aptr[u], aptr[u + 1], cptr[u]};
    const int32_t* wc = wcnt + (size_t)u * nwin;
    int32_t* out = pre + (size_t)ri * (nwin + 1);
    if (lane == 0) out[0] = 0;
    int32_t carry = 0;
    for (int32_t x0 = 0; x0 < nwin; x0 += kWave) {
      int32_t v = x0 + lane < nwin ? wc[x0 + lane] : 0;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const int32_t y = __shfl_up(v, o);
        if (lane >= o) v += y;
      }
      if (x0 + lane < nwin) out[x0 + lane + 1] = carry + v;
      carry += __shfl(v, kWave - 1);
    }
  }
}

// Tasks are software-pipelined: the task after next is claimed, and the next task's metadata
// loaded, while the current task runs; the next task's first-round A entries are loaded while the
// current task writes its output.  Only the B-row bounds and the gathers are left on a task's
// critical path.
// VT: the LDS accumulator (uint32_t only when the host has bounded every count of the product
// below 2^32: twice the columns per window for the same LDS, so half the tasks)
template <bool AV, bool BV, bool PRES, class VT = uint64_t>
__global__ void window_major_kernel(const int32_t* __restrict__ acol, const uint64_t* __restrict__ aval,
                                    const int64_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
                                    const uint64_t* __restrict__ bval, const WinRow* __restrict__ info,
                                    const int32_t* __restrict__ pre, int32_t nrows, int32_t ncols, int log2w,
                                    int32_t nwin, const int32_t* __restrict__ bnd,
                                    const longlong2* __restrict__ seg, int32_t nbrows,
                                    int32_t* __restrict__ queue, int32_t* __restrict__ ccol,
                                    uint64_t* __restrict__ cval) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t W = 1 << log2w;
  VT* vals = (VT*)smem;
  uint32_t* pres = (uint32_t*)(smem + (size_t)W * sizeof(VT));
  uint64_t* pres64 = (uint64_t*)pres;
  __shared__ TravLds L;
  // task records by parity: slot p holds the current task, slot p ^ 1 receives the next one
  // (id, A extent, output offset, column count, and the id of the task after it) before the
  // compaction barrier, so no barrier is needed between tasks
  __shared__ int32_t slot_t[2], slot_nt[2];
  __shared__ int64_t slot_m[2][4];
  __shared__ int32_t wcw[1024 / kWave];
  const int nthreads = blockDim.x, nwaves = nthreads / kWave;
  const int lane = lane_id(), wid = wave_id();
  const int chunk = min(kChunk, nthreads);
  const uint64_t below = (1ull << lane) - 1ull;
  const int64_t ntasks = (int64_t)nrows * nwin;
  // metadata of task t (thread 0): row extent, output offset of the window, its column count
  auto task_meta = [&](int32_t t, int64_t* m) {
    if (t >= ntasks) {
      m[0] = m[1] = m[2] = m[3] = 0;
      return;
    }
    const int32_t w = t / nrows, ri = t - w * nrows;
    const WinRow r = info[ri];
    const int32_t* p = pre + (size_t)ri * (nwin + 1) + w;
    const int32_t p0 = p[0], p1 = p[1];
    m[0] = r.rs;
    m[1] = r.re;
    m[2] = r.base + p0;
    m[3] = p1 - p0;
  };
  for (int i = threadIdx.x; i < W; i += nthreads) vals[i] = 0;
  for (int i = threadIdx.x; i < W / 32; i += nthreads) pres[i] = 0u;
  if (threadIdx.x == 0) {
    const int32_t t0 = atomicAdd(queue, 1);
    int64_t m[4];
    task_meta(t0, m);
    slot_t[0] = t0;
    for (int q = 0; q < 4; ++q) slot_m[0][q] = m[q];
    slot_nt[0] = atomicAdd(queue, 1);
  }
  __syncthreads();
  int par = 0;
  bool have_pref = false;  // the first round's entries of the current task are in registers
  int32_t pk = 0;          // prefetched entry: B row, A value
  uint64_t pa = 1ull;
  bool pvalid = false;
  while (true) {
    // (slot par was published before the previous task's compaction barrier or a skip barrier;
    // it is rewritten two tasks later, after this task's traversal barriers)
    const int32_t t = uniform_i32(slot_t[par]);
    if (t >= ntasks) break;
    const int32_t tn = uniform_i32(slot_nt[par]);
    const int64_t rs = uniform_i64(slot_m[par][0]), re = uniform_i64(slot_m[par][1]);
    const int64_t rowbase = uniform_i64(slot_m[par][2]);
    const int32_t target = (int32_t)uniform_i64(slot_m[par][3]);
    const int32_t w = t / nrows;
    // thread 0: claim the task after next and load the next task's metadata (registers)
    int32_t tnn = 0;
    int64_t nm[4] = {0, 0, 0, 0};
    if (threadIdx.x == 0) {
      tnn = atomicAdd(queue, 1);
      task_meta(tn, nm);
    }
    if (target == 0) {
      if (threadIdx.x == 0) {
        slot_t[par ^ 1] = tn;
        for (int q = 0; q < 4; ++q) slot_m[par ^ 1][q] = nm[q];
        slot_nt[par ^ 1] = tnn;
      }
      __syncthreads();
      par ^= 1;
      have_pref = false;
      continue;
    }
    const int64_t w0 = (int64_t)w << log2w;
    const int32_t* b_lo = bnd + (size_t)w * nbrows;
    const int32_t* b_hi = b_lo + nbrows;
    const longlong2* segw = seg ? seg + (size_t)w * nbrows : nullptr;  // one load per segment
    // PRES: presence bits from the start (one pass; the compaction touches only occupied slots)
    for (int pass = PRES ? 1 : 0; pass < 2; ++pass) {
      const bool with_pres = pass == 1;
      auto acc = [&](int32_t c, uint64_t v) {
        const int32_t x = (int32_t)(c - w0);
        if constexpr (sizeof(VT) == 8) atomicAdd((unsigned long long*)&vals[x], (unsigned long long)v);
        else atomicAdd((unsigned int*)&vals[x], (unsigned int)v);
        if (with_pres) atomicOr(&pres[x >> 5], 1u << (x & 31));
      };
      for (int64_t base = rs; base < re; base += chunk) {
        int64_t st = 0, cnt = 0;
        uint64_t a = 1ull;
        bool valid;
        int32_t k = 0;
        if (pass == (PRES ? 1 : 0) && base == rs && have_pref) {
          valid = pvalid;
          k = pk;
          a = pa;
        } else {
          const int64_t e = base + threadIdx.x;
          valid = (int)threadIdx.x < chunk && e < re;
          if (valid) {
            k = acol[e];
            a = AV ? aval[e] : 1ull;
          }
        }
        if (valid) {
          if (segw) {
            const longlong2 sg = segw[k];
            st = sg.x;
            cnt = sg.y - sg.x;
          } else {
            const int32_t s0 = b_lo[k];
            cnt = b_hi[k] - s0;
            st = bptr[k] + s0;
          }
        }
        trav_round<BV, 1, 8, 4>(st, cnt, a, bcol, bval, L, acc);
      }
      // compaction in column order: each wave owns a contiguous run of 64-column words; word
      // masks, counts and offsets are wave-uniform (scalar registers)
      const int64_t wend = min(w0 + (int64_t)W, (int64_t)ncols);
      const int32_t nwords = (int32_t)((wend - w0 + 63) / 64);
      const int32_t per = (nwords + nwaves - 1) / nwaves;
      const int32_t ja = min(wid * per, nwords), jb = min(ja + per, nwords);
      // presence bits: the wave's run of at most 64 words is read at once, a lane per word, and
      // kept for the write pass.  Values alone: the words are taken kCompactBatch at a time so that
      // their LDS reads are in flight together (a read past the wave's run is clamped to its last
      // word and not used).
      int32_t mine = 0;
      uint64_t pm = 0ull;  // presence mode: bits of word ja + lane
      if (with_pres) {
        if (ja + lane < jb) pm = pres64[ja + lane];
        mine = (int32_t)wave_sum_i64(__popcll(pm));
      } else {
        for (int32_t j = ja; j < jb; j += kCompactBatch) {
          uint64_t m[kCompactBatch];
#pragma unroll
          for (int u = 0; u < kCompactBatch; ++u) m[u] = __ballot(vals[min(j + u, jb - 1) * 64 + lane] != 0);
#pragma unroll
          for (int u = 0; u < kCompactBatch; ++u) mine += j + u < jb ? __popcll(m[u]) : 0;
        }
      }
      if (lane == 0) wcw[wid] = mine;
      if (threadIdx.x == 0) {
        slot_t[par ^ 1] = tn;
        for (int q = 0; q < 4; ++q) slot_m[par ^ 1][q] = nm[q];
        slot_nt[par ^ 1] = tnn;
      }
      __syncthreads();
      int32_t woff = 0, wtot = 0;
      for (int v2 = 0; v2 < nwaves; ++v2) {
        const int32_t cw = uniform_i32(wcw[v2]);
        woff += v2 < wid ? cw : 0;
        wtot += cw;
      }
      const bool write = with_pres || wtot == target;
      if (write) {
        // the next task's first-round entries: their loads overlap this task's output writes
        const int64_t nrs = uniform_i64(slot_m[par ^ 1][0]), nre = uniform_i64(slot_m[par ^ 1][1]);
        const int64_t e = nrs + threadIdx.x;
        pvalid = (int)threadIdx.x < chunk && e < nre;
        pk = 0;
        pa = 1ull;
        if (pvalid) {
          pk = acol[e];
          pa = AV ? aval[e] : 1ull;
        }
      }
      // this task's output run; slots are below the window's width, so 32-bit offsets from it
      int32_t* const cc = ccol + rowbase;
      uint64_t* const cv = cval + rowbase;
      const int32_t c0 = (int32_t)w0;
      int32_t pos0 = woff;
      if (with_pres) {
        // occupied words only, in column order
        for (uint64_t nzw = __ballot(pm != 0ull); nzw; nzw &= nzw - 1) {
          const int i = __ffsll((long long)nzw) - 1;
          const uint64_t m = readlane_u64(pm, i);
          const int32_t xi = (ja + i) * 64 + lane;
          if ((m >> lane) & 1ull) {
            const uint64_t v = (uint64_t)vals[xi];
            const uint32_t slot = (uint32_t)(pos0 + lanes_below(m));
            if ((int32_t)slot < target) {
              __builtin_nontemporal_store(c0 + xi, &cc[slot]);
              __builtin_nontemporal_store(v, &cv[slot]);
            }
            vals[xi] = 0;
          }
          if (lane == 0) pres64[ja + i] = 0ull;
          pos0 += __popcll(m);
        }
      } else {
        // written only when the nonzero slots number exactly `target`, so every slot is in range
        for (int32_t j = ja; j < jb; j += kCompactBatch) {
          uint64_t v[kCompactBatch];
#pragma unroll
          for (int u = 0; u < kCompactBatch; ++u) v[u] = (uint64_t)vals[min(j + u, jb - 1) * 64 + lane];
#pragma unroll
          for (int u = 0; u < kCompactBatch; ++u) {
            if (j + u < jb) {
              const int32_t xi = (j + u) * 64 + lane;
              const bool nz = v[u] != 0ull;
              const uint64_t m = __ballot(nz);
              if (write && nz) {
                const uint32_t slot = (uint32_t)(pos0 + lanes_below(m));
                __builtin_nontemporal_store(c0 + xi, &cc[slot]);
                __builtin_nontemporal_store(v[u], &cv[slot]);
              }
              if (m) vals[xi] = 0;
              pos0 += __popcll(m);
            }
          }
        }
      }
      // no barrier after a written window: the next task's first LDS writes (the traversal
      // scan, then accumulations) come after that scan's barrier, which every wave reaches only
      // when done with this window.  A redo (pass 1) reuses the slots at once, so it waits.
      if (write) break;
      __syncthreads();
    }
    par ^= 1;
    have_pref = true;
  }
}

// Window-major symbolic pass for the heavy rows: task t = w * nrows + ri counts the distinct
// columns of row rows[ri] in window w with an LDS bitmap of the window, writes the count to
// wcnt[u * nwin + w] (the window tier's output offsets) and adds it to the row's width.  Tasks
// run window by window for the same reason as in the numeric pass: the window's slice of B stays
// in L2.
struct SymRow {
  int64_t rs, re;
  int32_t u, pad;
};

__global__ void __launch_bounds__(kThreads) sym_prep_kernel(const int32_t* __restrict__ rows, int32_t nrows,
                                                            const int64_t* __restrict__ aptr,
                                                            SymRow* __restrict__ info) {
  for (int64_t ri = (int64_t)blockIdx.x * kThreads + threadIdx.x; ri < nrows; ri += (int64_t)gridDim.x * kThreads) {
    const int32_t u = rows[ri];
    info[ri] = SymRow{aptr[u], aptr[u + 1], u, 0};
  }
}

// A symbolic window spans 2^sub numeric windows (2^(log2w + sub) columns): fewer, larger tasks;
// the count of every numeric window inside it is taken from its part of the bitmap.
template <bool MERGE>
__global__ void window_symbolic_kernel(const int32_t* __restrict__ acol, const int64_t* __restrict__ bptr,
                                       const int32_t* __restrict__ bcol, const SymRow* __restrict__ info,
                                       int32_t nrows, int log2w, int sub, int32_t nwin,
                                       const int32_t* __restrict__ bnd, const longlong2* __restrict__ seg,
                                       int32_t nbrows, int32_t* __restrict__ queue,
                                       int32_t* __restrict__ wcnt, int64_t* __restrict__ row_nnz) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* bm = (uint32_t*)smem;
  const int32_t R = 1 << sub;                   // numeric windows per symbolic window
  const int32_t nw32 = (1 << (log2w + sub)) / 32;
  const int32_t sub32 = (1 << log2w) / 32;      // bitmap words per numeric window
  const int32_t lw = log2w - 5, wmask = sub32 - 1;
  const int32_t nwin_s = (nwin + R - 1) / R;
  __shared__ TravLds L;
  __shared__ int32_t claim;
  __shared__ int64_t meta[3];
  const int nthreads = blockDim.x, nwaves = nthreads / kWave;
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  const int chunk = min(kChunk, nthreads);
  const int64_t ntasks = (int64_t)nrows * nwin_s;
  for (int i = threadIdx.x; i < nw32; i += nthreads) bm[i] = 0u;
  if (threadIdx.x == 0) claim = atomicAdd(queue, 1);
  __syncthreads();
  while (true) {
    const int32_t t = uniform_i32(claim);
    if (t >= ntasks) break;
    const int32_t ws = t / nrows;
    __syncthreads();  // every thread has read `claim`
    if (threadIdx.x == 0) {
      const int32_t nx = atomicAdd(queue, 1);
      const SymRow r = info[t - ws * nrows];
      meta[0] = r.rs;
      meta[1] = r.re;
      meta[2] = r.u;
      claim = nx;
    }
    __syncthreads();
    const int64_t rs = uniform_i64(meta[0]), re = uniform_i64(meta[1]);
    const int32_t u = (int32_t)uniform_i64(meta[2]);
    const int32_t wa = ws * R, wb = min(wa + R, nwin);  // numeric windows [wa, wb)
    const int64_t w0 = (int64_t)wa << log2w;
    const int32_t* b_lo = bnd + (size_t)wa * nbrows;
    const int32_t* b_hi = bnd + (size_t)wb * nbrows;
    const longlong2* segw = seg ? seg + (size_t)ws * nbrows : nullptr;  // one load per segment
    for (int64_t base = rs; base < re; base += chunk) {
      int64_t st = 0, cnt = 0;
      const int64_t e = base + threadIdx.x;
      if ((int)threadIdx.x < chunk && e < re) {
        const int32_t k = acol[e];
        if (segw) {
          const longlong2 sg = segw[k];
          st = sg.x;
          cnt = sg.y - sg.x;
        } else {
          const int32_t s0 = b_lo[k];
          cnt = b_hi[k] - s0;
          st = bptr[k] + s0;
        }
      }
      if constexpr (MERGE) {
        // very dense rows: lanes whose columns share a 32-bit word are merged (segmented OR over
        // the wave) and one lane per word sets the bits
        trav_round<false, 4, 2, 2, true>(st, cnt, 1ull, bcol, nullptr, L, [&](bool valid, int32_t c, uint64_t) {
          const int32_t x = (int32_t)(c - w0);
          const int32_t word = valid ? (x >> 5) : -1;
          uint32_t bit = valid ? (1u << (x & 31)) : 0u;
#pragma unroll
          for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t ob = __shfl_up(bit, o);
            const int32_t ow = __shfl_up(word, o);
            if (lane >= o && ow == word) bit |= ob;
          }
          const int32_t nw = __shfl_down(word, 1);
          if (valid && (lane == kWave - 1 || nw != word)) atomicOr(&bm[word], bit);
        });
      } else {
        // bit layout inside a numeric window: column x sets bit x >> lw of word x & (words - 1), so
        // the consecutive columns of a dense B row fall in consecutive words (distinct banks)
        // instead of sharing one word; a window's words stay contiguous for its count
        trav_round<false, 4, 2>(st, cnt, 1ull, bcol, nullptr, L, [&](int32_t c, uint64_t) {
          const int32_t x = (int32_t)(c - w0);
          atomicOr(&bm[((x >> log2w) << lw) | (x & wmask)], 1u << ((x >> lw) & 31));
        });
      }
    }
    // count per numeric window (a wave each) and clear; trav_round ended with a barrier
    int64_t total = 0;
    for (int32_t x = wid; x < wb - wa; x += nwaves) {
      int32_t c = 0;
      for (int32_t i = lane; i < sub32; i += kWave) {
        c += __popc(bm[x * sub32 + i]);
        bm[x * sub32 + i] = 0u;
      }
      c = (int32_t)wave_sum_i64(c);
      if (lane == 0 && c > 0) wcnt[(size_t)u * nwin + wa + x] = c;
      total += c;
    }
    if (lane == 0 && total > 0) atomicAdd((unsigned long long*)&row_nnz[u], (unsigned long long)total);
    __syncthreads();  // the bitmap is clear before the next task's first accumulations
  }
}

// split the bitmap-tier rows (heavy first) by exact width: <= vcap stay in the bitmap kernel,
// wider rows go to the window kernel; per-class sums of A entries, products and widths
__global__ void __launch_bounds__(kThreads) split_wide_kernel(const int32_t* __restrict__ rows0, int32_t n0,
                                                              const int32_t* __restrict__ rows1, int32_t n1,
                                                              const int64_t* __restrict__ aptr,
                                                              const int64_t* __restrict__ S,
                                                              const int64_t* __restrict__ cptr, int64_t vcap,
                                                              int32_t* __restrict__ narrow, int32_t* __restrict__ wide,
                                                              int32_t* __restrict__ counts,
                                                              unsigned long long* __restrict__ sums) {
  const int32_t i = blockIdx.x * kThreads + threadIdx.x;
  int cls = -1;
  int32_t u = 0;
  int64_t d = 0, ub = 0, w = 0;
  if (i < n0 + n1) {
    u = i < n0 ? rows0[i] : rows1[i - n0];
    w = cptr[u + 1] - cptr[u];
    d = aptr[u + 1] - aptr[u];
    ub = S[aptr[u + 1]] - S[aptr[u]];
    cls = w > vcap ? 1 : 0;
  }
  const int lane = lane_id();
  {
    int32_t wmax = cls == 0 ? (int32_t)w : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o));
    if (lane == 0 && wmax > 0) atomicMax(&counts[2], wmax);
  }
  for (int c = 0; c < 2; ++c) {
    const uint64_t mask = __ballot(cls == c);
    if (!mask) continue;
    const int leader = __ffsll((long long)mask) - 1;
    const int64_t sd = wave_sum_i64(cls == c ? d : 0);
    const int64_t su = wave_sum_i64(cls == c ? ub : 0);
    const int64_t sw = wave_sum_i64(cls == c ? w : 0);
    int32_t base = 0;
    if (lane == leader) {
      base = atomicAdd(&counts[c], (int32_t)__popcll(mask));
      atomicAdd(&sums[3 * c], (unsigned long long)sd);
      atomicAdd(&sums[3 * c + 1], (unsigned long long)su);
      atomicAdd(&sums[3 * c + 2], (unsigned long long)sw);
    }
    base = __shfl(base, leader);
    if (cls == c) (c == 0 ? narrow : wide)[base + lanes_below(mask)] = u;
  }
}

// ------------------------------------------------------------------------------------------
// copy tier: a row of A with a single entry a (column k) whose B row repeats no column.  Row u of
// C is then a * B[k] with B's columns in B's (ascending) order, so its width is |B[k]| (no
// symbolic work) and its numeric pass is a streaming copy: the copy rows' outputs are laid end to
// end and cut into kCopyChunk-element chunks, one workgroup per chunk, consecutive threads on
// consecutive elements.
constexpr int64_t kCopyChunk = 8192;

__global__ void __launch_bounds__(kThreads) copy_width_kernel(const int32_t* __restrict__ rows, int32_t n,
                                                              const int64_t* __restrict__ aptr,
                                                              const int64_t* __restrict__ S,
                                                              int64_t* __restrict__ row_nnz) {
  for (int32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int32_t u = rows[i];
    row_nnz[u] = S[aptr[u] + 1] - S[aptr[u]];
  }
}

// per copy row: its length, first B element, first output slot and A value
__global__ void __launch_bounds__(kThreads) copy_prep_kernel(const int32_t* __restrict__ rows, int32_t n,
                                                             const int64_t* __restrict__ aptr,
                                                             const int32_t* __restrict__ acol,
                                                             const uint64_t* __restrict__ aval,
                                                             const int64_t* __restrict__ bptr,
                                                             const int64_t* __restrict__ cptr,
                                                             int64_t* __restrict__ len, int64_t* __restrict__ src,
                                                             int64_t* __restrict__ dst, uint64_t* __restrict__ av) {
  for (int32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int32_t u = rows[i];
    const int64_t e = aptr[u];
    const int32_t k = acol[e];
    len[i] = bptr[k + 1] - bptr[k];
    src[i] = bptr[k];
    dst[i] = cptr[u];
    av[i] = aval ? aval[e] : 1ull;
  }
}

// first copy row of every chunk: chunk c starts inside exactly one row (every copy row is non-empty)
__global__ void __launch_bounds__(kThreads) copy_chunks_kernel(const int64_t* __restrict__ pre, int32_t n,
                                                               int32_t* __restrict__ first) {
  for (int32_t r = blockIdx.x * kThreads + threadIdx.x; r < n; r += gridDim.x * kThreads) {
    const int64_t lo = pre[r], hi = pre[r + 1];
    for (int64_t c = (lo + kCopyChunk - 1) / kCopyChunk; c * kCopyChunk < hi; ++c) first[c] = r;
  }
}

template <bool BV>
__global__ void __launch_bounds__(kThreads) copy_tier_kernel(const int64_t* __restrict__ pre, int64_t total,
                                                             const int32_t* __restrict__ first,
                                                             const int64_t* __restrict__ src,
                                                             const int64_t* __restrict__ dst,
                                                             const uint64_t* __restrict__ av,
                                                             const int32_t* __restrict__ bcol,
                                                             const uint64_t* __restrict__ bval,
                                                             int32_t* __restrict__ ccol, uint64_t* __restrict__ cval) {
  const int64_t nch = (total + kCopyChunk - 1) / kCopyChunk;
  for (int64_t c = blockIdx.x; c < nch; c += gridDim.x) {
    int32_t r = first[c];
    int64_t lo = pre[r], hi = pre[r + 1];
    int64_t s0 = src[r] - lo, d0 = dst[r] - lo;  // element i of the run: B[s0 + i] -> C[d0 + i]
    uint64_t a = av[r];
    const int64_t end = min((c + 1) * kCopyChunk, total);
    for (int64_t i = c * kCopyChunk + threadIdx.x; i < end; i += kThreads) {
      while (i >= hi) {  // the next copy row of the chunk
        ++r;
        lo = hi;
        hi = pre[r + 1];
        s0 = src[r] - lo;
        d0 = dst[r] - lo;
        a = av[r];
      }
      const int32_t col = bcol[s0 + i];
      const uint64_t v = BV ? a * bval[s0 + i] : a;
      __builtin_nontemporal_store(col, &ccol[d0 + i]);
      __builtin_nontemporal_store(v, &cval[d0 + i]);
    }
  }
}

// Copy rows of at most kCopyRowMax elements: a wave per row, eight loads in flight per lane
constexpr int64_t kCopyRowMax = 8192;
template <bool AV, bool BV>
__global__ void __launch_bounds__(kThreads) copy_rows_kernel(const int32_t* __restrict__ rows, int32_t n,
                                                             const int64_t* __restrict__ aptr,
                                                             const int32_t* __restrict__ acol,
                                                             const uint64_t* __restrict__ aval,
                                                             const int64_t* __restrict__ bptr,
                                                             const int32_t* __restrict__ bcol,
                                                             const uint64_t* __restrict__ bval,
                                                             const int64_t* __restrict__ cptr,
                                                             int32_t* __restrict__ ccol, uint64_t* __restrict__ cval) {
  constexpr int kIn = 8;
  const int lane = lane_id();
  for (int64_t i = (int64_t)blockIdx.x * (kThreads / kWave) + wave_id(); i < n;
       i += (int64_t)gridDim.x * (kThreads / kWave)) {
    const int32_t u = rows[i];
    const int64_t e = aptr[u];
    const int32_t k = acol[e];
    const uint64_t a = AV ? aval[e] : 1ull;
    const int64_t s0 = bptr[k], len = bptr[k + 1] - s0, d0 = cptr[u];
    for (int64_t x0 = 0; x0 < len; x0 += kIn * kWave) {
      int32_t c[kIn];
      uint64_t v[kIn];
#pragma unroll
      for (int q = 0; q < kIn; ++q) {
        const int64_t x = x0 + q * kWave + lane;
        if (x < len) {
          c[q] = bcol[s0 + x];
          v[q] = BV ? bval[s0 + x] : 1ull;
        }
      }
#pragma unroll
      for (int q = 0; q < kIn; ++q) {
        const int64_t x = x0 + q * kWave + lane;
        if (x < len) {
          __builtin_nontemporal_store(c[q], &ccol[d0 + x]);
          __builtin_nontemporal_store(BV ? a * v[q] : a, &cval[d0 + x]);
        }
      }
    }
  }
}

// per-tier accounting of one pass: rows, A entries, products, output entries
struct TierCounts {
  int32_t rows[kNumBins] = {};
  int64_t a_nnz[kNumBins] = {};
  int64_t products[kNumBins] = {};
  int64_t c_nnz[kNumBins] = {};
  int64_t max_ub[kNumBins] = {};  // largest product count of a row
  int64_t max_d[kNumBins] = {};   // largest row of A (entries)
};

// scratch of the numeric bitmap/window split (device)
struct SplitScratch {
  int32_t* narrow = nullptr;
  int32_t* wide = nullptr;
  int32_t* counts = nullptr;          // 3: narrow rows, wide rows, widest narrow row
  unsigned long long* sums = nullptr;  // 6
  int32_t* bnd = nullptr;             // window bound table of B (built on first use)
  longlong2* nseg = nullptr;          // segment tables of B per numeric window and per symbolic
  longlong2* sseg = nullptr;          // window (built on first use; null above kSegMaxBytes)
  bool nseg_tried = false, sseg_tried = false;
  int32_t* wcnt = nullptr;            // columns per (row, window) from the symbolic pass, or null
                                      // (then the window tier runs one row at a time)
  WindowLayout wl;
  // with 4-byte windows: per row of A, 1 when its counts may reach 2^32 (those rows run in a
  // second launch with 8-byte accumulators over the same windows), or null when every row fits
  const uint8_t* big = nullptr;
  bool staged = false;  // split counts already copied to the host stage and synchronised
  struct SpgemmCore* owner = nullptr;
  DevAlloc* al = nullptr;             // the owner's allocator (per-launch task metadata)
  void build_bounds(const Csr& B, hipStream_t st);
  // segment table of G numeric windows per group (null when it would exceed kSegMaxBytes)
  longlong2* build_segs(const Csr& B, int32_t G, hipStream_t st);
};

// algorithmic bytes of a pass over one tier: A and C row pointers, A columns (+values) with
// the B row-pointer pair per A entry, every product's B column (+value), every output entry
// (column + value); the symbolic pass writes one int64 width per row instead of C
double tier_bytes(bool numeric, const TierCounts& tc, int t, bool av, bool bv) {
  const double r = tc.rows[t], d = (double)tc.a_nnz[t], p = (double)tc.products[t], w = (double)tc.c_nnz[t];
  if (!numeric) return 16.0 * (r + 1) + 20.0 * d + 4.0 * p;
  return 16.0 * (r + 1) + (20.0 + (av ? 8 : 0)) * d + (4.0 + (bv ? 8 : 0)) * p + 12.0 * w;
}

void launch_split(const int32_t* bins, int32_t m, int32_t nB, int32_t nH, const Csr& A, const int64_t* S,
                  const int64_t* cptr, const BitmapLayout& lay, const SpgemmOpts& o, SplitScratch* split,
                  hipStream_t st);

// widest row that keeps the bitmap kernel's LDS value slots: none when the product spans only
// a few windows (the window kernel's dense accumulator then beats rank lookups per product)
inline int64_t split_vcap(const BitmapLayout& lay, const SplitScratch* split, const SpgemmOpts& o) {
  const int64_t nwin = split->wl.nwin;
  return (nwin >= 2 && nwin <= o.window_all_nwin) ? 0 : lay.vcap;
}

template <bool NUMERIC, bool AV, bool BV>
void launch_tiers(const Csr& A, const Csr& B, const int64_t* S, const int32_t* bins, const TierCounts& tc,
                  int64_t* row_nnz, const int64_t* cptr, int32_t* ccol, uint64_t* cval, const SpgemmOpts& o,
                  const BitmapLayout& lay, SplitScratch* split, int32_t* queue, hipStream_t st, KernelTimer* tm) {
  const int32_t m = A.nrows;
  const int32_t nW = tc.rows[kWaveTier], nS = tc.rows[kSortTier], nB = tc.rows[kBitmapTier],
                nH = tc.rows[kHeavyTier];
  KernelTimer dummy;
  KernelTimer& T = tm ? *tm : dummy;
  if (nW > 0) {
    const unsigned g = (unsigned)ceil_div(nW, kThreads / kWave);
    const int tok = T.start(NUMERIC ? "spgemm.numeric.wave" : "spgemm.symbolic.wave", st);
    wave_tier_kernel<NUMERIC, AV, BV><<<g, kThreads, 0, st>>>(A.ptr, A.col, A.val, B.ptr, B.col, B.val, S,
                                                             bins + (size_t)kWaveTier * m, nW, row_nnz, cptr, ccol,
                                                             cval);
    SMCM_LAUNCH_CHECK();
    T.stop(tok, st, tier_bytes(NUMERIC, tc, kWaveTier, AV, BV));
  }
  const int32_t nC = tc.rows[kCopyTier];
  if (nC > 0) {
    const int32_t* crow = bins + (size_t)kCopyTier * m;
    const unsigned g = (unsigned)std::min<int64_t>(ceil_div(nC, kThreads), 65536);
    if constexpr (!NUMERIC) {
      const int tok = T.start("spgemm.symbolic.copy", st);
      copy_width_kernel<<<g, kThreads, 0, st>>>(crow, nC, A.ptr, S, row_nnz);
      SMCM_LAUNCH_CHECK();
      T.stop(tok, st, 16.0 * (nC + 1) + 20.0 * (double)tc.a_nnz[kCopyTier]);
    } else if (tc.max_ub[kCopyTier] <= kCopyRowMax) {
      const int tok = T.start("spgemm.numeric.copy", st);
      copy_rows_kernel<AV, BV><<<(unsigned)std::min<int64_t>(ceil_div(nC, kThreads / kWave), 65536), kThreads, 0, st>>>(
          crow, nC, A.ptr, A.col, A.val, B.ptr, B.col, B.val, cptr, ccol, cval);
      SMCM_LAUNCH_CHECK();
      T.stop(tok, st, tier_bytes(true, tc, kCopyTier, AV, BV));
    } else {
      // rows of any length: elements laid out flat over the copy rows, fixed-size chunks
      DevAlloc& al = *split->al;
      const int64_t total = tc.products[kCopyTier];
      DevBuf<int64_t> len(al, nC), pre(al, (int64_t)nC + 1), src(al, nC), dst(al, nC);
      DevBuf<uint64_t> av(al, nC);
      const int64_t nch = ceil_div(total, kCopyChunk);
      DevBuf<int32_t> first(al, std::max<int64_t>(nch, 1));
      const int tok = T.start("spgemm.numeric.copy", st);
      copy_prep_kernel<<<g, kThreads, 0, st>>>(crow, nC, A.ptr, A.col, A.val, B.ptr, cptr, len.p, src.p, dst.p, av.p);
      SMCM_LAUNCH_CHECK();
      scan_exclusive_i64(len.p, pre.p, nC, al);
      copy_chunks_kernel<<<g, kThreads, 0, st>>>(pre.p, nC, first.p);
      SMCM_LAUNCH_CHECK();
      const unsigned gc = (unsigned)std::min<int64_t>(std::max<int64_t>(nch, 1), (int64_t)o.cus * 8);
      copy_tier_kernel<BV><<<gc, kThreads, 0, st>>>(pre.p, total, first.p, src.p, dst.p, av.p, B.col, B.val, ccol, cval);
      SMCM_LAUNCH_CHECK();
      T.stop(tok, st, tier_bytes(true, tc, kCopyTier, AV, BV));
    }
  }
  if (nS > 0 && wave_bitmap_for(o, B.ncols)) {
    const WaveBitmapLayout wl = WaveBitmapLayout::make(B.ncols, (int32_t)o.wave_bitmap_cap, NUMERIC);
    const unsigned g = (unsigned)ceil_div(nS, kThreads / kWave);
    const int tok = T.start(NUMERIC ? "spgemm.numeric.wavebitmap" : "spgemm.symbolic.wavebitmap", st);
    wave_bitmap_kernel<NUMERIC, AV, BV><<<g, kThreads, wl.per_wave * (kThreads / kWave), st>>>(
        A.ptr, A.col, A.val, B.ptr, B.col, B.val, bins + (size_t)kSortTier * m, nS, wl, row_nnz, cptr, ccol, cval);
    SMCM_LAUNCH_CHECK();
    T.stop(tok, st, tier_bytes(NUMERIC, tc, kSortTier, AV, BV));
  } else if (nS > 0) {
    int cap = 64;
    while (cap < o.m_max) cap <<= 1;
    const size_t lds = (size_t)cap * (NUMERIC ? 12 : 4);
    const unsigned g = (unsigned)std::min<int64_t>(nS, (int64_t)o.cus * 16);
    const int tok = T.start(NUMERIC ? "spgemm.numeric.sort" : "spgemm.symbolic.sort", st);
    sort_tier_kernel<NUMERIC, AV, BV><<<g, kThreads, lds, st>>>(A.ptr, A.col, A.val, B.ptr, B.col, B.val, S,
                                                               bins + (size_t)kSortTier * m, nS, cap, row_nnz, cptr,
                                                               ccol, cval);
    SMCM_LAUNCH_CHECK();
    T.stop(tok, st, tier_bytes(NUMERIC, tc, kSortTier, AV, BV));
  }
  // heavy rows: window-major symbolic pass (needs the per-(row, window) count table)
  const bool sym_window = !NUMERIC && split->wcnt && nH > 0 && o.window_major;
  if (sym_window) {
    if (!split->bnd) split->build_bounds(B, st);
    const int32_t nwin = (int32_t)split->wl.nwin;
    DevBuf<SymRow> info(*split->al, nH);
    sym_prep_kernel<<<(unsigned)std::min<int64_t>(ceil_div(nH, kThreads), 65536), kThreads, 0, st>>>(
        bins + (size_t)kHeavyTier * m, nH, A.ptr, info.p);
    SMCM_LAUNCH_CHECK();
    fill_i32(queue, 1, 0, st);
    int sub = 0;
    while (((int64_t)split->wl.w << (sub + 1)) <= o.sym_window_cols && (1 << (sub + 1)) <= nwin) ++sub;
    const size_t lds = ((size_t)split->wl.w << sub) / 8;
    // rows whose products hit every column many times (dense chains) merge lanes per word
    const double per_col = (double)tc.products[kHeavyTier] / std::max(nH, 1) / std::max(B.ncols, 1);
    const bool merge = o.sym_merge == 1 || (o.sym_merge == 2 && per_col >= 8.0);
    auto skern = merge ? window_symbolic_kernel<true> : window_symbolic_kernel<false>;
    const unsigned g = persistent_grid(o, occupancy((const void*)skern, o.d_threads, lds));
    TierCounts heavy;
    heavy.rows[0] = nH;
    heavy.a_nnz[0] = tc.a_nnz[kHeavyTier];
    heavy.products[0] = tc.products[kHeavyTier];
    const int tok = T.start("spgemm.symbolic.window", st);
    if (!split->sseg_tried) {
      split->sseg = split->build_segs(B, 1 << sub, st);
      split->sseg_tried = true;
    }
    skern<<<g, o.d_threads, lds, st>>>(A.col, B.ptr, B.col, info.p, nH, split->wl.log2w, sub, nwin, split->bnd,
                                       split->sseg, B.nrows, queue, split->wcnt, row_nnz);
    SMCM_LAUNCH_CHECK();
    T.stop(tok, st, tier_bytes(false, heavy, 0, AV, BV));
  }
  const int32_t nHb = sym_window ? 0 : nH;  // heavy rows left to the bitmap kernel
  if (nB + nHb > 0 && !NUMERIC) {
    fill_i32(queue, 1, 0, st);
    const size_t lds = lay.bytes_sym();
    const int per_cu = std::max<int64_t>(1, std::min<int64_t>(2048 / o.d_threads, o.lds_bytes / (int64_t)(lds + 12 * 1024)));
    const unsigned g = (unsigned)std::min<int64_t>(nB + nHb, (int64_t)o.cus * per_cu);
    TierCounts both;
    both.rows[0] = nB + nHb;
    both.a_nnz[0] = tc.a_nnz[kBitmapTier] + (nHb ? tc.a_nnz[kHeavyTier] : 0);
    both.products[0] = tc.products[kBitmapTier] + (nHb ? tc.products[kHeavyTier] : 0);
    const int tok = T.start("spgemm.symbolic.bitmap", st);
    bitmap_tier_kernel<NUMERIC, AV, BV><<<g, o.d_threads, lds, st>>>(
        A.ptr, A.col, A.val, B.ptr, B.col, B.val, bins + (size_t)kHeavyTier * m, nHb, bins + (size_t)kBitmapTier * m,
        nB, B.ncols, lay.nw, lay.vcap, queue, row_nnz, cptr, ccol, cval, split->wl.log2w, (int32_t)split->wl.nwin,
        split->wcnt);
    SMCM_LAUNCH_CHECK();
    T.stop(tok, st, tier_bytes(false, both, 0, AV, BV));
  }
  if (nB + nH > 0 && NUMERIC) {
    // split by exact width: narrow rows keep the LDS bitmap + rank kernel, wide rows use windows
    // (launched and read back with the output size when the whole product is one block)
    if (!split->staged) {
      launch_split(bins, m, nB, nH, A, S, cptr, lay, o, split, st);
      SMCM_HIP(hipStreamSynchronize(st));
    }
    split->staged = false;
    int32_t cnt[3];
    unsigned long long sums[6];
    {
      const unsigned char* hs = host_stage() + kStageSplit;
      memcpy(cnt, hs, sizeof(cnt));
      memcpy(sums, hs + 64, sizeof(sums));
    }
    TierCounts cls;
    for (int c = 0; c < 2; ++c) {
      cls.rows[c] = cnt[c];
      cls.a_nnz[c] = (int64_t)sums[3 * c];
      cls.products[c] = (int64_t)sums[3 * c + 1];
      cls.c_nnz[c] = (int64_t)sums[3 * c + 2];
    }
    if (cnt[0] > 0) {
      fill_i32(queue, 1, 0, st);
      // value slots for the widest narrow row only: a smaller LDS footprint, more workgroups per CU
      BitmapLayout nl = lay;
      nl.vcap = std::min<int64_t>(lay.vcap, std::max<int64_t>(64, ((int64_t)cnt[2] + 63) / 64 * 64));
      const size_t lds = nl.bytes_num();
      const int per_cu = std::max<int64_t>(1, std::min<int64_t>(2048 / o.d_threads, o.lds_bytes / (int64_t)(lds + 12 * 1024)));
      const unsigned g = (unsigned)std::min<int64_t>(cnt[0], (int64_t)o.cus * per_cu);
      const int tok = T.start("spgemm.numeric.bitmap", st);
      bitmap_tier_kernel<NUMERIC, AV, BV><<<g, o.d_threads, lds, st>>>(A.ptr, A.col, A.val, B.ptr, B.col, B.val,
                                                                       nullptr, 0, split->narrow, cnt[0], B.ncols,
                                                                       nl.nw, nl.vcap, queue, row_nnz, cptr, ccol,
                                                                       cval);
      SMCM_LAUNCH_CHECK();
      T.stop(tok, st, tier_bytes(true, cls, 0, AV, BV));
    }
    // wide rows with few products per window: one pass over a column bitmap of the whole row
    // (values accumulated in the output by device atomics) beats many nearly empty windows
    const double per_task = cnt[1] > 0 ? (double)cls.products[1] / ((double)cnt[1] * (double)split->wl.nwin) : 0.0;
    if (cnt[1] > 0 && per_task < (double)o.window_min_products) {
      fill_i32(queue, 1, 0, st);
      BitmapLayout gl = lay;
      gl.vcap = 0;
      const size_t lds = gl.bytes_num();
      const int per_cu = std::max<int64_t>(1, std::min<int64_t>(2048 / o.d_threads, o.lds_bytes / (int64_t)(lds + 12 * 1024)));
      const unsigned g = (unsigned)std::min<int64_t>(cnt[1], (int64_t)o.cus * per_cu);
      const int tok = T.start("spgemm.numeric.bitmap.hbm", st);
      bitmap_tier_kernel<NUMERIC, AV, BV><<<g, o.d_threads, lds, st>>>(A.ptr, A.col, A.val, B.ptr, B.col, B.val,
                                                                       nullptr, 0, split->wide, cnt[1], B.ncols,
                                                                       gl.nw, gl.vcap, queue, row_nnz, cptr, ccol,
                                                                       cval);
      SMCM_LAUNCH_CHECK();
      T.stop(tok, st, tier_bytes(true, cls, 1, AV, BV));
    } else if (cnt[1] > 0) {
      if (!split->bnd) split->build_bounds(B, st);
      fill_i32(queue, 1, 0, st);
      const size_t lds = split->wl.bytes();
      // workgroups per CU: bounded by LDS and by 32 resident waves per CU
      const int64_t by_lds = o.lds_bytes / (int64_t)(lds + kStaticLds);
      const int64_t by_waves = 2048 / o.w_threads;
      const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(by_lds, by_waves));
      const unsigned g = (unsigned)std::min<int64_t>(cnt[1], (int64_t)o.cus * per_cu);
      const int tok = T.start("spgemm.numeric.window", st);
      if (split->wcnt) {
        const int32_t nwin = (int32_t)split->wl.nwin;
        // presence bits pay off when windows are sparsely filled (the compaction then touches few
        // slots); dense windows are cheaper with values alone (no per-product bit update)
        const double fill = (double)cls.c_nnz[1] / ((double)cnt[1] * nwin * (double)split->wl.w);
        const bool pres = o.window_presence == 1 || (o.window_presence == 2 && fill < 0.125);
        // the compaction's presence pass reads a wave's run of words with one lane each
        if ((split->wl.w / 64 + o.w_threads / kWave - 1) / (o.w_threads / kWave) > kWave)
          throw HipError(-22, "window tier: more than 64 words per wave in a window");
        if (!split->nseg_tried) {
          split->nseg = split->build_segs(B, 1, st);
          split->nseg_tried = true;
        }
        // one persistent launch over `rows` with accumulators of vbytes bytes
        auto run_rows = [&](const int32_t* rows, int32_t nrows, int vbytes) {
          DevBuf<WinRow> info(*split->al, nrows);
          DevBuf<int32_t> pre(*split->al, (int64_t)nrows * (nwin + 1));
          window_prep_kernel<<<(unsigned)std::min<int64_t>(ceil_div(nrows, kThreads / kWave), 65536), kThreads, 0,
                               st>>>(rows, nrows, A.ptr, cptr, split->wcnt, nwin, info.p, pre.p);
          SMCM_LAUNCH_CHECK();
          WindowLayout wl = split->wl;
          wl.vbytes = vbytes;
          const size_t lb = wl.bytes();
          auto kern = vbytes == 4
                          ? (pres ? window_major_kernel<AV, BV, true, uint32_t> : window_major_kernel<AV, BV, false, uint32_t>)
                          : (pres ? window_major_kernel<AV, BV, true> : window_major_kernel<AV, BV, false>);
          const unsigned gm = persistent_grid(o, occupancy((const void*)kern, o.w_threads, lb));
          fill_i32(queue, 1, 0, st);
          kern<<<gm, o.w_threads, lb, st>>>(A.col, A.val, B.ptr, B.col, B.val, info.p, pre.p, nrows, B.ncols,
                                            split->wl.log2w, nwin, split->bnd, split->nseg, B.nrows, queue, ccol,
                                            cval);
          SMCM_LAUNCH_CHECK();
        };
        if (split->wl.vbytes == 4 && split->big) {
          // rows whose counts may reach 2^32 take 8-byte accumulators over the same windows
          DevBuf<int32_t> lists(*split->al, 2 * (int64_t)cnt[1] + 16);
          int32_t* c2 = lists.p + 2 * (int64_t)cnt[1];
          fill_i32(c2, 2, 0, st);
          split_big_kernel<<<(unsigned)std::min<int64_t>(ceil_div(cnt[1], kThreads), 65536), kThreads, 0, st>>>(
              split->wide, cnt[1], split->big, lists.p, lists.p + cnt[1], c2);
          SMCM_LAUNCH_CHECK();
          int32_t hc[2] = {0, 0};
          SMCM_HIP(hipMemcpyAsync(hc, c2, sizeof(hc), hipMemcpyDeviceToHost, st));
          SMCM_HIP(hipStreamSynchronize(st));
          if (hc[0] < 0 || hc[1] < 0 || hc[0] + hc[1] != cnt[1])
            throw HipError(-22, "window tier: row split lost rows");
          if (hc[0] > 0) run_rows(lists.p, hc[0], 4);
          if (hc[1] > 0) run_rows(lists.p + cnt[1], hc[1], 8);
        } else {
          run_rows(split->wide, cnt[1], split->wl.vbytes);
        }
      } else {
        window_tier_kernel<AV, BV><<<g, o.w_threads, lds, st>>>(
            A.ptr, A.col, A.val, B.ptr, B.col, B.val, split->wide, cnt[1], B.ncols, split->wl.log2w,
            (int32_t)split->wl.nwin, split->bnd, B.nrows, queue, cptr, ccol, cval);
      }
      SMCM_LAUNCH_CHECK();
      T.stop(tok, st, tier_bytes(true, cls, 1, AV, BV));
    }
  }
}

// classify the bitmap-tier rows by exact width and copy the split counts to the host stage
// (no synchronisation)
void launch_split(const int32_t* bins, int32_t m, int32_t nB, int32_t nH, const Csr& A, const int64_t* S,
                  const int64_t* cptr, const BitmapLayout& lay, const SpgemmOpts& o, SplitScratch* split,
                  hipStream_t st) {
  fill_i32(split->counts, 3, 0, st);
  fill_i64((int64_t*)split->sums, 6, 0, st);
  split_wide_kernel<<<(unsigned)ceil_div(nB + nH, kThreads), kThreads, 0, st>>>(
      bins + (size_t)kHeavyTier * m, nH, bins + (size_t)kBitmapTier * m, nB, A.ptr, S, cptr,
      split_vcap(lay, split, o), split->narrow, split->wide, split->counts, split->sums);
  SMCM_LAUNCH_CHECK();
  // split->sums sits 64 bytes after split->counts (SpgemmCore's counter block): one copy
  SMCM_HIP(hipMemcpyAsync(host_stage() + kStageSplit, split->counts, 64 + 6 * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, st));
}

template <bool NUMERIC>
void dispatch(const Csr& A, const Csr& B, const int64_t* S, const int32_t* bins, const TierCounts& tc,
              int64_t* row_nnz, const int64_t* cptr, int32_t* ccol, uint64_t* cval, const SpgemmOpts& o,
              const BitmapLayout& lay, SplitScratch* split, int32_t* queue, hipStream_t st, KernelTimer* tm) {
  const bool av = A.val != nullptr, bv = B.val != nullptr;
  if (av && bv)
    launch_tiers<NUMERIC, true, true>(A, B, S, bins, tc, row_nnz, cptr, ccol, cval, o, lay, split, queue, st, tm);
  else if (av)
    launch_tiers<NUMERIC, true, false>(A, B, S, bins, tc, row_nnz, cptr, ccol, cval, o, lay, split, queue, st, tm);
  else if (bv)
    launch_tiers<NUMERIC, false, true>(A, B, S, bins, tc, row_nnz, cptr, ccol, cval, o, lay, split, queue, st, tm);
  else
    launch_tiers<NUMERIC, false, false>(A, B, S, bins, tc, row_nnz, cptr, ccol, cval, o, lay, split, queue, st,
                                        tm);
}

// raises the dynamic-LDS limit of one kernel; a failure names the kernel and the size
void lds_attr(const void* kernel, const char* name, size_t bytes) {
  const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw HipError((int)e, std::string("hipFuncSetAttribute(") + name + ", " + std::to_string(bytes) +
                               " bytes of LDS): " + hipGetErrorString(e));
  }
}

template <bool NUMERIC, bool AV, bool BV>
void set_lds_attr(size_t bytes) {
  lds_attr((const void*)bitmap_tier_kernel<NUMERIC, AV, BV>, "bitmap_tier_kernel", bytes);
  if constexpr (NUMERIC) {
    lds_attr((const void*)window_tier_kernel<AV, BV>, "window_tier_kernel", bytes);
    lds_attr((const void*)window_major_kernel<AV, BV, false>, "window_major_kernel", bytes);
    lds_attr((const void*)window_major_kernel<AV, BV, true>, "window_major_kernel", bytes);
    lds_attr((const void*)window_major_kernel<AV, BV, false, uint32_t>, "window_major_kernel", bytes);
    lds_attr((const void*)window_major_kernel<AV, BV, true, uint32_t>, "window_major_kernel", bytes);
  }
}

// every tiered kernel's limit raised to `bytes`; recorded as done only once all calls succeeded,
// so a failed call is retried by the next product
void set_all_lds_attrs(size_t bytes) {
  static std::mutex mu;  // server threads run products concurrently
  static size_t done = 0;
  std::lock_guard<std::mutex> lk(mu);
  if (bytes <= done) return;
  set_lds_attr<false, false, false>(bytes);
  set_lds_attr<false, true, false>(bytes);
  set_lds_attr<false, false, true>(bytes);
  set_lds_attr<false, true, true>(bytes);
  set_lds_attr<true, false, false>(bytes);
  set_lds_attr<true, true, false>(bytes);
  set_lds_attr<true, false, true>(bytes);
  set_lds_attr<true, true, true>(bytes);
  done = bytes;
}

BitmapLayout choose_layout(int64_t ncols, const SpgemmOpts& o) {
  // LDS budget for the dynamic part (the traversal's static arrays and scan scratch take the rest)
  const int64_t budget = o.lds_bytes - kStaticLds;
  BitmapLayout l;
  const int64_t want = std::max<int64_t>(128, (ncols + 127) / 128 * 128);
  // numeric: nw/8 + summary + nw/16 + 8*vcap <= budget with vcap >= 2048
  int64_t max_nw = (budget - 8 * 2048) * 16 / 3;
  max_nw = std::max<int64_t>(128, max_nw / 128 * 128);
  l.nw = std::min(want, max_nw);
  l.vcap = 2048;
  while (l.nw > 128 && (int64_t)l.bytes_num() > budget) l.nw -= 128;
  l.vcap = 0;
  l.vcap = (budget - (int64_t)l.bytes_num()) / 8;
  if (l.vcap > (1 << 16)) l.vcap = 1 << 16;
  if (l.vcap < 0) l.vcap = 0;
  if (o.bitmap_vcap >= 0) l.vcap = std::min(l.vcap, o.bitmap_vcap);
  return l;
}

// widest window whose accumulators (vbytes each) fit the LDS budget; window_cols caps the width
// of 8-byte windows, and of 4-byte ones at twice that (the same LDS)
WindowLayout choose_window(int64_t ncols, const SpgemmOpts& o, int vbytes = 8) {
  const int64_t budget = o.lds_bytes - kStaticLds;
  WindowLayout l;
  l.w = 64;
  l.log2w = 6;
  l.vbytes = vbytes;
  while (l.w < (1 << 20)) {
    WindowLayout n2;
    n2.w = l.w * 2;
    n2.vbytes = vbytes;
    if ((int64_t)n2.bytes() > budget || n2.w > std::max<int64_t>(64, 2 * ncols)) break;
    l.w = n2.w;
    l.log2w += 1;
  }
  const int64_t cap = o.window_cols > 0 ? o.window_cols * (8 / vbytes) : 0;
  while (cap > 0 && l.w > cap && l.w > 64) {
    l.w >>= 1;
    l.log2w -= 1;
  }
  l.nwin = std::max<int64_t>(1, (ncols + l.w - 1) / l.w);
  return l;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Analysis and binning of one product, reusable for row blocks of the same product.
struct SpgemmCore {
  const Csr& A;
  const Csr& B;
  const double* est_r;
  int est_stride;
  const SpgemmOpts& o;
  DevAlloc& al;
  KernelTimer* tm;
  DevBuf<int64_t> S;
  DevBuf<int32_t> bins;
  // counters of the pass in one block (laid out like the host stage's kStageBin / kStageSplit
  // sections, so each group is read back with one copy): tier row counts, queue, tier sums,
  // split counts, split sums
  DevBuf<unsigned char> meta;
  int32_t *counts = nullptr, *queue = nullptr;
  unsigned long long* sums = nullptr;
  DevBuf<int32_t> split_rows, bnd, wcnt;
  DevBuf<longlong2> segbuf[2];
  DevBuf<unsigned long long> rmax;  // largest value per row of B (+ a counter): mixed widths
  DevBuf<uint8_t> big_rows;         // per row of A: counts may reach 2^32
  SplitScratch split;
  TierCounts tc;
  BitmapLayout lay;
  int64_t products = 0;

  SpgemmCore(const Csr& a, const Csr& b, const double* er, int es, const SpgemmOpts& opts, DevAlloc& alloc,
             KernelTimer* timer)
      : A(a), B(b), est_r(er), est_stride(es), o(opts), al(alloc), tm(timer) {
    hipStream_t st = al.stream;
    DevBuf<int64_t> len(al, A.nnz);
    S = DevBuf<int64_t>(al, A.nnz + 1);
    if (A.nnz > 0) {
      const int tok = tm ? tm->start("spgemm.analysis", st) : -1;
      entry_len_kernel<<<(unsigned)std::min<int64_t>(ceil_div(A.nnz, kThreads), 65536), kThreads, 0, st>>>(
          A.col, A.nnz, B.ptr, len.p);
      SMCM_LAUNCH_CHECK();
      if (tm) tm->stop(tok, st, 16.0 * (double)A.nnz + 4.0 * (double)A.nnz);
    }
    scan_exclusive_i64(len.p, S.p, A.nnz, al);
    bins = DevBuf<int32_t>(al, (int64_t)kNumBins * (A.nrows > 0 ? A.nrows : 1));
    meta = DevBuf<unsigned char>(al, kMetaBytes);
    counts = (int32_t*)(meta.p + kMetaCounts);
    queue = (int32_t*)(meta.p + kMetaQueue);
    sums = (unsigned long long*)(meta.p + kMetaSums);
    lay = choose_layout(B.ncols, o);
    split.wl = choose_window(B.ncols, o);
    set_all_lds_attrs(std::max({lay.bytes_num(), lay.bytes_sym(), split.wl.bytes()}));
    split_rows = DevBuf<int32_t>(al, 2 * (int64_t)(A.nrows > 0 ? A.nrows : 1));

    split.narrow = split_rows.p;
    split.wide = split_rows.p + (A.nrows > 0 ? A.nrows : 1);
    split.counts = (int32_t*)(meta.p + kMetaSplit);
    split.sums = (unsigned long long*)(meta.p + kMetaSplit + 64);
    split.owner = this;
    split.al = &al;
    // per-(row, window) column counts for the window-major window tier; the task index
    // (rows x windows) must stay an int32 queue value
    const int64_t cells = (int64_t)A.nrows * split.wl.nwin;
    if (o.window_major && A.nrows > 0 && cells * 4 <= kWindowCountBytes && cells < (int64_t)INT32_MAX - (1 << 24)) {
      wcnt = DevBuf<int32_t>(al, cells);
      split.wcnt = wcnt.p;
    }
  }

  // bin rows [r0, r1) (bins hold absolute row ids); results in tc (synchronises)
  void bin(int32_t r0, int32_t r1) {
    hipStream_t st = al.stream;
    fill_i32(counts, kNumBins + 1, 0, st);
    fill_i64((int64_t*)sums, 5 * kNumBins + 2, 0, st);
    // the operands' largest values (a bound on every count, for 32-bit window accumulators):
    // scanned with the first binning of the whole product when the operands are small enough
    const bool whole = r0 == 0 && r1 == A.nrows;
    const bool scan_a = whole && A.val && A.nnz > 0 && A.nnz <= kMaxValueScan;
    const bool scan_b = whole && B.val && B.nnz > 0 && B.nnz <= kMaxValueScan;
    if (scan_a) {
      row_sum_max_kernel<<<(unsigned)std::min<int64_t>(ceil_div(A.nrows, kThreads / kWave), 1024), kThreads, 0,
                           st>>>(A.ptr, A.val, A.nrows, sums + 5 * kNumBins);
      SMCM_LAUNCH_CHECK();
    }
    if (scan_b) {
      max_u64_kernel<<<(unsigned)std::min<int64_t>(ceil_div(B.nnz, kThreads), 1024), kThreads, 0, st>>>(
          B.val, B.nnz, sums + 5 * kNumBins + 1);
      SMCM_LAUNCH_CHECK();
    }
    if (r1 > r0) {
      bin_rows_kernel<<<(unsigned)ceil_div(r1 - r0, (int64_t)kThreads * kBinRowsPerThread), kThreads, 0, st>>>(
          A.ptr, r0, r1, A.nrows, S.p, est_r, est_stride, (double)B.ncols, o.s_max, o.m_max, o.dense_div, o.heavy_ub,
          o.force_tier, wave_bitmap_for(o, B.ncols) ? (int)o.wave_bitmap_max_ub : 0,
          (B.distinct && o.copy_tier && o.force_tier < 0) ? 1 : 0, bins.p, counts, sums);
      SMCM_LAUNCH_CHECK();
    }
    unsigned char* hs = host_stage() + kStageBin;
    // tier row counts at +0, tier sums at +64 (largest product counts after the width sums) in
    // both blocks: one copy
    SMCM_HIP(hipMemcpyAsync(hs, meta.p + kMetaCounts, kMetaSums + sizeof(unsigned long long) * (5 * kNumBins + 2),
                            hipMemcpyDeviceToHost, st));
    SMCM_HIP(hipStreamSynchronize(st));
    unsigned long long sh[5 * kNumBins + 2];
    memcpy(tc.rows, hs, sizeof(int32_t) * kNumBins);
    memcpy(sh, hs + 64, sizeof(sh));
    for (int t = 0; t < kNumBins; ++t) {
      tc.a_nnz[t] = (int64_t)sh[2 * t];
      tc.products[t] = (int64_t)sh[2 * t + 1];
      tc.c_nnz[t] = 0;
      tc.max_ub[t] = (int64_t)sh[3 * kNumBins + t];
      tc.max_d[t] = (int64_t)sh[4 * kNumBins + t];
    }
    // every row of A binned: the products of all tiers are the product's total (S[nnz(A)])
    if (whole) {
      products = 0;
      for (int t = 0; t < kNumBins; ++t) products += tc.products[t];
      // 32-bit window accumulators when no count can reach 2^32: a count of row u is
      // sum_e A[u,e] * B[k_e, c] <= (row sum of A's values) x max(B); without values a row sums
      // to its length.  Unscanned operands bound nothing.
      int64_t widest = 0;
      for (int t = 0; t < kNumBins; ++t) widest = std::max(widest, tc.max_d[t]);
      const long double arow = A.val ? (scan_a ? (long double)sh[5 * kNumBins] : 0x1p64L) : (long double)widest;
      const long double bmax = B.val ? (scan_b ? (long double)sh[5 * kNumBins + 1] : 0x1p64L) : 1.0L;
      const bool window_rows = tc.rows[kBitmapTier] + tc.rows[kHeavyTier] > 0;
      if (window_rows && std::getenv("SMCM_TRACE"))
        fprintf(stderr, "[smcm] count bound %dx%d * %dx%d: row sum of A %.4Lg (%s) x largest B %.4Lg (%s)\n", A.nrows,
                A.ncols, B.nrows, B.ncols, arow, A.val ? (scan_a ? "scanned" : "not scanned") : "row length", bmax,
                B.val ? (scan_b ? "scanned" : "not scanned") : "no values");
      if (o.window_u32 && split.wcnt && window_rows && arow * bmax < 4294967296.0L) {
        const WindowLayout w4 = choose_window(B.ncols, o, 4);
        if (w4.w > split.wl.w) {
          split.wl = w4;
          set_all_lds_attrs(split.wl.bytes());
        }
      } else if (o.window_u32 && split.wcnt && window_rows && products >= o.window_mixed_min) {
        mixed_widths();
      }
    }
  }

  // Per-row count bounds (sum over the row of A[u,e] x the largest value of B's row k_e): when
  // few rows can reach 2^32, the product takes 4-byte windows and those rows run in a second
  // window launch with 8-byte accumulators over the same windows (which must fit the LDS).
  void mixed_widths() {
    hipStream_t st = al.stream;
    const WindowLayout w4 = choose_window(B.ncols, o, 4);
    WindowLayout w8 = w4;
    w8.vbytes = 8;
    if (w4.w <= split.wl.w || (int64_t)w8.bytes() > o.lds_bytes - kStaticLds) return;
    rmax = DevBuf<unsigned long long>(al, (int64_t)B.nrows + 3);
    unsigned long long* nbig = rmax.p + B.nrows;  // big rows, their largest product count, their total
    fill_i64((int64_t*)nbig, 3, 0, st);
    if (B.nrows > 0) {
      row_max_kernel<<<(unsigned)std::min<int64_t>(ceil_div(B.nrows, kThreads / kWave), 65536), kThreads, 0, st>>>(
          B.ptr, B.val, B.nrows, rmax.p);
      SMCM_LAUNCH_CHECK();
    }
    big_rows = DevBuf<uint8_t>(al, A.nrows);
    row_bound_kernel<<<(unsigned)std::min<int64_t>(ceil_div(A.nrows, kThreads / kWave), 65536), kThreads, 0, st>>>(
        A.ptr, A.col, A.val, A.nrows, B.ptr, rmax.p, big_rows.p, nbig);
    SMCM_LAUNCH_CHECK();
    unsigned long long hs[3] = {0, 0, 0};
    SMCM_HIP(hipMemcpyAsync(host_stage() + kStageNnz, nbig, sizeof(hs), hipMemcpyDeviceToHost, st));
    SMCM_HIP(hipStreamSynchronize(st));
    memcpy(hs, host_stage() + kStageNnz, sizeof(hs));
    const unsigned long long hb = hs[0];
    // The second launch runs the big rows alone: it pays off while they are a small share of the
    // rows and of the products, and while none of its (row, window) tasks, spread over the row's
    // windows, outweighs a small part of what each CU does in the first launch (a hub row alone in
    // a launch of few tasks would leave most CUs idle).
    const double task_max = (double)hs[1] / (double)std::max<int64_t>(w4.nwin, 1);
    const double cu_work = (double)products / (double)(std::max(o.cus, 1) * 2);
    const bool take = hb * 8 <= (unsigned long long)A.nrows && (double)hs[2] * 8.0 <= (double)products &&
                      task_max * 4.0 <= cu_work;
    if (std::getenv("SMCM_TRACE"))
      fprintf(stderr,
              "[smcm] count bounds per row %dx%d * %dx%d: %llu of %d rows may reach 2^32 (%llu of %lld products, "
              "largest row %llu): %s\n",
              A.nrows, A.ncols, B.nrows, B.ncols, hb, A.nrows, hs[2], (long long)products, hs[1],
              take ? "mixed widths" : "8-byte windows");
    if (!take) return;
    split.wl = w4;
    split.big = hb ? big_rows.p : nullptr;
    set_all_lds_attrs(std::max(w4.bytes(), w8.bytes()));
  }

  // widths of the binned rows into row_nnz (other entries untouched).  The per-tier widths are
  // copied to the host stage without a synchronisation: read_symbolic() fills tc.c_nnz after
  // the caller's next stream synchronisation.
  // the symbolic pass as ORs of B's rows kept as a bitmap: B's bitmap is small and reading a
  // row's nnz(A row) bitmap rows costs less than its products (checked over the whole product)
  // A row's entries are split over one workgroup's four waves, so a row of many entries (a hub)
  // would be one long serial walk: such products keep the tiered pass.
  bool use_bbits() const {
    if (o.bbits_bytes <= 0 || o.force_tier >= 0 || A.nrows == 0 || B.ncols == 0) return false;
    const int64_t nw = ceil_div(B.ncols, 64);
    if ((int64_t)B.nrows * nw * 8 > o.bbits_bytes) return false;
    int64_t widest = 0;
    for (int t = 0; t < kNumBins; ++t) widest = std::max(widest, tc.max_d[t]);
    if (widest > kBbitsMaxEntries) return false;
    return (double)A.nnz * (double)nw <= 4.0 * (double)products;
  }
  void symbolic(int64_t* row_nnz) {
    hipStream_t st = al.stream;
    if (split.wcnt) fill_i32(split.wcnt, (int64_t)A.nrows * split.wl.nwin, 0, st);
    if (use_bbits()) {
      const int32_t nw = (int32_t)ceil_div(B.ncols, 64);
      DevBuf<unsigned long long> bits(al, (int64_t)B.nrows * nw + 1);
      KernelTimer dummy;
      KernelTimer& T = tm ? *tm : dummy;
      DevBuf<int32_t> heavy(al, A.nrows);
      const int tok = T.start("spgemm.symbolic.bbits", st);
      // the bitmap and, after it, the heavy-row counter: zeroed by one fill
      fill_i64((int64_t*)bits.p, (int64_t)B.nrows * nw + 1, 0, st);
      int32_t* nheavy = (int32_t*)(bits.p + (int64_t)B.nrows * nw);
      const int wpb = kThreads / kWave;
      if (B.nrows > 0) {
        bbits_kernel<<<(unsigned)std::min<int64_t>(ceil_div(B.nrows, wpb), 65536), kThreads, 0, st>>>(
            B.ptr, B.col, B.nrows, nw, bits.p);
        SMCM_LAUNCH_CHECK();
      }
      // rows of every tier but the empty one (their widths stay 0)
      RowLists lists{};
      int32_t total = 0;
      for (int t = 0; t < kNumBins; ++t) {
        lists.rows[t] = bins.p + (size_t)t * A.nrows;
        lists.start[t] = total;
        total += t == kEmpty ? 0 : tc.rows[t];
      }
      lists.start[kNumBins] = total;
      if (total > 0) {
        bbits_symbolic_kernel<<<(unsigned)std::min<int64_t>(ceil_div(total, wpb), 65536), kThreads, 0, st>>>(
            A.ptr, A.col, bits.p, nw, lists, split.wl.log2w, (int32_t)split.wl.nwin, row_nnz, split.wcnt, heavy.p,
            nheavy);
        SMCM_LAUNCH_CHECK();
      }
      bbits_heavy_kernel<<<(unsigned)(o.cus * 8), kThreads, 0, st>>>(A.ptr, A.col, bits.p, nw, split.wl.log2w,
                                                                       (int32_t)split.wl.nwin, heavy.p, nheavy,
                                                                       row_nnz, split.wcnt);
      SMCM_LAUNCH_CHECK();
      T.stop(tok, st, 16.0 * ((double)A.nrows + 1) + 20.0 * (double)A.nnz + 4.0 * (double)products);
    } else {
      dispatch<false>(A, B, S.p, bins.p, tc, row_nnz, nullptr, nullptr, nullptr, o, lay, &split, queue, st, tm);
    }
    unsigned long long* tn = sums + 2 * kNumBins;
    fill_i64((int64_t*)tn, kNumBins, 0, st);
    tier_nnz_kernel<<<dim3(64, kNumBins - 1 - kEmpty), kThreads, 0, st>>>(bins.p, A.nrows, counts, row_nnz, tn);
    SMCM_LAUNCH_CHECK();
    SMCM_HIP(hipMemcpyAsync(host_stage() + kStageTierNnz, tn, sizeof(unsigned long long) * kNumBins,
                            hipMemcpyDeviceToHost, st));
  }
  // whole-product numeric pass ahead: classify its bitmap-tier rows now (cptr = scanned widths),
  // so the split counts come back with the caller's next synchronisation
  void stage_split(const int64_t* cptr) {
    const int32_t nB = tc.rows[kBitmapTier], nH = tc.rows[kHeavyTier];
    if (nB + nH == 0) return;
    launch_split(bins.p, A.nrows, nB, nH, A, S.p, cptr, lay, o, &split, al.stream);
    split.staged = true;
  }
  void read_symbolic() {
    unsigned long long h[kNumBins];
    memcpy(h, host_stage() + kStageTierNnz, sizeof(h));
    for (int t = 0; t < kNumBins; ++t) tc.c_nnz[t] = (int64_t)h[t];
  }

  // numeric pass of the binned rows; cptr holds absolute offsets, ccol/cval are shifted so
  // that ccol[cptr[u]] is the first output of row u
  void numeric(const int64_t* cptr, int32_t* ccol, uint64_t* cval) {
    dispatch<true>(A, B, S.p, bins.p, tc, nullptr, cptr, ccol, cval, o, lay, &split, queue, al.stream, tm);
  }

  void add_stats(SpgemmStats* st) const {
    if (!st) return;
    st->rows[0] += tc.rows[kEmpty];
    st->rows[1] += tc.rows[kWaveTier] + tc.rows[kCopyTier];
    st->rows[2] += tc.rows[kSortTier];
    st->rows[3] += tc.rows[kBitmapTier] + tc.rows[kHeavyTier];
  }
};

longlong2* SplitScratch::build_segs(const Csr& B, int32_t G, hipStream_t st) {
  const int32_t nwin = (int32_t)wl.nwin;
  const int32_t ngroups = (nwin + G - 1) / G;
  const int64_t n = (int64_t)std::max(B.nrows, 1) * ngroups;
  if (n * (int64_t)sizeof(longlong2) > kSegMaxBytes) return nullptr;
  if (!bnd) build_bounds(B, st);
  DevBuf<longlong2>& buf = owner->segbuf[G == 1 ? 0 : 1];
  buf = DevBuf<longlong2>(*al, n);
  if (B.nrows > 0) {
    window_segs_kernel<<<(unsigned)std::min<int64_t>(ceil_div(n, kThreads), 65536), kThreads, 0, st>>>(
        B.ptr, bnd, B.nrows, nwin, G, ngroups, buf.p);
    SMCM_LAUNCH_CHECK();
  }
  return buf.p;
}

void SplitScratch::build_bounds(const Csr& B, hipStream_t st) {
  owner->bnd = DevBuf<int32_t>(owner->al, (int64_t)std::max(B.nrows, 1) * (wl.nwin + 1));
  bnd = owner->bnd.p;
  if (B.nrows > 0) {
    window_bounds_kernel<<<(unsigned)std::min<int64_t>(ceil_div(B.nrows, kThreads / kWave), 65536), kThreads, 0, st>>>(
        B.ptr, B.col, B.nrows, wl.log2w, (int32_t)wl.nwin, bnd);
    SMCM_LAUNCH_CHECK();
  }
}

}  // namespace

void spgemm(const Csr& A, const Csr& B, const double* est_r, int est_stride, Csr& C, const SpgemmOpts& o,
            SpgemmStats* stats, DevAlloc& al, bool timing, KernelTimer* tm) {
  const bool trace = std::getenv("SMCM_TRACE") != nullptr;
  if (A.ncols != B.nrows) throw HipError(-22, "spgemm: inner dimensions differ");
  if (!B.sorted) throw HipError(-22, "spgemm: right operand rows must be sorted");
  hipStream_t st = al.stream;
  const int32_t m = A.nrows;
  auto t0 = std::chrono::steady_clock::now();
  C.nrows = m;
  C.ncols = B.ncols;
  C.sorted = true;
  C.distinct = true;
  C.ptr = al.alloc_n<int64_t>((int64_t)m + 1);
  SpgemmStats dstats;
  if (spgemm_dense(A, B, C, o, stats ? stats : &dstats, al, tm)) {
    if (timing || tm || trace) SMCM_HIP(hipStreamSynchronize(st));
    if (tm) tm->flush();
    if (stats) stats->ms_numeric = ms_since(t0);
    if (trace)
      fprintf(stderr, "[smcm] spgemm %dx%d * %dx%d: dense tier %.3f ms, entries %lld, products %lld, nnz %lld\n",
              A.nrows, A.ncols, B.nrows, B.ncols, ms_since(t0), (long long)A.nnz,
              (long long)(stats ? stats : &dstats)->products, (long long)C.nnz);
    return;
  }
  SpgemmCore core(A, B, est_r, est_stride, o, al, tm);
  core.bin(0, m);
  if (stats) {
    *stats = SpgemmStats{};
    stats->products = core.products;
    core.add_stats(stats);
    stats->windows = ceil_div(std::max<int64_t>(B.ncols, 1), core.lay.nw);
    stats->ms_analysis = ms_since(t0);
  }
  auto t1 = std::chrono::steady_clock::now();
  {
    DevBuf<int64_t> row_nnz(al, m > 0 ? m : 1);
    fill_i64(row_nnz.p, m, 0, st);
    core.symbolic(row_nnz.p);
    scan_exclusive_i64(row_nnz.p, C.ptr, m, al);
  }
  core.stage_split(C.ptr);
  int64_t nnz = 0;
  SMCM_HIP(hipMemcpyAsync(host_stage() + kStageNnz, C.ptr + m, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SMCM_HIP(hipStreamSynchronize(st));
  memcpy(&nnz, host_stage() + kStageNnz, sizeof(nnz));
  core.read_symbolic();
  if (stats) stats->ms_symbolic = ms_since(t1);
  auto t2 = std::chrono::steady_clock::now();
  C.nnz = nnz;
  C.col = al.alloc_n<int32_t>(nnz);
  C.val = al.alloc_n<uint64_t>(nnz);
  core.numeric(C.ptr, C.col, C.val);
  // callers order later work on the same stream; a synchronisation here only serves timing
  if (timing || tm || trace) SMCM_HIP(hipStreamSynchronize(st));
  if (tm) tm->flush();
  if (stats) {
    stats->ms_numeric = ms_since(t2);
    stats->nnz = nnz;
  }
  if (trace)
    fprintf(stderr,
            "[smcm] spgemm %dx%d * %dx%d: analysis %.3f symbolic %.3f numeric(+alloc) %.3f ms, products %lld, "
            "nnz %lld, windows of %lld columns with %d-byte accumulators\n",
            A.nrows, A.ncols, B.nrows, B.ncols, std::chrono::duration<double, std::milli>(t1 - t0).count(),
            std::chrono::duration<double, std::milli>(t2 - t1).count(), ms_since(t2), (long long)core.products,
            (long long)nnz, (long long)core.split.wl.w, core.split.wl.vbytes);
}

void spgemm_blocked(const Csr& A, const Csr& B, const double* est_r, int est_stride, int64_t block_bytes,
                    const BlockConsumer& consume, const SpgemmOpts& o, SpgemmStats* stats, DevAlloc& al,
                    KernelTimer* tm) {
  if (A.ncols != B.nrows) throw HipError(-22, "spgemm: inner dimensions differ");
  if (!B.sorted) throw HipError(-22, "spgemm: right operand rows must be sorted");
  hipStream_t st = al.stream;
  const int32_t m = A.nrows;
  const bool trace = std::getenv("SMCM_TRACE") != nullptr;
  auto t0 = std::chrono::steady_clock::now();
  SpgemmCore core(A, B, est_r, est_stride, o, al, tm);
  core.bin(0, m);
  if (stats) {
    *stats = SpgemmStats{};
    stats->products = core.products;
    core.add_stats(stats);
  }
  auto t1 = std::chrono::steady_clock::now();
  DevBuf<int64_t> ptr(al, (int64_t)m + 1);
  {
    DevBuf<int64_t> row_nnz(al, m > 0 ? m : 1);
    fill_i64(row_nnz.p, m, 0, st);
    core.symbolic(row_nnz.p);
    scan_exclusive_i64(row_nnz.p, ptr.p, m, al);
  }
  std::vector<int64_t> hptr((size_t)m + 1);
  SMCM_HIP(hipMemcpyAsync(hptr.data(), ptr.p, sizeof(int64_t) * (m + 1), hipMemcpyDeviceToHost, st));
  SMCM_HIP(hipStreamSynchronize(st));
  core.read_symbolic();
  if (tm) tm->flush();
  const int64_t max_entries = std::max<int64_t>(1, block_bytes / 12);
  // cut rows into blocks: widest [b, e) whose output fits; a single row always forms a block
  std::vector<std::pair<int32_t, int32_t>> blocks;
  int64_t widest = 1;
  for (int32_t b = 0; b < m;) {
    int32_t lo = b + 1, hi = m;
    while (lo < hi) {
      const int32_t mid = lo + (hi - lo + 1) / 2;
      if (hptr[mid] - hptr[b] <= max_entries) lo = mid;
      else hi = mid - 1;
    }
    const int32_t e = std::max(lo, b + 1);
    blocks.push_back({b, e});
    widest = std::max<int64_t>(widest, hptr[e] - hptr[b]);
    b = e;
  }
  auto t2 = std::chrono::steady_clock::now();
  // block buffers are allocated once and reused by every block
  DevBuf<int32_t> bcol(al, widest);
  DevBuf<uint64_t> bval(al, widest);
  DevBuf<int64_t> bptr(al, (int64_t)m + 1);
  if (trace) {
    SMCM_HIP(hipStreamSynchronize(st));
    fprintf(stderr,
            "[smcm] blocked %dx%d * %dx%d: analysis %.3f symbolic %.3f alloc %.3f ms, %zu blocks, products %lld, "
            "windows of %lld columns with %d-byte accumulators\n",
            A.nrows, A.ncols, B.nrows, B.ncols, std::chrono::duration<double, std::milli>(t1 - t0).count(),
            std::chrono::duration<double, std::milli>(t2 - t1).count(), ms_since(t2), blocks.size(),
            (long long)core.products, (long long)core.split.wl.w, core.split.wl.vbytes);
  }
  for (auto [b, e] : blocks) {
    auto tb = std::chrono::steady_clock::now();
    const int64_t p0 = hptr[b], cnt = hptr[e] - hptr[b];
    Csr blk;
    blk.nrows = e - b;
    blk.ncols = B.ncols;
    blk.nnz = cnt;
    blk.sorted = true;
    blk.distinct = true;
    core.bin(b, e);
    auto t1 = std::chrono::steady_clock::now();
    core.numeric(ptr.p, bcol.p - p0, bval.p - p0);
    rebase_ptr_kernel<<<(unsigned)ceil_div((int64_t)(e - b) + 1, kThreads), kThreads, 0, st>>>(ptr.p + b, e - b,
                                                                                                bptr.p);
    SMCM_LAUNCH_CHECK();
    blk.ptr = bptr.p;
    blk.col = bcol.p;
    blk.val = bval.p;
    consume(blk, b, e);
    SMCM_HIP(hipStreamSynchronize(st));
    if (trace)
      fprintf(stderr, "[smcm] block rows [%d, %d) nnz %lld bin %.3f ms numeric+consume %.3f ms\n", b, e,
              (long long)cnt, std::chrono::duration<double, std::milli>(t1 - tb).count(), ms_since(t1));
    if (tm) tm->flush();
  }
  if (stats) stats->nnz = hptr[m];
}

}  // namespace smcm
