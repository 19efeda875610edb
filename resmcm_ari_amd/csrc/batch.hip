// batch.hip - persistent executor for the products of many chains in one launch.
//
// Every product of a batch is a node (A * B, where each operand is a base relation, optionally
// through a row map for a source subset, or the output of another node).  A node's rows are cut
// into tiles of tile_rows rows; the work list is a device-side queue of (node, tile) items.
// The host seeds it with the tiles of nodes whose operands are all base relations; the last
// tile of a node to finish makes the node complete, and when every node operand of its
// consumer is complete, that tile pushes the consumer's tiles.  Workgroups are persistent: each
// claims queue slots in order until every tile of the batch has been claimed.
//
// No symbolic pass and no scan between products: a tile computes the product count (upper bound
// of the width) of each of its rows, reserves that much of the output arena with one atomic,
// and writes each row sorted and merged at its reserved offset; rows are described by
// (start, length) pairs, which is how later products read them.  Per row:
//   products <= 64         one wave, one product per lane, bitonic sort across the wave;
//   products <= sort_cap   one wave, products sorted in the wave's own LDS region;
//   longer                 the whole workgroup, dense LDS accumulators over column windows.
// All arithmetic is exact uint64 (mod 2^64), so results equal the per-chain path's bit for bit.
#include "kernels.h"

namespace smcm {

namespace {

constexpr int kMaxTileRows = 64;

struct Opnd {
  const int64_t* ptr;    // base CSR row pointers (null for a node output)
  const int32_t* rows;   // row map of a base CSR (null = identity)
  const int32_t* col;
  const uint64_t* val;   // null = every value 1
  const int64_t* start;  // node output rows (null for a base CSR)
  const int32_t* len;
};

__device__ __forceinline__ Opnd operand_a(const BatchNode& N, const BatchLaunch& bl) {
  if (N.a_node >= 0) {
    const BatchNode& S = bl.nodes[N.a_node];
    return Opnd{nullptr, nullptr, bl.arena_col, bl.arena_val, S.out_start, S.out_len};
  }
  return Opnd{N.a_ptr, N.a_rows, N.a_col, N.a_val, nullptr, nullptr};
}

__device__ __forceinline__ Opnd operand_b(const BatchNode& N, const BatchLaunch& bl) {
  if (N.b_node >= 0) {
    const BatchNode& S = bl.nodes[N.b_node];
    return Opnd{nullptr, nullptr, bl.arena_col, bl.arena_val, S.out_start, S.out_len};
  }
  return Opnd{N.b_ptr, nullptr, N.b_col, N.b_val, nullptr, nullptr};
}

__device__ __forceinline__ void row_of(const Opnd& o, int32_t u, int64_t& s, int32_t& n) {
  if (o.start) {
    s = o.start[u];
    n = o.len[u];
  } else {
    const int32_t r = o.rows ? o.rows[u] : u;
    s = o.ptr[r];
    n = (int32_t)(o.ptr[r + 1] - s);
  }
}

__device__ __forceinline__ uint64_t lanes_upto(int lane) {
  return lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);
}

// ub <= 64: one product per lane, bitonic sort across the wave, runs of equal columns summed
__device__ void row_wave(const Opnd& A, const Opnd& B, int64_t as, int32_t an, int32_t* ocol, uint64_t* oval,
                         int32_t& nnz) {
  const int lane = lane_id();
  int32_t k = INT32_MAX;
  uint64_t v = 0;
  int off = 0;
  for (int32_t x = 0; x < an; ++x) {
    const int32_t kb = A.col[as + x];
    const uint64_t av = A.val ? A.val[as + x] : 1ull;
    int64_t bs;
    int32_t bn;
    row_of(B, kb, bs, bn);
    if (lane >= off && lane < off + bn) {
      const int64_t b = bs + (lane - off);
      k = B.col[b];
      v = B.val ? av * B.val[b] : av;
    }
    off += bn;
  }
#pragma unroll
  for (int size = 2; size <= kWave; size <<= 1) {
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1) {
      const int32_t ok = __shfl_xor(k, j);
      const uint64_t ov = __shfl_xor(v, j);
      const bool asc = (lane & size) == 0;
      const bool lower = (lane & j) == 0;
      const bool take = (lower == asc) ? (ok < k) : (ok > k);
      if (take) {
        k = ok;
        v = ov;
      }
    }
  }
  const int32_t kprev = __shfl_up(k, 1);
  const bool valid = k != INT32_MAX;
  const bool head = valid && (lane == 0 || kprev != k);
  const uint64_t hmask = __ballot(head);
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint64_t y = __shfl_up(v, o);
    const int32_t ky = __shfl_up(k, o);
    if (lane >= o && ky == k) v += y;
  }
  const int32_t knext = __shfl_down(k, 1);
  const bool last = valid && (lane == kWave - 1 || knext != k);
  if (last) {
    const int pos = __popcll(hmask & lanes_upto(lane)) - 1;
    ocol[pos] = k;
    oval[pos] = v;
  }
  nnz = __popcll(hmask);
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 64 < ub <= cap: the wave expands the row into its LDS region, sorts it (bitonic) and merges
__device__ void row_lds(const Opnd& A, const Opnd& B, int64_t as, int32_t an, int32_t ub, int32_t* key,
                        uint64_t* val, int32_t* ocol, uint64_t* oval, int32_t& nnz) {
  const int lane = lane_id();
  int32_t off = 0;
  for (int32_t x = 0; x < an; ++x) {
    const int32_t kb = A.col[as + x];
    const uint64_t av = A.val ? A.val[as + x] : 1ull;
    int64_t bs;
    int32_t bn;
    row_of(B, kb, bs, bn);
    for (int32_t j = lane; j < bn; j += kWave) {
      key[off + j] = B.col[bs + j];
      val[off + j] = B.val ? av * B.val[bs + j] : av;
    }
    off += bn;
  }
  int32_t P = kWave;
  while (P < ub) P <<= 1;
  for (int32_t i = ub + lane; i < P; i += kWave) key[i] = INT32_MAX;
  wave_sync_lds();
  for (int32_t size = 2; size <= P; size <<= 1) {
    for (int32_t j = size >> 1; j > 0; j >>= 1) {
      for (int32_t t = lane; t < P / 2; t += kWave) {
        const int32_t i = 2 * j * (t / j) + (t % j);
        const int32_t p = i + j;
        const int32_t ki = key[i], kp = key[p];
        const bool asc = (i & size) == 0;
        if (asc ? (ki > kp) : (ki < kp)) {
          key[i] = kp;
          key[p] = ki;
          const uint64_t vi = val[i];
          val[i] = val[p];
          val[p] = vi;
        }
      }
      wave_sync_lds();
    }
  }
  int32_t count = 0, carry_k = INT32_MAX;
  uint64_t carry_v = 0;
  for (int32_t c0 = 0; c0 < ub; c0 += kWave) {
    const int32_t i = c0 + lane;
    const bool valid = i < ub;
    const int32_t k = valid ? key[i] : INT32_MAX;
    uint64_t v = valid ? val[i] : 0ull;
    int32_t kprev = __shfl_up(k, 1);
    if (lane == 0) kprev = carry_k;
    const bool head = valid && k != kprev;
    const uint64_t hmask = __ballot(head);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint64_t y = __shfl_up(v, o);
      const int32_t ky = __shfl_up(k, o);
      if (lane >= o && ky == k) v += y;
    }
    if (valid && k == carry_k) v += carry_v;  // the run left open by the previous chunk
    int32_t knext = __shfl_down(k, 1);
    if (lane == kWave - 1) knext = (c0 + kWave < ub) ? key[c0 + kWave] : INT32_MAX;
    const bool last = valid && knext != k;
    if (last) {
      const int32_t pos = count + __popcll(hmask & lanes_upto(lane)) - 1;
      ocol[pos] = k;
      oval[pos] = v;
    }
    count += __popcll(hmask);
    carry_k = __shfl(k, kWave - 1);
    carry_v = __shfl(v, kWave - 1);
  }
  nnz = count;
  wave_sync_lds();
}

// ub > cap: the workgroup accumulates the row window by window (window = 2^log2w columns) in
// LDS (values + presence bits), then writes each window's columns in order
template <int THREADS>
__device__ void row_windows(const Opnd& A, const Opnd& B, int64_t as, int32_t an, int32_t ncols, int log2w,
                            uint64_t* wval, uint32_t* wbit, int32_t* red, int32_t* ocol, uint64_t* oval,
                            int32_t& nnz) {
  constexpr int NW = THREADS / kWave;
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  const int32_t W = 1 << log2w;
  const int32_t nwords = W >> 5;
  const int32_t nwin = (ncols + W - 1) >> log2w;
  int32_t pos = 0;
  for (int32_t w = 0; w < nwin; ++w) {
    for (int32_t i = threadIdx.x; i < W; i += THREADS) wval[i] = 0;
    for (int32_t i = threadIdx.x; i < nwords; i += THREADS) wbit[i] = 0;
    __syncthreads();
    for (int32_t x = wid; x < an; x += NW) {
      const int32_t kb = A.col[as + x];
      const uint64_t av = A.val ? A.val[as + x] : 1ull;
      int64_t bs;
      int32_t bn;
      row_of(B, kb, bs, bn);
      for (int32_t j = lane; j < bn; j += kWave) {
        const int32_t c = B.col[bs + j];
        if ((c >> log2w) == w) {
          const int32_t lc = c & (W - 1);
          atomicAdd((unsigned long long*)&wval[lc], (unsigned long long)(B.val ? av * B.val[bs + j] : av));
          atomicOr(&wbit[lc >> 5], 1u << (lc & 31));
        }
      }
    }
    __syncthreads();
    // words per thread: contiguous runs keep the output in column order
    const int32_t wpt = (nwords + THREADS - 1) / THREADS;
    const int32_t w0 = threadIdx.x * wpt, w1 = min(w0 + wpt, nwords);
    int32_t cnt = 0;
    for (int32_t q = w0; q < w1; ++q) cnt += __popc(wbit[q]);
    int32_t total;
    const int32_t ex = [&] {
      int32_t incl = cnt;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const int32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
      }
      if (lane == kWave - 1) red[wid] = incl;
      __syncthreads();
      int32_t before = 0, all = 0;
      for (int q = 0; q < NW; ++q) {
        if (q < wid) before += red[q];
        all += red[q];
      }
      total = all;
      return before + incl - cnt;
    }();
    int32_t o = pos + ex;
    for (int32_t q = w0; q < w1; ++q) {
      uint32_t bits = wbit[q];
      while (bits) {
        const int b = __ffs(bits) - 1;
        bits &= bits - 1;
        const int32_t lc = (q << 5) | b;
        ocol[o] = (w << log2w) | lc;
        oval[o] = wval[lc];
        ++o;
      }
    }
    pos += total;
    __syncthreads();
  }
  nnz = pos;
}

template <int THREADS, int CAP>
__global__ void __launch_bounds__(THREADS) batch_kernel(BatchLaunch bl) {
  constexpr int NW = THREADS / kWave;
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int64_t s_ub[kMaxTileRows];
  __shared__ int64_t s_off[kMaxTileRows];
  __shared__ int32_t s_red[NW];
  __shared__ unsigned long long s_item;
  __shared__ int64_t s_base;
  __shared__ int32_t s_fail;
  __shared__ int32_t s_push_node, s_push_slot;
  const int lane = lane_id(), wid = threadIdx.x / kWave;
  // wave regions for the LDS sort: keys then values
  int32_t* wkey = (int32_t*)(lds + (size_t)wid * CAP * 12);
  uint64_t* wvalr = (uint64_t*)(lds + (size_t)wid * CAP * 12 + (size_t)CAP * 4);
  // window accumulators (the same LDS, used between barriers)
  uint64_t* wval = (uint64_t*)lds;
  uint32_t* wbit = (uint32_t*)(lds + ((size_t)8 << bl.log2w));

  for (;;) {
    if (threadIdx.x == 0) {
      unsigned long long item = 0;
      const int32_t slot = atomicAdd(&bl.state->head, 1);
      if (slot < bl.total_tiles) {
        const long long t0 = wall_clock64();
        while ((item = __hip_atomic_load(&bl.queue[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
          __builtin_amdgcn_s_sleep(8);
          if (wall_clock64() - t0 > bl.spin_limit) {
            atomicOr(&bl.state->error, 2);
            break;
          }
        }
      }
      s_item = item;
    }
    __syncthreads();
    __threadfence();
    const unsigned long long item = s_item;
    if (item == 0) break;  // every tile claimed (or a wait gave up)
    const int32_t n = (int32_t)((item - 1) >> 32);
    const int32_t t = (int32_t)((item - 1) & 0xffffffffull);
    const BatchNode& N = bl.nodes[n];
    const Opnd A = operand_a(N, bl), B = operand_b(N, bl);
    const int32_t r0 = t * bl.tile_rows;
    const int32_t nr = min(bl.tile_rows, N.nrows - r0);
    const bool in_fail =
        (N.a_node >= 0 && __hip_atomic_load(&bl.failed[N.a_node], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ||
        (N.b_node >= 0 && __hip_atomic_load(&bl.failed[N.b_node], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    // 1. products per row
    for (int32_t r = wid; r < nr; r += NW) {
      int64_t sum = 0;
      if (!in_fail) {
        int64_t as;
        int32_t an;
        row_of(A, r0 + r, as, an);
        for (int32_t x = lane; x < an; x += kWave) {
          int64_t bs;
          int32_t bn;
          row_of(B, A.col[as + x], bs, bn);
          sum += bn;
        }
        sum = wave_sum_i64(sum);
      }
      if (lane == 0) s_ub[r] = sum;
    }
    __syncthreads();
    // 2. reserve the tile's output
    if (threadIdx.x == 0) {
      int64_t tot = 0;
      for (int32_t r = 0; r < nr; ++r) {
        s_off[r] = tot;
        tot += s_ub[r];
      }
      int fail = in_fail ? 1 : 0;
      int64_t base = 0;
      if (!fail && tot > 0) {
        base = (int64_t)atomicAdd(&bl.state->arena_used, (unsigned long long)tot);
        if ((uint64_t)(base + tot) > bl.arena_cap) {
          fail = 1;
          atomicOr(&bl.state->error, 1);
        }
      }
      s_base = base;
      s_fail = fail;
    }
    __syncthreads();
    const bool fail = s_fail != 0;
    const int64_t base = s_base;
    // 3. rows that one wave computes
    for (int32_t r = wid; r < nr; r += NW) {
      const int64_t ub = s_ub[r];
      if (!fail && ub > CAP) continue;
      int32_t nnz = 0;
      const int64_t o = base + s_off[r];
      if (!fail && ub > 0) {
        int64_t as;
        int32_t an;
        row_of(A, r0 + r, as, an);
        if (ub <= kWave)
          row_wave(A, B, as, an, bl.arena_col + o, bl.arena_val + o, nnz);
        else
          row_lds(A, B, as, an, (int32_t)ub, wkey, wvalr, bl.arena_col + o, bl.arena_val + o, nnz);
      }
      if (lane == 0) {
        N.out_start[r0 + r] = fail ? 0 : o;
        N.out_len[r0 + r] = nnz;
      }
    }
    __syncthreads();
    // 4. rows that the workgroup computes
    if (!fail) {
      for (int32_t r = 0; r < nr; ++r) {
        if (s_ub[r] <= CAP) continue;
        const int64_t o = base + s_off[r];
        int64_t as;
        int32_t an;
        row_of(A, r0 + r, as, an);
        int32_t nnz = 0;
        row_windows<THREADS>(A, B, as, an, N.ncols, bl.log2w, wval, wbit, s_red, bl.arena_col + o,
                             bl.arena_val + o, nnz);
        if (threadIdx.x == 0) {
          N.out_start[r0 + r] = o;
          N.out_len[r0 + r] = nnz;
        }
      }
    }
    // 5. completion: every thread's writes are visible device-wide before the count moves
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
      s_push_node = -1;
      if (fail) __hip_atomic_store(&bl.failed[n], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      const int32_t d = atomicAdd(&bl.done_tiles[n], 1);
      if (d == N.ntiles - 1) {
        __threadfence();
        for (int32_t q = bl.parent_off[n]; q < bl.parent_off[n + 1]; ++q) {
          const int32_t p = bl.parents[q];
          const int32_t rdy = atomicAdd(&bl.ready[p], 1);
          if (rdy + 1 == bl.nodes[p].ninputs) {
            // at most one consumer becomes ready per completed node in a chain; a node shared by
            // several consumers pushes the others below, one per pass
            if (s_push_node < 0) {
              s_push_node = p;
              s_push_slot = atomicAdd(&bl.state->tail, bl.nodes[p].ntiles);
            } else {
              const int32_t slot = atomicAdd(&bl.state->tail, bl.nodes[p].ntiles);
              for (int32_t i = 0; i < bl.nodes[p].ntiles; ++i)
                __hip_atomic_store(&bl.queue[slot + i], (((unsigned long long)p << 32) | (unsigned)i) + 1ull,
                                   __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
          }
        }
      }
    }
    __syncthreads();
    const int32_t pn = s_push_node;
    if (pn >= 0) {
      const int32_t slot = s_push_slot;
      const int32_t cnt = bl.nodes[pn].ntiles;
      for (int32_t i = threadIdx.x; i < cnt; i += THREADS)
        __hip_atomic_store(&bl.queue[slot + i], (((unsigned long long)pn << 32) | (unsigned)i) + 1ull,
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
}

// per-row digests (same mixing as row_digest) and per-row lengths of padded rows
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) padded_digest_kernel(const int64_t* start, const int32_t* len,
                                                           const int32_t* col, const uint64_t* val,
                                                           const int32_t* colmap, int32_t nrows, uint64_t* out) {
  const int lane = lane_id();
  const int32_t u = blockIdx.x * 4 + threadIdx.x / kWave;
  if (u >= nrows) return;
  const int64_t s = start[u];
  const int32_t n = len[u];
  uint64_t h = 0;
  for (int32_t x = lane; x < n; x += kWave) {
    const int32_t c = colmap ? colmap[col[s + x]] : col[s + x];
    h += mix64(mix64((uint64_t)(uint32_t)c ^ 0x9E3779B97F4A7C15ull) ^ val[s + x]);
  }
  h = (uint64_t)wave_sum_i64((int64_t)h);
  if (lane == 0) out[u] = mix64((uint64_t)n ^ 0x243F6A8885A308D3ull) + h;
}

__global__ void padded_copy_kernel(const int64_t* start, const int32_t* len, const int64_t* ptr,
                                   const int32_t* col, const uint64_t* val, int32_t nrows, int32_t* ocol,
                                   uint64_t* oval) {
  const int lane = lane_id();
  const int32_t u = blockIdx.x * 4 + threadIdx.x / kWave;
  if (u >= nrows) return;
  const int64_t s = start[u], d = ptr[u];
  const int32_t n = len[u];
  for (int32_t x = lane; x < n; x += kWave) {
    ocol[d + x] = col[s + x];
    oval[d + x] = val[s + x];
  }
}

__global__ void len_to_i64_kernel(const int32_t* len, int32_t n, int64_t* out) {
  const int32_t u = blockIdx.x * 256 + threadIdx.x;
  if (u < n) out[u] = len[u];
}

template <int THREADS, int CAP>
void launch(const BatchLaunch& bl, hipStream_t st) {
  const size_t sort_lds = (size_t)(THREADS / kWave) * CAP * 12;
  const size_t win_lds = ((size_t)8 << bl.log2w) + ((size_t)1 << bl.log2w) / 8;
  const size_t lds = std::max(sort_lds, win_lds);
  batch_kernel<THREADS, CAP><<<(unsigned)bl.grid, THREADS, lds, st>>>(bl);
  SMCM_LAUNCH_CHECK();
}

}  // namespace

size_t batch_lds_bytes(int threads, int sort_cap, int log2w) {
  const size_t sort_lds = (size_t)(threads / kWave) * sort_cap * 12;
  const size_t win_lds = ((size_t)8 << log2w) + ((size_t)1 << log2w) / 8;
  return std::max(sort_lds, win_lds);
}

bool batch_shape_supported(int threads, int sort_cap) {
  return (threads == 256 && (sort_cap == 512 || sort_cap == 1024 || sort_cap == 2048)) ||
         (threads == 512 && (sort_cap == 512 || sort_cap == 1024));
}

void batch_execute(const BatchLaunch& bl, hipStream_t st) {
  if (bl.tile_rows < 1 || bl.tile_rows > kMaxTileRows) throw HipError(-22, "batch: tile_rows must be in [1, 64]");
  if (bl.log2w < 8 || bl.log2w > 16) throw HipError(-22, "batch: window must be 2^8..2^16 columns");
#define SMCM_B(T, C)                  \
  if (bl.threads == T && bl.sort_cap == C) \
    return launch<T, C>(bl, st);
  SMCM_B(256, 512) SMCM_B(256, 1024) SMCM_B(256, 2048) SMCM_B(512, 512) SMCM_B(512, 1024)
#undef SMCM_B
  throw HipError(-22, "batch: unsupported workgroup shape");
}

void padded_digest(const int64_t* start, const int32_t* len, const int32_t* col, const uint64_t* val,
                   const int32_t* colmap, int32_t nrows, uint64_t* out, hipStream_t st) {
  if (nrows <= 0) return;
  padded_digest_kernel<<<(unsigned)ceil_div(nrows, 4), 256, 0, st>>>(start, len, col, val, colmap, nrows, out);
  SMCM_LAUNCH_CHECK();
}

void padded_to_csr(const int64_t* start, const int32_t* len, const int32_t* col, const uint64_t* val, int32_t nrows,
                   int32_t ncols, Csr& out, DevAlloc& al) {
  hipStream_t st = al.stream;
  out = Csr{};
  out.nrows = nrows;
  out.ncols = ncols;
  out.ptr = al.alloc_n<int64_t>((int64_t)nrows + 1);
  DevBuf<int64_t> l64(al, nrows > 0 ? nrows : 1);
  if (nrows > 0) {
    len_to_i64_kernel<<<(unsigned)ceil_div(nrows, 256), 256, 0, st>>>(len, nrows, l64.p);
    SMCM_LAUNCH_CHECK();
  }
  scan_exclusive_i64(l64.p, out.ptr, nrows, al);
  int64_t nnz = 0;
  SMCM_HIP(hipMemcpyAsync(&nnz, out.ptr + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SMCM_HIP(hipStreamSynchronize(st));
  out.nnz = nnz;
  out.col = al.alloc_n<int32_t>(nnz);
  out.val = al.alloc_n<uint64_t>(nnz);
  if (nrows > 0) {
    padded_copy_kernel<<<(unsigned)ceil_div(nrows, 4), 256, 0, st>>>(start, len, out.ptr, col, val, nrows, out.col,
                                                                      out.val);
    SMCM_LAUNCH_CHECK();
  }
}

}  // namespace smcm
