// device_graph.hip - on-device construction of typed relation matrices from the resident HIN.
//
// The HIN lives in HBM as a global adjacency CSR (vertex types, int64 row pointers, int32
// neighbour ids, int32 entry types).  A meta-path edge (ts --et--> td) becomes a CSR over the
// vertices of type ts (ascending global id) whose columns are the type-local ids of the
// matching neighbours, in adjacency order with duplicates kept (BoolMatrix constructor,
// efficient/resmcm.cpp:374-396).  Rows are extracted one wave per row: each lane tests one
// adjacency entry, a ballot gives the matching lanes and their order-preserving output slots.
#include "kernels.h"

namespace smcm {

namespace {

constexpr int kThreads = 256;

__global__ void fill_i64_kernel(int64_t* p, int64_t n, int64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void type_flag_kernel(const int32_t* __restrict__ vtype, int32_t n, int32_t t, int64_t* __restrict__ flag) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) flag[v] = vtype[v] == t ? 1 : 0;
}

__global__ void type_scatter_kernel(const int32_t* __restrict__ vtype, int32_t n, int32_t t,
                                    const int64_t* __restrict__ pos, int32_t* __restrict__ local,
                                    int32_t* __restrict__ bins_t) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && vtype[v] == t) {
    int32_t p = (int32_t)pos[v];
    local[v] = p;
    bins_t[p] = v;
  }
}

// one wave per row: count matching entries
__global__ void __launch_bounds__(kThreads) typed_count_kernel(
    const int32_t* __restrict__ vtype, const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj_nbr,
    const int32_t* __restrict__ adj_et, const int32_t* __restrict__ bins_ts, const int32_t* __restrict__ rows_sel,
    int32_t nrows, int32_t et, int32_t td, int64_t* __restrict__ cnt) {
  const int lane = lane_id();
  const int64_t wpb = kThreads / kWave;
  for (int64_t r = (int64_t)blockIdx.x * wpb + threadIdx.x / kWave; r < nrows; r += (int64_t)gridDim.x * wpb) {
    const int32_t g = bins_ts[rows_sel ? rows_sel[r] : r];
    const int64_t p0 = adj_ptr[g], p1 = adj_ptr[g + 1];
    int64_t c = 0;
    for (int64_t b = p0; b < p1; b += kWave) {
      const int64_t x = b + lane;
      bool m = false;
      if (x < p1) m = adj_et[x] == et && vtype[adj_nbr[x]] == td;
      c += __popcll(__ballot(m));
    }
    if (lane == 0) cnt[r] = c;
  }
}

// one wave per row: order-preserving compaction of matching entries; detects unsorted rows
__global__ void __launch_bounds__(kThreads) typed_fill_kernel(
    const int32_t* __restrict__ vtype, const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj_nbr,
    const int32_t* __restrict__ adj_et, const int32_t* __restrict__ local, const int32_t* __restrict__ bins_ts,
    const int32_t* __restrict__ rows_sel, int32_t nrows, int32_t et, int32_t td, const int64_t* __restrict__ ptr,
    int32_t* __restrict__ col, int32_t* __restrict__ unsorted) {
  const int lane = lane_id();
  const int64_t wpb = kThreads / kWave;
  for (int64_t r = (int64_t)blockIdx.x * wpb + threadIdx.x / kWave; r < nrows; r += (int64_t)gridDim.x * wpb) {
    const int32_t g = bins_ts[rows_sel ? rows_sel[r] : r];
    const int64_t p0 = adj_ptr[g], p1 = adj_ptr[g + 1];
    int64_t out = ptr[r];
    int32_t carry = INT32_MIN;  // last column written by the previous chunk
    bool bad = false;
    for (int64_t b = p0; b < p1; b += kWave) {
      const int64_t x = b + lane;
      bool m = false;
      int32_t c = 0;
      if (x < p1) {
        int32_t nb = adj_nbr[x];
        m = adj_et[x] == et && vtype[nb] == td;
        if (m) c = local[nb];
      }
      const uint64_t mask = __ballot(m);
      const uint64_t below = mask & ((1ull << lane) - 1ull);
      const int src = below ? 63 - __clzll(below) : lane;
      const int32_t prev_in_chunk = __shfl(c, src);
      if (m) {
        col[out + __popcll(below)] = c;
        const int32_t prev = below ? prev_in_chunk : carry;
        bad |= c < prev;
      }
      if (mask) carry = __shfl(c, 63 - __clzll(mask));
      out += __popcll(mask);
    }
    if (__ballot(bad)) {
      if (lane == 0) atomicOr(unsorted, 1);
    }
  }
}

__global__ void __launch_bounds__(kThreads) gather_len_kernel(const int64_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ rows_sel, int32_t nsel,
                                                              int64_t* __restrict__ len) {
  int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nsel) {
    int32_t s = rows_sel[r];
    len[r] = ptr[s + 1] - ptr[s];
  }
}

__global__ void __launch_bounds__(kThreads) gather_rows_kernel(
    const int64_t* __restrict__ iptr, const int32_t* __restrict__ icol, const uint64_t* __restrict__ ival,
    const int32_t* __restrict__ rows_sel, int32_t nsel, const int64_t* __restrict__ optr, int32_t* __restrict__ ocol,
    uint64_t* __restrict__ oval) {
  const int lane = lane_id();
  const int64_t wpb = kThreads / kWave;
  for (int64_t r = (int64_t)blockIdx.x * wpb + threadIdx.x / kWave; r < nsel; r += (int64_t)gridDim.x * wpb) {
    const int32_t s = rows_sel[r];
    const int64_t p0 = iptr[s], p1 = iptr[s + 1], o = optr[r];
    for (int64_t x = p0 + lane; x < p1; x += kWave) {
      ocol[o + (x - p0)] = icol[x];
      if (ival) oval[o + (x - p0)] = ival[x];
    }
  }
}

__global__ void map_columns_kernel(const int32_t* __restrict__ src, int64_t n, const int32_t* __restrict__ map,
                                   int32_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = map[src[i]];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// row digest = mix(width) + sum over entries of mix(mix(col) ^ val), mod 2^64: one wave per row
__global__ void __launch_bounds__(kThreads) row_digest_kernel(const int64_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ col,
                                                              const uint64_t* __restrict__ val, int32_t nrows,
                                                              uint64_t* __restrict__ out) {
  const int lane = lane_id();
  const int64_t wpb = kThreads / kWave;
  for (int64_t u = (int64_t)blockIdx.x * wpb + threadIdx.x / kWave; u < nrows; u += (int64_t)gridDim.x * wpb) {
    const int64_t p0 = ptr[u], p1 = ptr[u + 1];
    uint64_t h = 0;
    for (int64_t x = p0 + lane; x < p1; x += kWave)
      h += mix64(mix64((uint64_t)(uint32_t)col[x] ^ 0x9E3779B97F4A7C15ull) ^ (val ? val[x] : 1ull));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
    if (lane == 0) out[u] = h + mix64((uint64_t)(p1 - p0) ^ 0x243F6A8885A308D3ull);
  }
}

// plain device copy in 16-byte vectors (bandwidth reference)
__global__ void __launch_bounds__(kThreads) copy_kernel(const int4* __restrict__ src, int4* __restrict__ dst,
                                                        int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    dst[i] = src[i];
}

unsigned grid_for(int64_t items, int per_block, int64_t cap = 1 << 20) {
  int64_t g = ceil_div(items, per_block);
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace

void fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t st) {
  if (n <= 0) return;
  fill_i64_kernel<<<grid_for(n, kThreads, 8192), kThreads, 0, st>>>(p, n, v);
  SMCM_LAUNCH_CHECK();
}

void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
  if (n <= 0) return;
  fill_i32_kernel<<<grid_for(n, kThreads, 8192), kThreads, 0, st>>>(p, n, v);
  SMCM_LAUNCH_CHECK();
}

void build_type_index(const int32_t* vtype, int32_t n, int32_t ntypes, int32_t* local,
                      int32_t* const* bins, DevAlloc& al) {
  if (n <= 0) return;
  DevBuf<int64_t> flag(al, n), pos(al, (int64_t)n + 1);
  for (int32_t t = 0; t < ntypes; ++t) {
    if (!bins[t]) continue;
    type_flag_kernel<<<grid_for(n, kThreads), kThreads, 0, al.stream>>>(vtype, n, t, flag.p);
    SMCM_LAUNCH_CHECK();
    scan_exclusive_i64(flag.p, pos.p, n, al);
    type_scatter_kernel<<<grid_for(n, kThreads), kThreads, 0, al.stream>>>(vtype, n, t, pos.p, local, bins[t]);
    SMCM_LAUNCH_CHECK();
  }
}

void build_typed_csr(const int32_t* vtype, const int64_t* adj_ptr, const int32_t* adj_nbr,
                     const int32_t* adj_et, const int32_t* local, const int32_t* bins_ts,
                     const int32_t* rows_sel, int32_t nrows, int32_t et, int32_t td, int32_t ncols,
                     Csr& out, DevAlloc& al) {
  out.nrows = nrows;
  out.ncols = ncols;
  out.val = nullptr;
  out.distinct = false;  // adjacency duplicates are kept
  out.ptr = al.alloc_n<int64_t>((int64_t)nrows + 1);
  const int wpb = kThreads / kWave;
  if (nrows == 0) {
    fill_i64(out.ptr, 1, 0, al.stream);
    out.nnz = 0;
    out.col = al.alloc_n<int32_t>(1);
    out.sorted = true;
    return;
  }
  DevBuf<int64_t> cnt(al, nrows);
  typed_count_kernel<<<grid_for(nrows, wpb, 65536), kThreads, 0, al.stream>>>(vtype, adj_ptr, adj_nbr, adj_et, bins_ts,
                                                                               rows_sel, nrows, et, td, cnt.p);
  SMCM_LAUNCH_CHECK();
  scan_exclusive_i64(cnt.p, out.ptr, nrows, al);
  int64_t nnz = 0;
  SMCM_HIP(hipMemcpyAsync(&nnz, out.ptr + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  out.nnz = nnz;
  out.col = al.alloc_n<int32_t>(nnz);
  DevBuf<int32_t> flag(al, 1);
  fill_i32(flag.p, 1, 0, al.stream);
  typed_fill_kernel<<<grid_for(nrows, wpb, 65536), kThreads, 0, al.stream>>>(
      vtype, adj_ptr, adj_nbr, adj_et, local, bins_ts, rows_sel, nrows, et, td, out.ptr, out.col, flag.p);
  SMCM_LAUNCH_CHECK();
  int32_t unsorted = 0;
  SMCM_HIP(hipMemcpyAsync(&unsorted, flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  out.sorted = unsorted == 0;
}

void csr_gather_rows(const Csr& in, const int32_t* rows_sel, int32_t nsel, Csr& out, DevAlloc& al) {
  out.nrows = nsel;
  out.ncols = in.ncols;
  out.sorted = in.sorted;
  out.distinct = in.distinct;
  out.ptr = al.alloc_n<int64_t>((int64_t)nsel + 1);
  DevBuf<int64_t> len(al, nsel);
  if (nsel > 0) {
    gather_len_kernel<<<grid_for(nsel, kThreads), kThreads, 0, al.stream>>>(in.ptr, rows_sel, nsel, len.p);
    SMCM_LAUNCH_CHECK();
  }
  scan_exclusive_i64(len.p, out.ptr, nsel, al);
  int64_t nnz = 0;
  SMCM_HIP(hipMemcpyAsync(&nnz, out.ptr + nsel, sizeof(int64_t), hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  out.nnz = nnz;
  out.col = al.alloc_n<int32_t>(nnz);
  out.val = in.val ? al.alloc_n<uint64_t>(nnz) : nullptr;
  if (nsel > 0) {
    const int wpb = kThreads / kWave;
    gather_rows_kernel<<<grid_for(nsel, wpb, 65536), kThreads, 0, al.stream>>>(in.ptr, in.col, in.val, rows_sel,
                                                                                nsel, out.ptr, out.col, out.val);
    SMCM_LAUNCH_CHECK();
  }
}

void row_digest(const Csr& c, uint64_t* out, hipStream_t st) {
  if (c.nrows <= 0) return;
  row_digest_kernel<<<grid_for(c.nrows, kThreads / kWave, 65536), kThreads, 0, st>>>(c.ptr, c.col, c.val, c.nrows,
                                                                                     out);
  SMCM_LAUNCH_CHECK();
}

void device_copy(const void* src, void* dst, int64_t bytes, hipStream_t st) {
  const int64_t n = bytes / 16;
  if (n <= 0) return;
  copy_kernel<<<grid_for(n, kThreads, 256 * 64), kThreads, 0, st>>>((const int4*)src, (int4*)dst, n);
  SMCM_LAUNCH_CHECK();
}

void map_columns(const int32_t* src, int64_t n, const int32_t* map, int32_t* dst, hipStream_t st) {
  if (n <= 0) return;
  map_columns_kernel<<<grid_for(n, kThreads, 16384), kThreads, 0, st>>>(src, n, map, dst);
  SMCM_LAUNCH_CHECK();
}

}  // namespace smcm
