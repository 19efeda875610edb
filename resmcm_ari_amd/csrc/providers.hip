// providers.hip - the reference's alternative DP cost providers on the GPU: MetaAC
// (efficient/meta.cpp), MNC sketches (efficient/mnc.cpp), DMap density maps (efficient/dm.cpp)
// and LGraph rank propagation (efficient/lg.cpp), plus the estimate-only variants of
// accuracy/{meta,mnc,dm,lg}.cpp.
//
// All four work in the reference's global index space: M_i is an n x n matrix whose rows are the
// vertices of type t_i, columns global vertex ids.  The typed CSRs here have local rows and
// columns; rowg[i] / colg[i] (device) map them to global ids.  The sketch work (histograms,
// dot products, density-map products, rank propagation) runs in kernels; the DP over the
// L x L table is host code with the reference's evaluation order and lowest-k tie rule.
//
// Arithmetic: MetaAC and DMap compute in binary64 as the reference does; MNC and LGraph use
// long double in the reference (x87 80-bit) and binary64 here.  Integer parts (sketch counts,
// dot products, block counts) are exact.  Sums whose order the reference leaves to OpenMP or
// evaluates sequentially are tree sums here; the union recurrence sp = sp + x - sp*x of MNC
// is evaluated as 1 - prod(1 - x) (as -expm1(sum log1p(-x)) when every x < 1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "kernels.h"
#include "providers.h"

namespace smcm {
namespace {

constexpr int kT = 256;
constexpr int kBlocks = 512;  // partial sums of the deterministic reductions
constexpr double kAlph = -130.0, kBt = -16.0, kGmm = 83.0;  // efficient/meta.cpp:44-46
constexpr int kDmapB = 256;   // DenseBlock::_b (dm.cpp:56)
constexpr int kLgR = 32;      // LayeredGraph::_r (lg.cpp:61)

__device__ __forceinline__ double block_sum_f64(double x, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) red[w] = x;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)blockDim.x / kWave; ++i) t += red[i];
  __syncthreads();
  return t;
}

// ---- MetaAC -----------------------------------------------------------------------------
// partial[b] = sum over this block's rows of (double)len / n / n
__global__ void __launch_bounds__(kT) metaac_terms_kernel(const int64_t* __restrict__ ptr, int32_t nrows, double n,
                                                          double* __restrict__ partial) {
  __shared__ double red[kT / kWave];
  double s = 0;
  for (int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * kT)
    s += (double)(ptr[r + 1] - ptr[r]) / n / n;
  s = block_sum_f64(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---- MNC --------------------------------------------------------------------------------
// a wave per row (lanes over its entries), so hub rows are spread over 64 lanes
__global__ void __launch_bounds__(kT) mnc_counts_kernel(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                                        int32_t nrows, const int32_t* __restrict__ rowg,
                                                        const int32_t* __restrict__ colg, int32_t* __restrict__ h_r,
                                                        int32_t* __restrict__ h_c) {
  const int lane = lane_id();
  for (int64_t r = ((int64_t)blockIdx.x * kT + threadIdx.x) / kWave; r < nrows; r += (int64_t)gridDim.x * (kT / kWave)) {
    const int64_t a = ptr[r], b = ptr[r + 1];
    if (lane == 0) h_r[rowg[r]] = (int32_t)(b - a);
    for (int64_t e = a + lane; e < b; e += kWave) atomicAdd(&h_c[colg[col[e]]], 1);
  }
}

// extended sketches (mnc.cpp:640-656): h_er[u] counts entries whose column has count 1;
// h_ec[v] counts entries whose ROW id has column count 1 (the reference tests colSketch at u).
// A wave per row, as above.
__global__ void __launch_bounds__(kT) mnc_ext_kernel(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                                     int32_t nrows, const int32_t* __restrict__ rowg,
                                                     const int32_t* __restrict__ colg, const int32_t* __restrict__ h_c,
                                                     int32_t* __restrict__ h_er, int32_t* __restrict__ h_ec) {
  const int lane = lane_id();
  for (int64_t r = ((int64_t)blockIdx.x * kT + threadIdx.x) / kWave; r < nrows; r += (int64_t)gridDim.x * (kT / kWave)) {
    const int32_t u = rowg[r];
    const bool urow1 = h_c[u] == 1;
    int64_t er = 0;
    for (int64_t e = ptr[r] + lane; e < ptr[r + 1]; e += kWave) {
      const int32_t v = colg[col[e]];
      er += h_c[v] == 1;
      if (urow1) atomicAdd(&h_ec[v], 1);
    }
    er = wave_sum_i64(er);
    if (lane == 0 && er) atomicAdd(&h_er[u], (int32_t)er);
  }
}

// out: [0] nonzeros, [1] max, [2] entries == 1, [3] sum, [4] entries > thr
__global__ void __launch_bounds__(kT) mnc_stats_kernel(const int32_t* __restrict__ h, int64_t n, int32_t thr,
                                                       unsigned long long* __restrict__ out) {
  int64_t nz = 0, one = 0, sum = 0, gt = 0;
  int32_t mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const int32_t x = h[i];
    nz += x != 0;
    one += x == 1;
    sum += x;
    gt += x > thr;
    mx = max(mx, x);
  }
  nz = wave_sum_i64(nz);
  one = wave_sum_i64(one);
  sum = wave_sum_i64(sum);
  gt = wave_sum_i64(gt);
  mx = (int32_t)wave_max_i64(mx);
  if (lane_id() == 0) {
    atomicAdd(&out[0], (unsigned long long)nz);
    atomicMax(&out[1], (unsigned long long)mx);
    atomicAdd(&out[2], (unsigned long long)one);
    atomicAdd(&out[3], (unsigned long long)sum);
    atomicAdd(&out[4], (unsigned long long)gt);
  }
}

// MNCSketches::getSparsity terms (mnc.cpp:399-441).  mode 1: sum of the int32 products
// h_c(a)[j] * h_r(b)[j] (wrapping like the reference's int arithmetic); mode 2 (extended
// sketches): exact integer part and the union of lsp_j = (h_c - h1)(h_r - h2) / scale;
// mode 3: union of lsp_j = h_c(a)[j] h_r(b)[j] / scale.
// ipart[b], logs[b], prods[b], flags[b]: per-block partials (fixed block order on the host)
__global__ void __launch_bounds__(kT) mnc_terms_kernel(int mode, const int32_t* __restrict__ a_hc,
                                                       const int32_t* __restrict__ b_hr,
                                                       const int32_t* __restrict__ h1,
                                                       const int32_t* __restrict__ h2, int64_t n, double scale,
                                                       long long* __restrict__ ipart, double* __restrict__ logs,
                                                       double* __restrict__ prods, int* __restrict__ flags) {
  __shared__ double red[kT / kWave];
  __shared__ long long redi[kT / kWave];
  long long ip = 0;
  double lg = 0, pr = 1;
  int fl = 0;
  for (int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x; j < n; j += (int64_t)gridDim.x * kT) {
    const int32_t ac = a_hc[j], br = b_hr[j];
    if (mode == 1) {
      ip += (long long)(int32_t)((uint32_t)ac * (uint32_t)br);
      continue;
    }
    double x;
    if (mode == 2) {
      const int32_t e1 = h1[j], e2 = h2[j];
      ip += (long long)e1 * br + (long long)(ac - e1) * e2;
      x = (double)(ac - e1) * (double)(br - e2) / scale;
    } else {
      x = (double)ac * (double)br / scale;
    }
    if (x != 0.0) {
      if (!(x < 1.0)) fl = 1;
      else lg += log1p(-x);
      pr *= 1.0 - x;
    }
  }
  // integer part: exact wave sums, then per block
  long long w = ip;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
  if (lane_id() == 0) redi[threadIdx.x / kWave] = w;
  double p = pr;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) p *= __shfl_xor(p, o);
  const double lsum = block_sum_f64(lg, red);  // (contains barriers: redi is complete)
  if (lane_id() == 0) red[threadIdx.x / kWave] = p;
  int fw = fl;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) fw |= __shfl_xor(fw, o);
  __shared__ int fblk;
  if (threadIdx.x == 0) fblk = 0;
  __syncthreads();
  if (lane_id() == 0 && fw) atomicOr(&fblk, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    double pp = 1;
    for (int i = 0; i < kT / kWave; ++i) {
      t += redi[i];
      pp *= red[i];
    }
    ipart[blockIdx.x] = t;
    logs[blockIdx.x] = lsum;
    prods[blockIdx.x] = pp;
    flags[blockIdx.x] = fblk;
  }
}

// propagate*Sketch (mnc.cpp:495-548): out[i] = (int) round(v[i] * sparsity * m * l / sum(v))
__global__ void __launch_bounds__(kT) mnc_propagate_kernel(const int32_t* __restrict__ v, int64_t n, double sp,
                                                           double m, double l, double sum, int32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double x = round((double)v[i] * sp * m * l / sum);
    // a NaN or out-of-range value converts to INT_MIN as on the reference's x86 host
    out[i] = (x == x && x < 2147483648.0 && x >= -2147483648.0) ? (int32_t)x : INT32_MIN;
  }
}

// ---- DMap -------------------------------------------------------------------------------
// counts of the 256 x 256 blocks of the global matrix, over the map's block window
__global__ void __launch_bounds__(kT) dmap_count_kernel(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                                        int32_t nrows, const int32_t* __restrict__ rowg,
                                                        const int32_t* __restrict__ colg, int32_t r0, int32_t c0,
                                                        int32_t nc, double* __restrict__ map) {
  // a wave per row (lanes over its entries); counts are exact in binary64, so the order of the
  // additions does not matter
  const int lane = lane_id();
  for (int64_t r = ((int64_t)blockIdx.x * kT + threadIdx.x) / kWave; r < nrows; r += (int64_t)gridDim.x * (kT / kWave)) {
    const int64_t bi = rowg[r] / kDmapB - r0;
    for (int64_t e = ptr[r] + lane; e < ptr[r + 1]; e += kWave)
      atomicAdd(&map[bi * nc + (colg[col[e]] / kDmapB - c0)], 1.0);
  }
}

__global__ void __launch_bounds__(kT) dmap_scale_kernel(double* __restrict__ map, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    map[i] = map[i] / kDmapB / kDmapB;
}

// dense_map_estimator (dm.cpp:539-561): out[i][j] combines, in ascending k, every nonzero pair
// tmp1 = 1 - (1 - L[i][k] R[k][j])^256 as out = tmp1 + out - tmp1 * out.  One thread per
// (i, j); the left value is uniform across the block.
__global__ void __launch_bounds__(kT) dmap_product_kernel(const double* __restrict__ Lm, int32_t l_k0, int32_t l_nk,
                                                          const double* __restrict__ Rm, int32_t r_k0, int32_t r_nc,
                                                          int32_t k_lo, int32_t k_hi, int32_t nr, int32_t row0,
                                                          double* __restrict__ out) {
  const int32_t i = row0 + (int32_t)blockIdx.y;
  const int32_t j = blockIdx.x * kT + threadIdx.x;
  if (i >= nr) return;
  double res = 0;
  for (int32_t k = k_lo; k < k_hi; ++k) {
    const double sp1 = Lm[(int64_t)i * l_nk + (k - l_k0)];
    if (sp1 == 0) continue;
    if (j < r_nc) {
      const double sp2 = Rm[(int64_t)(k - r_k0) * r_nc + j];
      if (sp2 == 0) continue;
      const double t1 = 1.0 - pow(1.0 - sp1 * sp2, (double)kDmapB);
      res = t1 + res - t1 * res;
    }
  }
  if (j < r_nc) out[(int64_t)i * r_nc + j] = res;
}

__global__ void __launch_bounds__(kT) sum_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ partial) {
  __shared__ double red[kT / kWave];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) s += x[i];
  s = block_sum_f64(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---- LGraph -----------------------------------------------------------------------------
__device__ __forceinline__ double lg_rank(uint64_t seed, int32_t layer, int64_t u, int idx) {
  const uint64_t key = ((uint64_t)layer << 58) ^ ((uint64_t)u << 5) ^ (uint64_t)idx;
  uint64_t z = seed + (key + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const double uni = ((double)(z >> 11) + 1.0) * 0x1.0p-53;
  return -log(uni);
}

__global__ void __launch_bounds__(kT) lg_fill_kernel(double* __restrict__ R, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) R[i] = v;
}

// ranks of the leaf's row nodes: every row (all) or the rows with entries
__global__ void __launch_bounds__(kT) lg_ranks_kernel(const int64_t* __restrict__ ptr, int32_t nrows,
                                                      const int32_t* __restrict__ rowg, int all, uint64_t seed,
                                                      int32_t layer, double* __restrict__ R) {
  const int64_t total = (int64_t)nrows * kLgR;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int64_t r = t / kLgR;
    const int idx = (int)(t % kLgR);
    if (!all && ptr[r + 1] == ptr[r]) continue;
    const int64_t u = rowg[r];
    R[u * kLgR + idx] = lg_rank(seed, layer, u, idx);
  }
}

// row of every entry: a wave per row, lanes over its entries
__global__ void __launch_bounds__(kT) lg_rowid_kernel(const int64_t* __restrict__ ptr, int32_t nrows,
                                                      int32_t* __restrict__ rowid) {
  const int lane = threadIdx.x % kWave;
  for (int64_t r = ((int64_t)blockIdx.x * kT + threadIdx.x) / kWave; r < nrows; r += (int64_t)gridDim.x * (kT / kWave))
    for (int64_t e = ptr[r] + lane; e < ptr[r + 1]; e += kWave) rowid[e] = (int32_t)r;
}

// out[v] = min over entries (u, v) of in[u], elementwise over the 32 ranks (non-negative
// doubles order like their bit patterns).  A thread per (entry, rank), the ranks of an entry in
// consecutive lanes, so a hub row is spread over many threads and each wave's atomics touch two
// columns; the minimum does not depend on the order.
__global__ void __launch_bounds__(kT) lg_minprop_kernel(const int32_t* __restrict__ rowid,
                                                        const int32_t* __restrict__ col, int64_t nnz,
                                                        const int32_t* __restrict__ rowg,
                                                        const int32_t* __restrict__ colg, const double* __restrict__ in,
                                                        double* __restrict__ out) {
  const int64_t total = nnz * kLgR;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int64_t e = t / kLgR;
    const int idx = (int)(t % kLgR);
    const double x = in[(int64_t)rowg[rowid[e]] * kLgR + idx];
    atomicMin((unsigned long long*)&out[(int64_t)colg[col[e]] * kLgR + idx], (unsigned long long)__double_as_longlong(x));
  }
}

// mult[v] += 1 for every entry's column (the column-node list of distinct-type matrices)
__global__ void __launch_bounds__(kT) lg_indeg_kernel(const int32_t* __restrict__ col, int64_t nnz,
                                                      const int32_t* __restrict__ colg, int32_t* __restrict__ mult) {
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kT)
    atomicAdd(&mult[colg[col[e]]], 1);
}

// listed vertices (mult > 0) take src's ranks (copy) or keep theirs; unlisted ones get `big`
// (reset) or keep theirs
__global__ void __launch_bounds__(kT) lg_select_kernel(double* __restrict__ R, const double* __restrict__ src,
                                                       const int32_t* __restrict__ mult, int64_t n, double big) {
  const int64_t total = n * kLgR;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const bool listed = mult[t / kLgR] > 0;
    if (src) {
      if (listed) R[t] = src[t];
    } else if (!listed) {
      R[t] = big;
    }
  }
}

// getEstiNNZ (lg.cpp:111-124): sum over the column-node list of (r - 1) / sum of the ranks
__global__ void __launch_bounds__(kT) lg_estimate_kernel(const double* __restrict__ R, const int32_t* __restrict__ mult,
                                                         int64_t n, double* __restrict__ partial) {
  __shared__ double red[kT / kWave];
  double s = 0;
  for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n; v += (int64_t)gridDim.x * kT) {
    const int32_t c = mult[v];
    if (c <= 0) continue;
    double t = 0;
    for (int idx = 0; idx < kLgR; ++idx) t += R[v * kLgR + idx];
    s += (double)c * ((double)(kLgR - 1) / t);
  }
  s = block_sum_f64(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---- host helpers ---------------------------------------------------------------------------
unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kT), 4096)); }

double sum_partials(const double* dev, int nb, hipStream_t st) {
  std::vector<double> h(nb);
  SMCM_HIP(hipMemcpyAsync(h.data(), dev, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
  SMCM_HIP(hipStreamSynchronize(st));
  double s = 0;
  for (double x : h) s += x;
  return s;
}

double dsum(const double* x, int64_t n, DevAlloc& al) {
  DevBuf<double> part(al, kBlocks);
  sum_kernel<<<kBlocks, kT, 0, al.stream>>>(x, n, part.p);
  SMCM_LAUNCH_CHECK();
  return sum_partials(part.p, kBlocks, al.stream);
}

// ---- MetaAC -----------------------------------------------------------------------------
void run_metaac(const ProviderChain& ch, bool accuracy, ProviderOut& out, DevAlloc& al) {
  const int L = ch.L;
  const double n = (double)ch.n;
  std::vector<double> dens(L);
  DevBuf<double> part(al, kBlocks);
  for (int i = 0; i < L; ++i) {
    metaac_terms_kernel<<<kBlocks, kT, 0, al.stream>>>(ch.mats[i].ptr, ch.mats[i].nrows, n, part.p);
    SMCM_LAUNCH_CHECK();
    dens[i] = sum_partials(part.p, kBlocks, al.stream);
  }
  std::vector<double> sp((size_t)L * L, 0.0), m((size_t)L * L, 0.0);
  std::vector<int32_t> s((size_t)L * L, 0);
  const int N = ch.n;
  auto cost = [&](double a, double b, double c) {
    // compute_sparse_cost(a, b, c, n, n, n) (meta.cpp:493-501); int a_rows etc.
    const double tA = N * (double)(a * N);
    const double tB = ((((double)N * N) * a) * N) * b;
    const double tC = (double)(N * c) * N;
    return accuracy ? tA + tB + tC : kAlph * tA + kBt * tB + kGmm * tC;
  };
  for (int l = 1; l < L; ++l)
    for (int i = 0; i < L - l; ++i) {
      const int j = i + l;
      m[i * L + j] = DBL_MAX;
      for (int k = i; k < j; ++k) {
        const double a = sp[i * L + k] != 0 ? sp[i * L + k] : dens[i];
        const double b = sp[(k + 1) * L + j] != 0 ? sp[(k + 1) * L + j] : dens[j];
        const double r = 1.0 - pow(1.0 - a * b, (double)N);
        const double c = m[i * L + k] + m[(k + 1) * L + j] + cost(a, b, r);
        if (c < m[i * L + j]) {
          m[i * L + j] = c;
          s[i * L + j] = k;
          sp[i * L + j] = r;
        }
      }
    }
  out.c = sp;
  out.m = m;
  out.s = s;
  out.estimate = sp[L - 1] * n * n;  // sparsities[len - 1] * hn * hn (accuracy/meta.cpp:546, :721)
}

// ---- MNC --------------------------------------------------------------------------------
struct Sk {
  DevBuf<int32_t> h_r, h_c, h_er, h_ec;
  bool ext = false;
  int64_t r_nnz = 0, c_nnz = 0, max_r = 0, max_c = 0, rN1 = 0, cN1 = 0, sum_r = 0, sum_c = 0, r_gt = 0, c_gt = 0;
  long double es_nnz = 0;
};

void sk_stats(Sk& k, int64_t n, DevAlloc& al) {
  DevBuf<unsigned long long> st(al, 10);
  fill_i64((int64_t*)st.p, 10, 0, al.stream);
  mnc_stats_kernel<<<grid_for(n), kT, 0, al.stream>>>(k.h_r.p, n, (int32_t)(n / 2), st.p);
  mnc_stats_kernel<<<grid_for(n), kT, 0, al.stream>>>(k.h_c.p, n, (int32_t)(n / 2), st.p + 5);
  SMCM_LAUNCH_CHECK();
  unsigned long long h[10];
  SMCM_HIP(hipMemcpyAsync(h, st.p, sizeof(h), hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  k.r_nnz = (int64_t)h[0];
  k.max_r = (int64_t)h[1];
  k.rN1 = (int64_t)h[2];
  k.sum_r = (int64_t)h[3];
  k.r_gt = (int64_t)h[4];
  k.c_nnz = (int64_t)h[5];
  k.max_c = (int64_t)h[6];
  k.cN1 = (int64_t)h[7];
  k.sum_c = (int64_t)h[8];
  k.c_gt = (int64_t)h[9];
}

void mnc_leaf(const ProviderChain& ch, int i, Sk& k, DevAlloc& al) {
  const int64_t n = ch.n;
  const Csr& M = ch.mats[i];
  k.h_r = DevBuf<int32_t>(al, n);
  k.h_c = DevBuf<int32_t>(al, n);
  k.h_er = DevBuf<int32_t>(al, n);
  k.h_ec = DevBuf<int32_t>(al, n);
  fill_i32(k.h_r.p, n, 0, al.stream);
  fill_i32(k.h_c.p, n, 0, al.stream);
  fill_i32(k.h_er.p, n, 0, al.stream);
  fill_i32(k.h_ec.p, n, 0, al.stream);
  if (M.nrows > 0) {
    mnc_counts_kernel<<<grid_for((int64_t)M.nrows * kWave), kT, 0, al.stream>>>(M.ptr, M.col, M.nrows, ch.rowg[i], ch.colg[i], k.h_r.p,
                                                              k.h_c.p);
    mnc_ext_kernel<<<grid_for((int64_t)M.nrows * kWave), kT, 0, al.stream>>>(M.ptr, M.col, M.nrows, ch.rowg[i], ch.colg[i], k.h_c.p,
                                                           k.h_er.p, k.h_ec.p);
    SMCM_LAUNCH_CHECK();
  }
  k.ext = true;
  sk_stats(k, n, al);
}

struct Terms {
  long long ip = 0;
  double logsum = 0, prod = 1;
  bool big = false;
};

Terms mnc_terms(int mode, const Sk& a, const Sk& b, double scale, int64_t n, DevAlloc& al) {
  DevBuf<long long> ip(al, kBlocks);
  DevBuf<double> lg(al, kBlocks), pr(al, kBlocks);
  DevBuf<int> fl(al, kBlocks);
  mnc_terms_kernel<<<kBlocks, kT, 0, al.stream>>>(mode, a.h_c.p, b.h_r.p, mode == 2 ? a.h_ec.p : nullptr,
                                                  mode == 2 ? a.h_er.p : nullptr, n, scale, ip.p, lg.p, pr.p, fl.p);
  SMCM_LAUNCH_CHECK();
  std::vector<long long> hip_(kBlocks);
  std::vector<double> hl(kBlocks), hp(kBlocks);
  std::vector<int> hf(kBlocks);
  SMCM_HIP(hipMemcpyAsync(hip_.data(), ip.p, sizeof(long long) * kBlocks, hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipMemcpyAsync(hl.data(), lg.p, sizeof(double) * kBlocks, hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipMemcpyAsync(hp.data(), pr.p, sizeof(double) * kBlocks, hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipMemcpyAsync(hf.data(), fl.p, sizeof(int) * kBlocks, hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  Terms t;
  for (int q = 0; q < kBlocks; ++q) {
    t.ip += hip_[q];
    t.logsum += hl[q];
    t.prod *= hp[q];
    t.big |= hf[q] != 0;
  }
  return t;
}

// 1 - prod(1 - lsp_j): the reference's sequential union recurrence in closed form
long double union_of(const Terms& t) {
  if (!t.big) return (long double)(-expm1(t.logsum));
  return 1.0L - (long double)t.prod;
}

inline int32_t wrap32(int64_t x) { return (int32_t)(uint32_t)(uint64_t)x; }

// MNCSketches::getSparsity(a, b, n, n, n), mnc.cpp:399-441
long double mnc_sparsity(const Sk& a, const Sk& b, int64_t n, DevAlloc& al) {
  long double nnz;
  if (a.max_r <= 1 || b.max_c <= 1) {
    nnz = (long double)mnc_terms(1, a, b, 1.0, n, al).ip;
  } else if (a.ext && b.ext) {
    const long double mn_out = (long double)wrap32((a.r_nnz - a.rN1) * (b.c_nnz - b.cN1));
    const Terms t = mnc_terms(2, a, b, (double)mn_out, n, al);
    nnz = (long double)t.ip + union_of(t) * mn_out;
  } else {
    const long double p = (long double)a.r_nnz * (long double)b.c_nnz;
    const Terms t = mnc_terms(3, a, b, (double)p, n, al);
    nnz = union_of(t) * p;
  }
  const long double larger = (long double)wrap32(a.r_gt * b.c_gt);
  nnz = nnz < larger ? larger : nnz;
  return (nnz / (long double)n) * 1 / (long double)n;
}

void mnc_propagate(const Sk& a, const Sk& b, int64_t n, Sk& out, DevAlloc& al) {
  const long double sp = mnc_sparsity(a, b, n, al);
  out.es_nnz = sp * (long double)n * (long double)n;
  out.h_r = DevBuf<int32_t>(al, n);
  out.h_c = DevBuf<int32_t>(al, n);
  out.h_er = DevBuf<int32_t>();
  out.h_ec = DevBuf<int32_t>();
  out.ext = false;
  const double sd = (double)sp;
  mnc_propagate_kernel<<<grid_for(n), kT, 0, al.stream>>>(a.h_r.p, n, sd, (double)n, (double)n, (double)a.sum_r,
                                                          out.h_r.p);
  mnc_propagate_kernel<<<grid_for(n), kT, 0, al.stream>>>(b.h_c.p, n, sd, (double)n, (double)n, (double)b.sum_c,
                                                          out.h_c.p);
  SMCM_LAUNCH_CHECK();
  sk_stats(out, n, al);
  out.rN1 = out.cN1 = 0;  // not counted for propagated sketches
}

void run_mnc(const ProviderChain& ch, bool accuracy, ProviderOut& out, DevAlloc& al) {
  const int L = ch.L;
  const int64_t n = ch.n;
  std::vector<Sk> E((size_t)L * L);
  for (int i = 0; i < L; ++i) mnc_leaf(ch, i, E[i * L + i], al);
  std::vector<long double> m((size_t)L * L, 0.0L);
  std::vector<int32_t> s((size_t)L * L, 0);
  std::vector<double> c((size_t)L * L, 0.0);
  for (int l = 1; l < L; ++l)
    for (int i = 0; i < L - l; ++i) {
      const int j = i + l;
      m[i * L + j] = LDBL_MAX;
      bool have = false;
      for (int k = i; k < j; ++k) {
        const Sk& a = E[i * L + k];
        const Sk& b = E[(k + 1) * L + j];
        Sk cand;
        long double tmp;
        if (accuracy) {
          tmp = (long double)mnc_terms(1, a, b, 1.0, n, al).ip;  // dot of the sketches (accuracy/mnc.cpp:794)
        } else {
          mnc_propagate(a, b, n, cand, al);
          tmp = cand.es_nnz;
        }
        const long double cost = m[i * L + k] + m[(k + 1) * L + j] + tmp;
        if (cost < m[i * L + j]) {
          m[i * L + j] = cost;
          s[i * L + j] = k;
          if (accuracy) mnc_propagate(a, b, n, cand, al);
          E[i * L + j] = std::move(cand);
          have = true;
        }
      }
      // every cost NaN: the reference keeps a null sketch here (and fails on it later); the
      // product at split i stands in
      if (!have) mnc_propagate(E[i * L + i], E[(i + 1) * L + j], n, E[i * L + j], al);
      c[i * L + j] = (double)E[i * L + j].es_nnz;
    }
  out.m.assign((size_t)L * L, 0.0);
  for (size_t q = 0; q < m.size(); ++q) out.m[q] = (double)m[q];
  out.s = s;
  out.c = c;
  out.estimate = L > 1 ? (double)E[L - 1].es_nnz : 0.0;
}

// ---- DMap -------------------------------------------------------------------------------
struct BMap {
  int32_t r0 = 0, nr = 0, c0 = 0, nc = 0;
  DevBuf<double> d;
};

void dmap_leaf(const ProviderChain& ch, int i, BMap& b, DevAlloc& al) {
  const Csr& M = ch.mats[i];
  b.r0 = ch.blo[i];
  b.nr = ch.bhi[i] - ch.blo[i] + 1;
  b.c0 = ch.blo[i + 1];
  b.nc = ch.bhi[i + 1] - ch.blo[i + 1] + 1;
  const int64_t cells = (int64_t)b.nr * b.nc;
  b.d = DevBuf<double>(al, cells);
  fill_i64((int64_t*)b.d.p, cells, 0, al.stream);
  if (M.nrows > 0) {
    dmap_count_kernel<<<grid_for((int64_t)M.nrows * kWave), kT, 0, al.stream>>>(M.ptr, M.col, M.nrows, ch.rowg[i], ch.colg[i], b.r0,
                                                              b.c0, b.nc, b.d.p);
    SMCM_LAUNCH_CHECK();
  }
  dmap_scale_kernel<<<grid_for(cells), kT, 0, al.stream>>>(b.d.p, cells);
  SMCM_LAUNCH_CHECK();
}

void dmap_product(const BMap& A, const BMap& B, BMap& out, DevAlloc& al) {
  out.r0 = A.r0;
  out.nr = A.nr;
  out.c0 = B.c0;
  out.nc = B.nc;
  const int64_t cells = (int64_t)out.nr * out.nc;
  out.d = DevBuf<double>(al, cells);
  const int32_t k_lo = std::max(A.c0, B.r0), k_hi = std::min(A.c0 + A.nc, B.r0 + B.nr);
  if (k_hi <= k_lo || cells == 0) {
    fill_i64((int64_t*)out.d.p, cells, 0, al.stream);
    return;
  }
  for (int32_t row0 = 0; row0 < out.nr; row0 += 65535) {  // grid.y limit
    dim3 grid((unsigned)ceil_div(out.nc, kT), (unsigned)std::min<int32_t>(65535, out.nr - row0));
    dmap_product_kernel<<<grid, kT, 0, al.stream>>>(A.d.p, A.c0, A.nc, B.d.p, B.r0, B.nc, k_lo, k_hi, out.nr, row0,
                                                    out.d.p);
    SMCM_LAUNCH_CHECK();
  }
}

void run_dmap(const ProviderChain& ch, ProviderOut& out, DevAlloc& al) {
  const int L = ch.L;
  const int64_t bn = ceil_div(ch.n, kDmapB);
  std::vector<BMap> leaf(L);
  for (int i = 0; i < L; ++i) dmap_leaf(ch, i, leaf[i], al);
  std::vector<double> c((size_t)L * L, 0.0);
  auto density = [&](const BMap& b) { return dsum(b.d.p, (int64_t)b.nr * b.nc, al) / (double)bn / (double)bn; };
  std::vector<BMap> row_prev;  // maps (i+1, j) for j > i+1
  for (int i = L - 1; i >= 0; --i) {
    c[i * L + i] = density(leaf[i]);
    std::vector<BMap> row_cur(L);
    for (int j = i + 1; j < L; ++j) {
      const BMap& R = (j == i + 1) ? leaf[i + 1] : row_prev[j];
      dmap_product(leaf[i], R, row_cur[j], al);
      c[i * L + j] = density(row_cur[j]);
    }
    row_prev = std::move(row_cur);
  }
  std::vector<double> m((size_t)L * L, 0.0);
  std::vector<int32_t> s((size_t)L * L, 0);
  const double n = (double)ch.n;
  for (int l = 1; l < L; ++l)
    for (int i = 0; i < L - l; ++i) {
      const int j = i + l;
      m[i * L + j] = DBL_MAX;
      for (int k = i; k < j; ++k) {
        const double tmp = kAlph * c[i * L + k] + kBt / ch.n * c[i * L + k] * c[(k + 1) * L + j] + kGmm * c[i * L + j];
        const double cost = m[i * L + k] + m[(k + 1) * L + j] + tmp;
        if (cost < m[i * L + j]) {
          m[i * L + j] = cost;
          s[i * L + j] = k;
        }
      }
    }
  out.c = c;
  out.m = m;
  out.s = s;
  out.estimate = c[L - 1] * n * n;  // c_[len - 1] * hn * hn (accuracy/dm.cpp:892)
}

// ---- LGraph -----------------------------------------------------------------------------
// the column-node list of buildLayeredGraph as multiplicities: distinct row and column types
// list every entry's column (in-degree); otherwise the list depends on the row flags
// (lg.cpp:150-166) and is built on the host
void lg_column_list(const ProviderChain& ch, int i, bool accuracy, DevBuf<int32_t>& mult, DevAlloc& al) {
  const int64_t n = ch.n;
  const Csr& M = ch.mats[i];
  mult = DevBuf<int32_t>(al, n);
  fill_i32(mult.p, n, 0, al.stream);
  if (accuracy) {
    // every vertex of the column type, once (accuracy/lg.cpp:638)
    std::vector<int32_t> h(n, 0);
    for (int32_t v : ch.col_type_ids[i]) h[v] = 1;
    SMCM_HIP(hipMemcpyAsync(mult.p, h.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, al.stream));
    SMCM_HIP(hipStreamSynchronize(al.stream));
    return;
  }
  if (!ch.same_type[i]) {
    if (M.nnz > 0) {
      lg_indeg_kernel<<<grid_for(M.nnz), kT, 0, al.stream>>>(M.col, M.nnz, ch.colg[i], mult.p);
      SMCM_LAUNCH_CHECK();
    }
    return;
  }
  std::vector<int64_t> ptr(M.nrows + 1);
  std::vector<int32_t> col(M.nnz), rg(M.nrows), cg(ch.col_type_ids[i].begin(), ch.col_type_ids[i].end());
  SMCM_HIP(hipMemcpyAsync(ptr.data(), M.ptr, sizeof(int64_t) * ptr.size(), hipMemcpyDeviceToHost, al.stream));
  if (M.nnz) SMCM_HIP(hipMemcpyAsync(col.data(), M.col, sizeof(int32_t) * M.nnz, hipMemcpyDeviceToHost, al.stream));
  if (M.nrows) SMCM_HIP(hipMemcpyAsync(rg.data(), ch.rowg[i], sizeof(int32_t) * M.nrows, hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  std::vector<int32_t> order(M.nrows);
  for (int32_t r = 0; r < M.nrows; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return rg[x] < rg[y]; });
  std::vector<char> flag(n, 0);
  std::vector<int32_t> h(n, 0);
  for (int32_t r : order) {
    const int32_t u = rg[r];
    for (int64_t e = ptr[r]; e < ptr[r + 1]; ++e) {
      const int32_t v = cg[col[e]];
      if (!flag[v]) {
        h[v] += 1;
        flag[u] = 1;
      }
    }
  }
  SMCM_HIP(hipMemcpyAsync(mult.p, h.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
}

double lg_estimate(const double* R, const int32_t* mult, int64_t n, DevAlloc& al) {
  DevBuf<double> part(al, kBlocks);
  lg_estimate_kernel<<<kBlocks, kT, 0, al.stream>>>(R, mult, n, part.p);
  SMCM_LAUNCH_CHECK();
  return sum_partials(part.p, kBlocks, al.stream);
}

void run_lgraph(const ProviderChain& ch, bool accuracy, ProviderOut& out, DevAlloc& al) {
  const int L = ch.L;
  const int64_t n = ch.n;
  const int64_t cells = n * kLgR;
  std::vector<DevBuf<int32_t>> mult(L), rowid(L);
  for (int i = 0; i < L; ++i) {
    lg_column_list(ch, i, accuracy, mult[i], al);
    const Csr& M = ch.mats[i];
    rowid[i] = DevBuf<int32_t>(al, M.nnz);
    if (M.nnz > 0) {
      lg_rowid_kernel<<<grid_for((int64_t)M.nrows * kWave), kT, 0, al.stream>>>(M.ptr, M.nrows, rowid[i].p);
      SMCM_LAUNCH_CHECK();
    }
  }
  std::vector<long double> c((size_t)L * L, 0.0L);
  DevBuf<double> R(al, cells), T(al, cells), U(al, cells);
  for (int i = 0; i < L; ++i) {
    const Csr& M = ch.mats[i];
    // leaf: ranks at the row nodes, then each listed column takes the minimum of its rows'
    lg_fill_kernel<<<grid_for(cells), kT, 0, al.stream>>>(R.p, cells, DBL_MAX);
    lg_fill_kernel<<<grid_for(cells), kT, 0, al.stream>>>(T.p, cells, DBL_MAX);
    if (M.nrows > 0) {
      lg_ranks_kernel<<<grid_for((int64_t)M.nrows * kLgR), kT, 0, al.stream>>>(M.ptr, M.nrows, ch.rowg[i],
                                                                                accuracy ? 1 : 0, ch.seed, i, R.p);
      if (M.nnz > 0)
        lg_minprop_kernel<<<grid_for(M.nnz * kLgR), kT, 0, al.stream>>>(rowid[i].p, M.col, M.nnz, ch.rowg[i], ch.colg[i],
                                                                 R.p, T.p);
    }
    lg_select_kernel<<<grid_for(cells), kT, 0, al.stream>>>(R.p, T.p, mult[i].p, n, DBL_MAX);
    SMCM_LAUNCH_CHECK();
    if (!accuracy) c[i * L + i] = lg_estimate(R.p, mult[i].p, n, al);
    // chain i .. j: ranks propagated through M_j (only the listed columns in the efficient
    // variant, lg.cpp:671-679)
    double* left = R.p;
    for (int j = i + 1; j < L; ++j) {
      const Csr& Mj = ch.mats[j];
      double* dst = left == T.p ? U.p : T.p;
      lg_fill_kernel<<<grid_for(cells), kT, 0, al.stream>>>(dst, cells, DBL_MAX);
      if (Mj.nnz > 0)
        lg_minprop_kernel<<<grid_for(Mj.nnz * kLgR), kT, 0, al.stream>>>(rowid[j].p, Mj.col, Mj.nnz, ch.rowg[j], ch.colg[j],
                                                                  left, dst);
      if (!accuracy) lg_select_kernel<<<grid_for(cells), kT, 0, al.stream>>>(dst, nullptr, mult[j].p, n, DBL_MAX);
      SMCM_LAUNCH_CHECK();
      c[i * L + j] = lg_estimate(dst, mult[j].p, n, al);
      left = dst;
    }
  }
  std::vector<long double> m((size_t)L * L, 0.0L);
  std::vector<int32_t> s((size_t)L * L, 0);
  for (int l = 1; l < L; ++l)
    for (int i = 0; i < L - l; ++i) {
      const int j = i + l;
      m[i * L + j] = LDBL_MAX;
      for (int k = i; k < j; ++k) {
        const long double tmp = c[i * L + k] + c[(k + 1) * L + j] + c[i * L + j];
        const long double cost = m[i * L + k] + m[(k + 1) * L + j] + tmp;
        if (cost < m[i * L + j]) {
          m[i * L + j] = cost;
          s[i * L + j] = k;
        }
      }
    }
  out.c.assign((size_t)L * L, 0.0);
  out.m.assign((size_t)L * L, 0.0);
  for (size_t q = 0; q < c.size(); ++q) {
    out.c[q] = (double)c[q];
    out.m[q] = (double)m[q];
  }
  out.s = s;
  out.estimate = (double)c[L - 1];
}

}  // namespace

void run_provider(int provider, bool accuracy, const ProviderChain& ch, ProviderOut& out, DevAlloc& al) {
  if (ch.L < 1) throw HipError(-22, "cost provider: empty chain");
  switch (provider) {
    case kProviderMetaAC: run_metaac(ch, accuracy, out, al); break;
    case kProviderMNC: run_mnc(ch, accuracy, out, al); break;
    case kProviderDMap: run_dmap(ch, out, al); break;
    case kProviderLGraph: run_lgraph(ch, accuracy, out, al); break;
    default: throw HipError(-22, "unknown cost provider");
  }
}

}  // namespace smcm
