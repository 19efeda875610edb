// estimator.hip - row-wise sparsity estimator (one pass per chain position i, all j > i fused).
//
// Reference: BoolMatrix::getSparsities (efficient/resmcm.cpp:442-452) and
// DynamicOptimizer::compute_sparsities (:468-485), scheduled by :506-517:
//   r_ii[u] = min(1, deg_u(M_i) / dims[t_{i+1}])
//   r_ij[u] = 1 - prod_{v in row_u(M_i)} (1 - r_{i+1,j}[v])      (product in row order)
//   c_ij    = sum_u r_ij[u] * dims[t_{j+1}]
// Every r_ij[u] is computed with the reference's operation order in IEEE binary64 (the file is
// built with -ffp-contract=off), so per-row densities are bit-identical to a one-thread run of
// the reference.  The product for one row is a dependent chain and cannot be reassociated; rows
// of degree <= kLightDeg use one lane each.  Longer rows take two steps: every factor (1 - r) of
// their entries is gathered in parallel into a chain-major buffer (items of kItem entries), then
// one wave per row streams its factors through LDS, several chunks in flight, and lane j multiplies
// chain j through them in row order.
// Once all running products are exactly 0 they stay 0, so a row stops early - this is exact.
//
// c_ij is reduced in a fixed tree order (deterministic run to run); exact_sum instead sums in
// row order on one lane, reproducing the reference's one-thread sum bit for bit.
#include "kernels.h"

namespace smcm {

namespace {

constexpr int kThreads = 256;
constexpr int kLightDeg = 32;
constexpr int kItem = 2048;  // entries per factor item of a heavy row
constexpr int kMaxW = 15;

template <int W>
__global__ void __launch_bounds__(kThreads) est_light_kernel(const int64_t* __restrict__ ptr,
                                                             const int32_t* __restrict__ col, int32_t nrows,
                                                             double leaf_dim, const double* __restrict__ Rn,
                                                             double* __restrict__ Ri, int32_t* __restrict__ heavy,
                                                             int32_t* __restrict__ nheavy, int2* __restrict__ items,
                                                             int32_t* __restrict__ nitems) {
  const int32_t u = blockIdx.x * kThreads + threadIdx.x;
  if (u >= nrows) return;
  const int64_t p0 = ptr[u], p1 = ptr[u + 1];
  const int64_t deg = p1 - p0;
  double s = (double)deg / leaf_dim;
  double* out = Ri + (size_t)u * (W + 1);
  out[0] = s > 1 ? 1 : s;
  if constexpr (W > 0) {
    if (deg > kLightDeg) {
      heavy[atomicAdd(nheavy, 1)] = u;
      const int32_t nit = (int32_t)((deg + kItem - 1) / kItem);
      const int32_t base = atomicAdd(nitems, nit);
      for (int32_t k = 0; k < nit; ++k) items[base + k] = make_int2(u, k);
      return;
    }
    double acc[W];
#pragma unroll
    for (int j = 0; j < W; ++j) acc[j] = 1;
#pragma unroll 4
    for (int64_t e = p0; e < p1; ++e) {
      const double* r = Rn + (size_t)col[e] * W;
#pragma unroll
      for (int j = 0; j < W; ++j) acc[j] *= (1 - r[j]);
    }
#pragma unroll
    for (int j = 0; j < W; ++j) out[1 + j] = 1 - acc[j];
  }
}

// factors of heavy rows: Fh[j * nnz + e] = 1 - R_{i+1}[col[e]][j], one block per item (kItem
// consecutive entries of one row), every gather independent
template <int W>
__global__ void __launch_bounds__(kThreads) est_factor_kernel(const int64_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ col, int64_t nnz,
                                                              const double* __restrict__ Rn,
                                                              const int2* __restrict__ items,
                                                              const int32_t* __restrict__ nitems,
                                                              double* __restrict__ Fh) {
  constexpr int kPer = kItem / kThreads;
  const int32_t nit = *nitems;
  for (int32_t it = blockIdx.x; it < nit; it += gridDim.x) {
    const int2 item = items[it];
    const int64_t p1 = ptr[item.x + 1];
    const int64_t b = ptr[item.x] + (int64_t)item.y * kItem + threadIdx.x;
    int32_t c[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int64_t x = b + q * kThreads;
      c[q] = x < p1 ? col[x] : -1;
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      if (c[q] < 0) continue;
      const int64_t x = b + q * kThreads;
      const double* r = Rn + (size_t)c[q] * W;
#pragma unroll
      for (int j = 0; j < W; ++j) Fh[(size_t)j * nnz + x] = 1 - r[j];
    }
  }
}

// one wave per heavy row: the wave loads chunks of 64 entries' factors (chain-major, coalesced,
// kDepth chunks in flight in registers) and stages one chunk at a time in LDS; lane j < W
// multiplies chain j through the chunk in row order (the reference's order).  Entries past the
// row's end load as 1.0, which leaves a product unchanged.
template <int W>
__global__ void __launch_bounds__(kThreads) est_chain_kernel(const int64_t* __restrict__ ptr, int64_t nnz,
                                                             const double* __restrict__ Fh,
                                                             double* __restrict__ Ri,
                                                             const int32_t* __restrict__ heavy,
                                                             const int32_t* __restrict__ nheavy) {
  constexpr int kDepth = 24 / W < 2 ? 2 : (24 / W > 8 ? 8 : 24 / W);
  __shared__ __align__(16) double F[kThreads / kWave][W * kWave];
  const int lane = lane_id();
  const int32_t nh = *nheavy;
  const int32_t wpb = kThreads / kWave;
  double* fw = F[threadIdx.x / kWave];
  for (int32_t h = blockIdx.x * wpb + threadIdx.x / kWave; h < nh; h += gridDim.x * wpb) {
    const int32_t u = uniform_i32(heavy[h]);
    const int64_t p0 = uniform_i64(ptr[u]), p1 = uniform_i64(ptr[u + 1]);
    const int64_t nch = (p1 - p0 + kWave - 1) / kWave;
    double r[kDepth][W];
#pragma unroll
    for (int d = 0; d < kDepth; ++d) {
      const int64_t x = p0 + (int64_t)d * kWave + lane;
#pragma unroll
      for (int j = 0; j < W; ++j) r[d][j] = x < p1 ? Fh[(size_t)j * nnz + x] : 1.0;
    }
    double acc = 1.0;  // lane j < W: the running product of chain j
    bool done = false;
    for (int64_t c0 = 0; c0 < nch && !done; c0 += kDepth) {
#pragma unroll
      for (int d = 0; d < kDepth; ++d) {
        if (c0 + d >= nch) {
          done = true;
          break;
        }
#pragma unroll
        for (int j = 0; j < W; ++j) fw[j * kWave + lane] = r[d][j];
        {
          const int64_t x = p0 + (c0 + d + kDepth) * kWave + lane;
#pragma unroll
          for (int j = 0; j < W; ++j) r[d][j] = x < p1 ? Fh[(size_t)j * nnz + x] : 1.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < W) {
          const double2* q = (const double2*)(fw + lane * kWave);
          double2 v[kWave / 2];
#pragma unroll
          for (int k = 0; k < kWave / 2; ++k) v[k] = q[k];
#pragma unroll
          for (int k = 0; k < kWave / 2; ++k) {
            acc *= v[k].x;
            acc *= v[k].y;
          }
        }
        // every chain exactly 0: it stays 0 (exact early exit)
        const bool live = lane < W && acc != 0.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (__ballot(live) == 0) {
          done = true;
          break;
        }
      }
    }
    if (lane < W) Ri[(size_t)u * (W + 1) + 1 + lane] = 1 - acc;
  }
}

struct DimJ {
  double d[kMaxW > 0 ? kMaxW : 1];
};

template <int W>
__global__ void __launch_bounds__(kThreads) est_partial_kernel(const double* __restrict__ Ri, int32_t nrows,
                                                               const DimJ dimj, double* __restrict__ partial) {
  __shared__ double sh[kThreads];
  const int32_t u = blockIdx.x * kThreads + threadIdx.x;
  for (int j = 0; j < W; ++j) {
    double x = 0.0;
    if (u < nrows) {
      double r = Ri[(size_t)u * (W + 1) + 1 + j];
      x = r * dimj.d[j];
    }
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * W + j] = sh[0];
    __syncthreads();
  }
}

// c_j = sum of block partials in block order (one thread per j)
__global__ void est_final_kernel(const double* __restrict__ partial, int32_t nblocks, int W, double* __restrict__ c) {
  const int j = threadIdx.x;
  if (j >= W) return;
  double s = 0.0;
  for (int32_t b = 0; b < nblocks; ++b) s = s + partial[(size_t)b * W + j];
  c[j] = s;
}

// exact mode: c_j = sequential sum over rows in row order, starting from 0 (one lane per j)
__global__ void est_exact_kernel(const double* __restrict__ Ri, int32_t nrows, int W, const DimJ dimj,
                                 double* __restrict__ c) {
  const int j = threadIdx.x;
  if (j >= W) return;
  const double d = dimj.d[j];
  double s = 0.0;
  for (int32_t u = 0; u < nrows; ++u) {
    double x = Ri[(size_t)u * (W + 1) + 1 + j] * d;
    s = s + x;
  }
  c[j] = s;
}


// ---- product-cost estimate ---------------------------------------------------------------
// Column analogue of the row estimator, for the cost of a product rather than its size.  For
// the sub-chain X_ik = M_i ... M_k (rows: those of M_i), rho_ik[v] is the estimated fraction of
// rows of X_ik that have column v:
//   rho_ii[v] = min(1, colcount_{M_i}[v] / rows_i)
//   rho_ik[v] = 1 - prod_{w : M_k[w,v] != 0} (1 - rho_{i,k-1}[w])      (independence, as above)
// and the products of X_ik * X_{k+1,j} are estimated as
//   P(i,k,j) = rows_i * sum_v rho_ik[v] * r_{k+1,j}[v] * dims[t_{j+1}].
// For two leaves this is sum_v colcount[v] * min(deg[v], dim), the exact product count when no
// row holds more entries than the column type has vertices.  The product over w is accumulated
// as a sum of logarithms in 2^-32 fixed point with integer atomics, so it does not depend on the
// order of the additions; the dot products use a fixed tree order.  Results are deterministic.
constexpr double kFx = 4294967296.0;
constexpr double kLogFloor = -64.0;

__device__ inline int64_t fx_log1m(double rho) {
  double l = log1p(-rho);
  if (!(l > kLogFloor)) l = kLogFloor;
  return (int64_t)llrint(l * kFx);
}

// one thread per entry of M_k: column v of the entry collects, for every start i < k, the
// logarithm of (1 - rho_{i,k-1}[w]) of the entry's row w, and the entry count for i = k
__global__ void __launch_bounds__(kThreads) colest_scatter_kernel(const int64_t* __restrict__ ptr,
                                                                  const int32_t* __restrict__ col, int32_t nrows,
                                                                  int64_t nnz, const int64_t* __restrict__ lin, int nin,
                                                                  int64_t* __restrict__ S) {
  const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= nnz) return;
  int32_t lo = 0, hi = nrows;  // row w: ptr[w] <= e < ptr[w+1]
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (ptr[mid] <= e) lo = mid;
    else hi = mid;
  }
  const int32_t w = lo;
  const int64_t v = col[e];
  unsigned long long* out = (unsigned long long*)(S + v * (nin + 1));
  for (int i = 0; i < nin; ++i) atomicAdd(out + i, (unsigned long long)lin[(int64_t)w * nin + i]);
  atomicAdd(out + nin, 1ull);
}

// blockIdx.y = start i (0..k): rho_ik[v] for this block's columns, the next pass's logarithms,
// and per-block partial sums of rows_i * rho_ik[v] * r_{k+1,j}[v] * dims[t_{j+1}]
__global__ void __launch_bounds__(kThreads) colest_finalize_kernel(const int64_t* __restrict__ S, int32_t ncols, int k,
                                                                   const DimJ nrows_i, const double* __restrict__ Rk1,
                                                                   int Wj, const DimJ dim_j, int64_t* __restrict__ lout,
                                                                   double* __restrict__ partial) {
  __shared__ double sh[kThreads];
  const int i = blockIdx.y;
  const int32_t v = blockIdx.x * kThreads + threadIdx.x;
  const int nk = k + 1;
  double rho = 0.0;
  if (v < ncols) {
    const int64_t x = S[(int64_t)v * nk + i];
    if (i == k) {
      const double r = nrows_i.d[i] > 0 ? (double)x / nrows_i.d[i] : 0.0;
      rho = r > 1 ? 1 : r;
    } else {
      rho = -expm1((double)x / kFx);
    }
    if (lout) lout[(int64_t)v * nk + i] = fx_log1m(rho);
  }
  for (int jj = 0; jj < Wj; ++jj) {
    double y = 0.0;
    if (v < ncols) y = nrows_i.d[i] * rho * (Rk1[(int64_t)v * Wj + jj] * dim_j.d[jj]);
    sh[threadIdx.x] = y;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[((int64_t)i * gridDim.x + blockIdx.x) * Wj + jj] = sh[0];
    __syncthreads();
  }
}

// P[(i*L + k)*L + k+1+jj] = sum of block partials in block order
__global__ void colest_final_kernel(const double* __restrict__ partial, int32_t nblocks, int nk, int Wj, int k, int L,
                                    double* __restrict__ P) {
  const int t = threadIdx.x;
  if (t >= nk * Wj) return;
  const int i = t / Wj, jj = t % Wj;
  double s = 0.0;
  for (int32_t b = 0; b < nblocks; ++b) s = s + partial[((int64_t)i * nblocks + b) * Wj + jj];
  P[((int64_t)i * L + k) * L + k + 1 + jj] = s;
}

template <int W>
void run_pass(const EstimatorPass& p, DevAlloc& al) {
  const Csr& M = *p.M;
  const int32_t n = M.nrows;
  hipStream_t st = al.stream;
  // heavy rows have more than kLightDeg entries: fewer than nnz / (kLightDeg + 1) rows, and
  // fewer than nnz / kItem + (heavy rows) items
  const int64_t cap_items = M.nnz / kItem + M.nnz / (kLightDeg + 1) + 1;
  DevBuf<int32_t> heavy(al, n > 0 ? n : 1), counts(al, 2);
  DevBuf<int2> items(al, W > 0 ? cap_items : 1);
  int32_t* nheavy = counts.p;
  int32_t* nitems = counts.p + 1;
  fill_i32(counts.p, 2, 0, st);
  const unsigned blocks = (unsigned)ceil_div(n > 0 ? n : 1, kThreads);
  // algorithmic bytes of the pass: row pointers, column ids, W gathered densities per entry,
  // W+1 densities written per row
  const double bytes = 16.0 * (n + 1) + 4.0 * (double)M.nnz + 8.0 * W * (double)M.nnz + 8.0 * (W + 1) * n;
  const int tok = p.tm ? p.tm->start("estimator.pass", st) : -1;
  if (n > 0) {
    est_light_kernel<W><<<blocks, kThreads, 0, st>>>(M.ptr, M.col, n, p.leaf_dim, p.Rn, p.Ri, heavy.p, nheavy,
                                                     items.p, nitems);
    SMCM_LAUNCH_CHECK();
  }
  if constexpr (W > 0) {
    if (n > 0 && M.nnz > kLightDeg) {
      DevBuf<double> Fh(al, (int64_t)W * M.nnz);
      const unsigned fblocks = (unsigned)(cap_items < 4096 ? cap_items : 4096);
      est_factor_kernel<W><<<fblocks, kThreads, 0, st>>>(M.ptr, M.col, M.nnz, p.Rn, items.p, nitems, Fh.p);
      SMCM_LAUNCH_CHECK();
      est_chain_kernel<W><<<2048, kThreads, 0, st>>>(M.ptr, M.nnz, Fh.p, p.Ri, heavy.p, nheavy);
      SMCM_LAUNCH_CHECK();
    }
    if (p.tm) p.tm->stop(tok, st, bytes);
    DimJ dimj{};
    for (int j = 0; j < W; ++j) dimj.d[j] = p.dimj[j];
    if (p.exact_sum || n == 0) {
      est_exact_kernel<<<1, 64, 0, st>>>(p.Ri, n, W, dimj, p.c_dev);
      SMCM_LAUNCH_CHECK();
    } else {
      DevBuf<double> partial(al, (int64_t)blocks * W);
      est_partial_kernel<W><<<blocks, kThreads, 0, st>>>(p.Ri, n, dimj, partial.p);
      SMCM_LAUNCH_CHECK();
      est_final_kernel<<<1, 64, 0, st>>>(partial.p, (int32_t)blocks, W, p.c_dev);
      SMCM_LAUNCH_CHECK();
    }
  } else {
    if (p.tm) p.tm->stop(tok, st, bytes);
  }
  if (p.tm) {
    SMCM_HIP(hipStreamSynchronize(st));
    p.tm->flush();
  }
}

}  // namespace

void estimator_pass(const EstimatorPass& p, DevAlloc& al) {
  switch (p.W) {
#define SMCM_W(w) \
  case w:         \
    run_pass<w>(p, al); \
    break;
    SMCM_W(0) SMCM_W(1) SMCM_W(2) SMCM_W(3) SMCM_W(4) SMCM_W(5) SMCM_W(6) SMCM_W(7)
    SMCM_W(8) SMCM_W(9) SMCM_W(10) SMCM_W(11) SMCM_W(12) SMCM_W(13) SMCM_W(14) SMCM_W(15)
#undef SMCM_W
    default:
      throw HipError(-22, "estimator: chains longer than 16 matrices are not supported");
  }
}

void product_estimate(const ProductEstimate& pe, DevAlloc& al) {
  const int L = pe.L;
  if (L < 2) return;
  if (L - 1 > kMaxW) throw HipError(-22, "product estimate: chains longer than 16 matrices are not supported");
  hipStream_t st = al.stream;
  DimJ rows{};
  for (int i = 0; i < L - 1; ++i) rows.d[i] = (double)pe.mats[i].nrows;
  DevBuf<int64_t> lin;
  const int tok = pe.tm ? pe.tm->start("estimator.products", st) : -1;
  double bytes = 0;
  for (int k = 0; k <= L - 2; ++k) {
    const Csr& M = pe.mats[k];
    const int nk = k + 1;
    const int32_t nc = M.ncols;
    const int Wj = L - 1 - k;
    DevBuf<int64_t> S(al, (int64_t)nc * nk);
    fill_i64(S.p, (int64_t)nc * nk, 0, st);
    if (M.nnz > 0) {
      colest_scatter_kernel<<<(unsigned)ceil_div(M.nnz, kThreads), kThreads, 0, st>>>(M.ptr, M.col, M.nrows, M.nnz,
                                                                                   lin.p, k, S.p);
      SMCM_LAUNCH_CHECK();
    }
    DevBuf<int64_t> lout;
    if (k + 1 <= L - 2) lout = DevBuf<int64_t>(al, (int64_t)nc * nk);
    DimJ dj{};
    for (int jj = 0; jj < Wj; ++jj) dj.d[jj] = pe.dims[k + 2 + jj];
    const int32_t nb = (int32_t)ceil_div(nc > 0 ? nc : 1, kThreads);
    DevBuf<double> partial(al, (int64_t)nk * nb * Wj);
    colest_finalize_kernel<<<dim3((unsigned)nb, (unsigned)nk), kThreads, 0, st>>>(S.p, nc, k, rows, pe.R[k + 1], Wj, dj,
                                                                                 lout.p, partial.p);
    SMCM_LAUNCH_CHECK();
    colest_final_kernel<<<1, 256, 0, st>>>(partial.p, nb, nk, Wj, k, L, pe.p_dev);
    SMCM_LAUNCH_CHECK();
    lin = std::move(lout);
    bytes += 16.0 * (M.nrows + 1) + (4.0 + 8.0 * k) * (double)M.nnz + 8.0 * nk * nc * (2 + Wj);
  }
  if (pe.tm) {
    pe.tm->stop(tok, st, bytes);
    SMCM_HIP(hipStreamSynchronize(st));
    pe.tm->flush();
  }
}

}  // namespace smcm
