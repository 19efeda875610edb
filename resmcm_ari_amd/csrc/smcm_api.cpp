// smcm_api.cpp - C ABI (include/smcm.h) and the host-side chain executor.
//
// The executor restates SSP_dynamic (efficient/resmcm.cpp:653-772) for a device-resident graph:
//   1. typed matrices of the meta-path (built on the device once per (ts, et, td) and cached;
//      not timed, as in the reference where they are built before the timer, :664-680);
//   2. estimator passes i = L-1..0 on the device (es_time, :506-517);
//   3. chain-order DP on the host (dp_time, :519-538) and plan extraction (:563-573);
//   4. the plan's SpGEMMs on the device with the reference's operand stack (:709-760);
//      consumed intermediates are released immediately (:718-757).
// total = estimator + DP + products (smcm_time, :679-771); all phases end synchronised, times
// are host wall-clock seconds.
#include "smcm.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "kernels.h"

using smcm::Csr;
using smcm::DevAlloc;
using smcm::DevBuf;
using smcm::HipError;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct ApiError {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw ApiError{code, msg}; }

template <class F>
int guarded(F&& f) {
  try {
    f();
    return SMCM_OK;
  } catch (const ApiError& e) {
    return set_err(e.code, e.msg);
  } catch (const HipError& e) {
    if (e.code == -22) return set_err(SMCM_EINVAL, e.what());
    if (e.code == (int)hipErrorOutOfMemory) return set_err(SMCM_ENOMEM, e.what());
    return set_err(SMCM_EDEVICE, e.what());
  } catch (const std::bad_alloc&) {
    return set_err(SMCM_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return set_err(SMCM_EDEVICE, e.what());
  }
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void free_csr(DevAlloc& al, Csr& c) {
  al.free(c.ptr);
  al.free(c.col);
  al.free(c.val);
  c = Csr{};
}

}  // namespace

struct smcm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  DevAlloc al;
  uint64_t budget = 0;
  smcm::SpgemmOpts opts;
  bool exact_cost_sum = false;
  int plan_cost = 0;  // 0: the reference DP over estimated sizes; 1: estimated products + sizes
  smcm_stats last{};
  int64_t live = 0;  // graphs and results created from this context and not yet freed
  smcm::KernelTimer timer;
};

struct smcm_graph {
  smcm_ctx* ctx = nullptr;
  int32_t n = 0;
  int64_t m = 0;
  int32_t ntypes = 0;
  int32_t* vtype = nullptr;
  int64_t* adj_ptr = nullptr;
  int32_t* adj_nbr = nullptr;
  int32_t* adj_et = nullptr;
  int32_t* local = nullptr;
  std::vector<int32_t> h_vtype, h_local, dims;
  std::vector<int32_t*> bins;
  std::map<std::tuple<int32_t, int32_t, int32_t>, Csr> cache;
};

struct smcm_result {
  smcm_ctx* ctx = nullptr;
  Csr C;
  int32_t* col_map = nullptr;  // device: local column id -> global id (null = identity)
  std::vector<int32_t> row_ids;
};

namespace {

void check_ctx(smcm_ctx* c) {
  if (!c) fail(SMCM_EINVAL, "null context");
  SMCM_HIP(hipSetDevice(c->device));
}

// host-side fallback that sorts every row of a (base) matrix whose adjacency order is not
// ascending; used only for the right operand of a product, never for the estimator
Csr sorted_copy(DevAlloc& al, const Csr& a) {
  std::vector<int64_t> ptr(a.nrows + 1);
  std::vector<int32_t> col(a.nnz);
  SMCM_HIP(hipMemcpyAsync(ptr.data(), a.ptr, sizeof(int64_t) * ptr.size(), hipMemcpyDeviceToHost, al.stream));
  if (a.nnz) SMCM_HIP(hipMemcpyAsync(col.data(), a.col, sizeof(int32_t) * a.nnz, hipMemcpyDeviceToHost, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  for (int32_t r = 0; r < a.nrows; ++r) std::sort(col.begin() + ptr[r], col.begin() + ptr[r + 1]);
  Csr s = a;
  s.ptr = al.alloc_n<int64_t>(a.nrows + 1);
  s.col = al.alloc_n<int32_t>(a.nnz);
  s.val = nullptr;
  s.sorted = true;
  SMCM_HIP(hipMemcpyAsync(s.ptr, ptr.data(), sizeof(int64_t) * ptr.size(), hipMemcpyHostToDevice, al.stream));
  if (a.nnz) SMCM_HIP(hipMemcpyAsync(s.col, col.data(), sizeof(int32_t) * a.nnz, hipMemcpyHostToDevice, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  return s;
}

// typed relation matrix (ts --et--> td) over all rows of ts, cached on the graph
const Csr& typed(smcm_graph* g, int32_t ts, int32_t et, int32_t td) {
  auto key = std::make_tuple(ts, et, td);
  auto it = g->cache.find(key);
  if (it != g->cache.end()) return it->second;
  DevAlloc& al = g->ctx->al;
  Csr c;
  smcm::build_typed_csr(g->vtype, g->adj_ptr, g->adj_nbr, g->adj_et, g->local, g->bins[ts], nullptr, g->dims[ts],
                        et, td, g->dims[td], c, al);
  return g->cache.emplace(key, c).first->second;
}

// sorted version of a cached typed matrix (for right operands); cached under et + 1<<20
const Csr& typed_sorted(smcm_graph* g, int32_t ts, int32_t et, int32_t td) {
  const Csr& c = typed(g, ts, et, td);
  if (c.sorted) return c;
  auto key = std::make_tuple(ts, et | (1 << 20), td);
  auto it = g->cache.find(key);
  if (it != g->cache.end()) return it->second;
  Csr s = sorted_copy(g->ctx->al, c);
  return g->cache.emplace(key, s).first->second;
}

struct Chain {
  smcm_ctx* ctx;
  smcm_graph* g;
  int32_t L = 0;
  std::vector<int32_t> vts, ets;
  std::vector<int32_t> src_local;  // local ids of the source rows (empty = all rows)
  bool subset = false;
  std::vector<Csr> mats;           // estimator operands (file order)
  Csr m0_subset;                   // owned copy of M_0 restricted to the source rows
  DevBuf<int32_t> d_src;
  std::vector<DevBuf<double>> R;   // R[i]: rows_i x (L-i)
  std::vector<double> c;           // L*L
  std::vector<double> ptab;        // L*L*L estimated products per split (plan_cost 1)
  bool want_products = false;
  std::vector<double> mtab;
  std::vector<int32_t> stab, pairs;
  double es_time = 0, dp_time = 0, spgemm_time = 0, total = 0;

  ~Chain() {
    if (m0_subset.ptr) free_csr(ctx->al, m0_subset);
  }

  void prepare(const int32_t* vtypes, const int32_t* etypes, int32_t len, const int32_t* src_rows, int64_t n_src) {
    if (!g) fail(SMCM_EINVAL, "null graph");
    if (!vtypes || !etypes) fail(SMCM_EINVAL, "null meta-path");
    if (len < 1 || len > 16) fail(SMCM_EINVAL, "the meta-path is incorrect: L must be in [1, 16]");
    L = len;
    want_products = ctx->plan_cost == 1;
    vts.assign(vtypes, vtypes + L + 1);
    ets.assign(etypes, etypes + L);
    for (int32_t t : vts)
      if (t < 0 || t >= g->ntypes) fail(SMCM_EINVAL, "the meta-path is incorrect: vertex type not in graph");
    for (int32_t e : ets)
      if (e < 0) fail(SMCM_EINVAL, "the meta-path is incorrect: negative edge type");
    DevAlloc& al = ctx->al;
    mats.resize(L);
    for (int32_t i = 0; i < L; ++i) mats[i] = typed(g, vts[i], ets[i], vts[i + 1]);
    if (src_rows) {
      if (n_src < 0 || n_src > INT32_MAX) fail(SMCM_EINVAL, "bad source row count");
      subset = true;
      src_local.resize(n_src);
      for (int64_t r = 0; r < n_src; ++r) {
        const int32_t v = src_rows[r];
        if (v < 0 || v >= g->n || g->h_vtype[v] != vts[0])
          fail(SMCM_EINVAL, "source row " + std::to_string(v) + " is not a vertex of the start type");
        src_local[r] = g->h_local[v];
      }
      d_src = DevBuf<int32_t>(al, n_src);
      if (n_src)
        SMCM_HIP(hipMemcpyAsync(d_src.p, src_local.data(), sizeof(int32_t) * n_src, hipMemcpyHostToDevice, al.stream));
      smcm::csr_gather_rows(mats[0], d_src.p, (int32_t)n_src, m0_subset, al);
      mats[0] = m0_subset;
    }
  }

  void estimate() {
    DevAlloc& al = ctx->al;
    const double t0 = now_s();
    c.assign((size_t)L * L, 0.0);
    R.clear();
    R.resize(L);
    DevBuf<double> cdev(al, (int64_t)L * L);
    smcm::fill_i64((int64_t*)cdev.p, (int64_t)L * L, 0, al.stream);  // +0.0 bit pattern
    for (int32_t i = L - 1; i >= 0; --i) {
      const int W = L - 1 - i;
      R[i] = DevBuf<double>(al, (int64_t)mats[i].nrows * (W + 1));
      std::vector<double> dimj(W > 0 ? W : 1);
      for (int q = 0; q < W; ++q) dimj[q] = (double)g->dims[vts[i + 2 + q]];
      smcm::EstimatorPass p;
      p.M = &mats[i];
      p.leaf_dim = (double)g->dims[vts[i + 1]];
      p.Rn = W > 0 ? R[i + 1].p : nullptr;
      p.W = W;
      p.dimj = dimj.data();
      p.Ri = R[i].p;
      p.c_dev = cdev.p + (size_t)i * L + i + 1;
      p.exact_sum = ctx->exact_cost_sum;
      p.tm = ctx->timer.enabled ? &ctx->timer : nullptr;
      smcm::estimator_pass(p, al);
    }
    DevBuf<double> pdev;
    if (want_products && L > 1) {
      const int64_t n3 = (int64_t)L * L * L;
      pdev = DevBuf<double>(al, n3);
      smcm::fill_i64((int64_t*)pdev.p, n3, 0, al.stream);
      std::vector<const double*> rp(L, nullptr);
      for (int32_t q = 1; q < L; ++q) rp[q] = R[q].p;
      std::vector<double> dq(L + 1);
      for (int32_t q = 0; q <= L; ++q) dq[q] = (double)g->dims[vts[q]];
      smcm::ProductEstimate pe;
      pe.mats = mats.data();
      pe.L = L;
      pe.R = rp.data();
      pe.dims = dq.data();
      pe.p_dev = pdev.p;
      pe.tm = ctx->timer.enabled ? &ctx->timer : nullptr;
      smcm::product_estimate(pe, al);
      ptab.assign((size_t)n3, 0.0);
      SMCM_HIP(hipMemcpyAsync(ptab.data(), pdev.p, sizeof(double) * n3, hipMemcpyDeviceToHost, al.stream));
    }
    SMCM_HIP(hipMemcpyAsync(c.data(), cdev.p, sizeof(double) * c.size(), hipMemcpyDeviceToHost, al.stream));
    SMCM_HIP(hipStreamSynchronize(al.stream));
    es_time = now_s() - t0;
  }

  // resmcm.cpp:519-538 and :563-573
  void dp() {
    const double t0 = now_s();
    mtab.assign((size_t)L * L, 0.0);
    stab.assign((size_t)L * L, 0);
    if (want_products && !ptab.empty()) {
      // cost(i,j) = min_k cost(i,k) + cost(k+1,j) + P(i,k,j), plus the estimated size of every
      // intermediate (not of the root); same evaluation order and lowest-k tie rule as the
      // reference DP.  Any plan gives the same product (exact integer arithmetic), so this only
      // changes the amount of work.
      for (int32_t l = 1; l < L; ++l) {
        for (int32_t i = 0; i < L - l; ++i) {
          const int32_t j = i + l;
          double best = DBL_MAX;
          for (int32_t k = i; k < j; ++k) {
            const double cost = mtab[i * L + k] + mtab[(k + 1) * L + j] + ptab[((size_t)i * L + k) * L + j];
            if (cost < best) {
              best = cost;
              stab[i * L + j] = k;
            }
          }
          mtab[i * L + j] = best + (l < L - 1 ? c[i * L + j] : 0.0);
        }
      }
      pairs.clear();
      if (L > 1) plan_rec(0, L - 1);
      dp_time = now_s() - t0;
      return;
    }
    for (int32_t l = 1; l < L; ++l) {
      for (int32_t i = 0; i < L - l; ++i) {
        const int32_t j = i + l;
        mtab[i * L + j] = DBL_MAX;
        for (int32_t k = i; k < j; ++k) {
          const double tmp = c[i * L + k] + c[(k + 1) * L + j] + c[i * L + j];
          const double cost = mtab[i * L + k] + mtab[(k + 1) * L + j] + tmp;
          if (cost < mtab[i * L + j]) {
            mtab[i * L + j] = cost;
            stab[i * L + j] = k;
          }
        }
      }
    }
    pairs.clear();
    if (L > 1) plan_rec(0, L - 1);
    dp_time = now_s() - t0;
  }

  int32_t plan_rec(int32_t i, int32_t j) {
    if (i == j) return i;
    const int32_t s = stab[i * L + j];
    const int32_t k = plan_rec(i, s);
    const int32_t l = plan_rec(s + 1, j);
    pairs.push_back(k);
    pairs.push_back(l);
    return -1;
  }

  void check_plan(const int32_t* p) {
    // simulate the operand stack: every leaf used once, in order, intervals adjacent
    std::vector<std::pair<int32_t, int32_t>> st;
    std::vector<bool> used(L, false);
    for (int32_t q = 0; q < L - 1; ++q) {
      const int32_t k = p[2 * q], l = p[2 * q + 1];
      auto leaf = [&](int32_t x) {
        if (x < 0 || x >= L || used[x]) fail(SMCM_EINVAL, "invalid plan: bad leaf index");
        used[x] = true;
        return std::make_pair(x, x);
      };
      std::pair<int32_t, int32_t> a, b;
      if (k >= 0 && l >= 0) {
        a = leaf(k);
        b = leaf(l);
        st.push_back({a.first, b.second});
      } else if (k < 0 && l >= 0) {
        if (st.empty()) fail(SMCM_EINVAL, "invalid plan: empty stack");
        a = st.back();
        b = leaf(l);
        st.back() = {a.first, b.second};
      } else if (k >= 0 && l < 0) {
        if (st.empty()) fail(SMCM_EINVAL, "invalid plan: empty stack");
        a = leaf(k);
        b = st.back();
        st.back() = {a.first, b.second};
      } else {
        if (st.size() < 2) fail(SMCM_EINVAL, "invalid plan: stack underflow");
        a = st[st.size() - 2];
        b = st.back();
        st.pop_back();
        st.back() = {a.first, b.second};
      }
      if (a.second + 1 != b.first) fail(SMCM_EINVAL, "invalid plan: operands are not adjacent");
    }
    if (st.size() != 1 || st[0].first != 0 || st[0].second != L - 1) fail(SMCM_EINVAL, "invalid plan");
  }

  struct Opd {
    Csr m;
    int32_t i, j;
    bool owned;
  };

  // resmcm.cpp:709-760 with device products; returns the product (owned by the caller).
  // With a consumer, the root product is materialised in row blocks of at most block_bytes
  // and handed to the consumer (the returned Csr is then empty).
  Csr execute(const std::vector<int32_t>& plan, smcm_stats* stats, const smcm::BlockConsumer* consumer = nullptr,
              int64_t block_bytes = 0) {
    DevAlloc& al = ctx->al;
    const double t0 = now_s();
    auto base = [&](int32_t x, bool right) -> Opd {
      if (right) {
        const Csr& s = (x == 0 && subset) ? mats[0] : typed_sorted(g, vts[x], ets[x], vts[x + 1]);
        if (x == 0 && subset && !s.sorted) fail(SMCM_EINVAL, "unsorted source matrix as right operand");
        return Opd{s, x, x, false};
      }
      return Opd{mats[x], x, x, false};
    };
    std::vector<Opd> st;
    smcm::KernelTimer* tm = ctx->timer.enabled ? &ctx->timer : nullptr;
    const size_t nprod = plan.size() / 2;
    size_t done_products = 0;
    auto product = [&](Opd a, Opd b) -> Opd {
      Csr C;
      smcm::SpgemmStats ss;
      const int32_t i = a.i, j = b.j;
      const double* est = R[i].p + (j - i);
      const bool root = ++done_products == nprod;
      if (root && consumer)
        smcm::spgemm_blocked(a.m, b.m, est, L - i, block_bytes, *consumer, ctx->opts, &ss, al, tm);
      else
        smcm::spgemm(a.m, b.m, est, L - i, C, ctx->opts, &ss, al, true, tm);
      if (stats) {
        stats->products += ss.products;
        for (int t = 0; t < 4; ++t) stats->rows_tier[t] += ss.rows[t];
        stats->n_spgemm += 1;
      }
      if (a.owned) free_csr(al, a.m);
      if (b.owned) free_csr(al, b.m);
      return Opd{C, i, j, true};
    };
    try {
      for (size_t q = 0; q + 1 < plan.size() + 1 && 2 * q < plan.size(); ++q) {
        const int32_t k = plan[2 * q], l = plan[2 * q + 1];
        if (k >= 0 && l >= 0) {
          st.push_back(product(base(k, false), base(l, true)));
        } else if (k < 0 && l >= 0) {
          Opd top = st.back();
          st.pop_back();
          st.push_back(product(top, base(l, true)));
        } else if (k >= 0 && l < 0) {
          Opd top = st.back();
          st.pop_back();
          st.push_back(product(base(k, false), top));
        } else {
          Opd r = st.back();
          st.pop_back();
          Opd lft = st.back();
          st.pop_back();
          st.push_back(product(lft, r));
        }
      }
    } catch (...) {
      for (auto& o : st)
        if (o.owned) free_csr(al, o.m);
      throw;
    }
    spgemm_time = now_s() - t0;
    return st.at(0).m;
  }
};

Csr identity_csr(DevAlloc& al, int32_t n) {
  std::vector<int64_t> ptr(n + 1);
  std::vector<int32_t> col(n);
  for (int32_t i = 0; i <= n; ++i) ptr[i] = i;
  for (int32_t i = 0; i < n; ++i) col[i] = i;
  Csr I;
  I.nrows = I.ncols = n;
  I.nnz = n;
  I.ptr = al.alloc_n<int64_t>(n + 1);
  I.col = al.alloc_n<int32_t>(n);
  SMCM_HIP(hipMemcpyAsync(I.ptr, ptr.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, al.stream));
  if (n) SMCM_HIP(hipMemcpyAsync(I.col, col.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, al.stream));
  SMCM_HIP(hipStreamSynchronize(al.stream));
  return I;
}

smcm_result* make_result(Chain& ch, Csr C) {
  smcm_ctx* ctx = ch.ctx;
  auto* r = new smcm_result;
  r->ctx = ctx;
  r->C = C;
  ctx->live++;
  const int32_t tcol = ch.vts[ch.L];
  const int32_t dcol = ch.g->dims[tcol];
  r->col_map = ctx->al.alloc_n<int32_t>(dcol);
  if (dcol)
    SMCM_HIP(hipMemcpyAsync(r->col_map, ch.g->bins[tcol], sizeof(int32_t) * dcol, hipMemcpyDeviceToDevice,
                            ctx->al.stream));
  const int32_t t0 = ch.vts[0];
  if (ch.subset) {
    r->row_ids.resize(ch.src_local.size());
    std::vector<int32_t> b(ch.g->dims[t0]);
    // global id of local row x of type t0
    for (int32_t v = 0; v < ch.g->n; ++v)
      if (ch.g->h_vtype[v] == t0) b[ch.g->h_local[v]] = v;
    for (size_t q = 0; q < ch.src_local.size(); ++q) r->row_ids[q] = b[ch.src_local[q]];
  } else {
    r->row_ids.reserve(ch.g->dims[t0]);
    for (int32_t v = 0; v < ch.g->n; ++v)
      if (ch.g->h_vtype[v] == t0) r->row_ids.push_back(v);
  }
  SMCM_HIP(hipStreamSynchronize(ctx->al.stream));
  return r;
}

// Runs a chain: estimator, DP (or the forced plan) and the products.  Without a consumer the
// product is returned in *C; with one, the root product is streamed in row blocks.
void run_core(Chain& ch, const int32_t* forced_pairs, Csr* C, const smcm::BlockConsumer* consumer,
              int64_t block_bytes, smcm_times* times) {
  smcm_ctx* ctx = ch.ctx;
  const int32_t L = ch.L;
  smcm_stats st{};
  smcm::KernelTimer* tm = ctx->timer.enabled ? &ctx->timer : nullptr;
  Csr out;
  if (L == 1) {
    // one matrix: merge duplicates and sort by multiplying with the identity
    Csr I = identity_csr(ctx->al, ch.mats[0].ncols);
    smcm::SpgemmStats ss;
    const double t0 = now_s();
    try {
      if (consumer)
        smcm::spgemm_blocked(ch.mats[0], I, nullptr, 0, block_bytes, *consumer, ctx->opts, &ss, ctx->al, tm);
      else
        smcm::spgemm(ch.mats[0], I, nullptr, 0, out, ctx->opts, &ss, ctx->al, true, tm);
    } catch (...) {
      free_csr(ctx->al, I);
      throw;
    }
    free_csr(ctx->al, I);
    ch.spgemm_time = now_s() - t0;
    st.products = ss.products;
    st.nnz = ss.nnz;
    st.ref_nnz = ch.mats[0].nnz;
    st.n_spgemm = 0;
  } else {
    const double t0 = now_s();
    if (forced_pairs) ch.want_products = false;
    ch.estimate();
    ch.dp();
    std::vector<int32_t> plan = ch.pairs;
    if (forced_pairs) {
      ch.check_plan(forced_pairs);
      plan.assign(forced_pairs, forced_pairs + 2 * (L - 1));
    }
    out = ch.execute(plan, &st, consumer, block_bytes);
    ch.total = now_s() - t0;
    if (std::getenv("SMCM_TRACE"))
      fprintf(stderr, "[smcm] chain L=%d: est %.3f dp %.3f spgemm %.3f total %.3f ms\n", L, ch.es_time * 1e3,
              ch.dp_time * 1e3, ch.spgemm_time * 1e3, ch.total * 1e3);
    for (size_t q = 0; q < plan.size() && q < 30; ++q) st.pairs[q] = plan[q];
    st.c_root = ch.c[L - 1];
    if (!consumer) st.nnz = out.nnz;
  }
  if (!consumer) st.nnz = out.nnz;
  if (L > 1) st.ref_nnz = st.nnz;
  ctx->last = st;
  if (times) {
    times->total = ch.total;
    times->spgemm = ch.spgemm_time;
    times->dp = ch.dp_time;
    times->est = ch.es_time;
  }
  if (C) *C = out;
}

int run_chain(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
              const int32_t* src_rows, int64_t n_src, const int32_t* forced_pairs, smcm_result** out,
              smcm_times* times) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out) fail(SMCM_EINVAL, "null output");
    *out = nullptr;
    Chain ch{ctx, g};
    ch.prepare(vtypes, etypes, L, src_rows, n_src);
    Csr C;
    run_core(ch, forced_pairs, &C, nullptr, 0, times);
    try {
      *out = make_result(ch, C);
    } catch (...) {
      free_csr(ctx->al, C);
      throw;
    }
  });
}

}  // namespace

extern "C" {

const char* smcm_last_error(void) { return g_err.c_str(); }

int smcm_ctx_create(int device, uint64_t hbm_budget, smcm_ctx** out) {
  return guarded([&] {
    if (!out) fail(SMCM_EINVAL, "null output");
    *out = nullptr;
    int ndev = 0;
    SMCM_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) fail(SMCM_EINVAL, "no such device");
    SMCM_HIP(hipSetDevice(device));
    auto ctx = std::make_unique<smcm_ctx>();
    ctx->device = device;
    SMCM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->al.stream = ctx->stream;
    // keep freed blocks in the device's default pool
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
      uint64_t thr = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    (void)hipGetLastError();
    int lds = 65536, cus = 256;
    SMCM_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    SMCM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    ctx->opts.lds_bytes = lds;
    ctx->opts.cus = cus;
    size_t freeb = 0, totalb = 0;
    SMCM_HIP(hipMemGetInfo(&freeb, &totalb));
    ctx->budget = hbm_budget ? hbm_budget : (uint64_t)(0.9 * (double)freeb);
    *out = ctx.release();
  });
}

int smcm_ctx_destroy(smcm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    if (ctx->live != 0)
      fail(SMCM_EINVAL, std::to_string(ctx->live) + " graph/result handle(s) still alive; free them first");
    (void)hipSetDevice(ctx->device);
    ctx->al.release_cache();
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int smcm_ctx_sync(smcm_ctx* ctx) {
  return guarded([&] {
    check_ctx(ctx);
    SMCM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int smcm_set_option(smcm_ctx* ctx, const char* key, int64_t v) {
  return guarded([&] {
    if (!ctx || !key) fail(SMCM_EINVAL, "null argument");
    std::string k(key);
    auto& o = ctx->opts;
    if (k == "s_max") {
      if (v < 0 || v > 64) fail(SMCM_EINVAL, "s_max must be in [0, 64]");
      o.s_max = (int)v;
    } else if (k == "m_max") {
      if (v < 64 || v > 4096) fail(SMCM_EINVAL, "m_max must be in [64, 4096]");
      o.m_max = (int)v;
    } else if (k == "dense_div") {
      if (v < 1) fail(SMCM_EINVAL, "dense_div must be >= 1");
      o.dense_div = (int)v;
    } else if (k == "heavy_ub") {
      o.heavy_ub = v;
    } else if (k == "d_threads") {
      if (v != 256 && v != 512 && v != 1024) fail(SMCM_EINVAL, "d_threads must be 256, 512 or 1024");
      o.d_threads = (int)v;
    } else if (k == "w_threads") {
      if (v != 256 && v != 512 && v != 1024) fail(SMCM_EINVAL, "w_threads must be 256, 512 or 1024");
      o.w_threads = (int)v;
    } else if (k == "window_cols") {
      if (v != 0 && (v < 64 || (v & (v - 1)) != 0)) fail(SMCM_EINVAL, "window_cols must be 0 or a power of two >= 64");
      o.window_cols = v;
    } else if (k == "bitmap_vcap") {
      if (v < -1) fail(SMCM_EINVAL, "bitmap_vcap must be -1 or >= 0");
      o.bitmap_vcap = v;
    } else if (k == "force_tier") {
      if (v < -1 || v > 2) fail(SMCM_EINVAL, "force_tier must be -1..2");
      o.force_tier = (int)v;
    } else if (k == "exact_cost_sum") {
      ctx->exact_cost_sum = v != 0;
    } else if (k == "plan_cost") {
      if (v != 0 && v != 1) fail(SMCM_EINVAL, "plan_cost must be 0 (reference) or 1 (products)");
      ctx->plan_cost = (int)v;
    } else if (k == "kernel_timing") {
      ctx->timer.enabled = v != 0;
    } else if (k == "lds_bytes") {
      if (v < 16384) fail(SMCM_EINVAL, "lds_bytes too small");
      int dev_max = 65536;
      SMCM_HIP(hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
      if (v > dev_max) fail(SMCM_EINVAL, "lds_bytes above the device limit");
      o.lds_bytes = v;
    } else {
      fail(SMCM_EINVAL, "unknown option " + k);
    }
  });
}

int smcm_get_option(smcm_ctx* ctx, const char* key, int64_t* v) {
  return guarded([&] {
    if (!ctx || !key || !v) fail(SMCM_EINVAL, "null argument");
    std::string k(key);
    auto& o = ctx->opts;
    if (k == "s_max") *v = o.s_max;
    else if (k == "m_max") *v = o.m_max;
    else if (k == "dense_div") *v = o.dense_div;
    else if (k == "heavy_ub") *v = o.heavy_ub;
    else if (k == "d_threads") *v = o.d_threads;
    else if (k == "w_threads") *v = o.w_threads;
    else if (k == "window_cols") *v = o.window_cols;
    else if (k == "bitmap_vcap") *v = o.bitmap_vcap;
    else if (k == "force_tier") *v = o.force_tier;
    else if (k == "exact_cost_sum") *v = ctx->exact_cost_sum;
    else if (k == "plan_cost") *v = ctx->plan_cost;
    else if (k == "kernel_timing") *v = ctx->timer.enabled;
    else if (k == "lds_bytes") *v = o.lds_bytes;
    else if (k == "cus") *v = o.cus;
    else if (k == "budget") *v = (int64_t)ctx->budget;
    else fail(SMCM_EINVAL, "unknown option " + k);
  });
}

int smcm_last_stats(smcm_ctx* ctx, smcm_stats* out) {
  return guarded([&] {
    if (!ctx || !out) fail(SMCM_EINVAL, "null argument");
    *out = ctx->last;
  });
}

int smcm_graph_upload(smcm_ctx* ctx, int32_t n, const int32_t* vtype, const int64_t* adj_ptr,
                      const int32_t* adj_nbr, const int32_t* adj_etype, smcm_graph** out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out) fail(SMCM_EINVAL, "null output");
    *out = nullptr;
    if (n < 0 || (n > 0 && (!vtype || !adj_ptr))) fail(SMCM_EINVAL, "null graph arrays");
    const int64_t m = n > 0 ? adj_ptr[n] : 0;
    if (n > 0 && adj_ptr[0] != 0) fail(SMCM_EINVAL, "adj_ptr[0] must be 0");
    if (m > 0 && (!adj_nbr || !adj_etype)) fail(SMCM_EINVAL, "null adjacency arrays");
    auto g = std::make_unique<smcm_graph>();
    g->ctx = ctx;
    g->n = n;
    g->m = m;
    g->h_vtype.assign(vtype, vtype + n);
    int32_t maxt = -1;
    for (int32_t v = 0; v < n; ++v) {
      if (vtype[v] < 0 || vtype[v] >= 4096) fail(SMCM_EINVAL, "vertex type out of range");
      maxt = std::max(maxt, vtype[v]);
      if (adj_ptr[v + 1] < adj_ptr[v]) fail(SMCM_EINVAL, "adj_ptr is not monotone");
    }
    for (int64_t e = 0; e < m; ++e)
      if (adj_nbr[e] < 0 || adj_nbr[e] >= n) fail(SMCM_EINVAL, "neighbour id out of range");
    g->ntypes = maxt + 1;
    g->dims.assign(g->ntypes, 0);
    g->h_local.resize(n);
    for (int32_t v = 0; v < n; ++v) g->h_local[v] = g->dims[vtype[v]]++;
    DevAlloc& al = ctx->al;
    g->vtype = al.alloc_n<int32_t>(n);
    g->adj_ptr = al.alloc_n<int64_t>((int64_t)n + 1);
    g->adj_nbr = al.alloc_n<int32_t>(m);
    g->adj_et = al.alloc_n<int32_t>(m);
    g->local = al.alloc_n<int32_t>(n);
    if (n) {
      SMCM_HIP(hipMemcpyAsync(g->vtype, vtype, sizeof(int32_t) * n, hipMemcpyHostToDevice, al.stream));
      SMCM_HIP(hipMemcpyAsync(g->adj_ptr, adj_ptr, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, al.stream));
    }
    if (m) {
      SMCM_HIP(hipMemcpyAsync(g->adj_nbr, adj_nbr, sizeof(int32_t) * m, hipMemcpyHostToDevice, al.stream));
      SMCM_HIP(hipMemcpyAsync(g->adj_et, adj_etype, sizeof(int32_t) * m, hipMemcpyHostToDevice, al.stream));
    }
    g->bins.assign(g->ntypes, nullptr);
    for (int32_t t = 0; t < g->ntypes; ++t) g->bins[t] = al.alloc_n<int32_t>(g->dims[t]);
    smcm::build_type_index(g->vtype, n, g->ntypes, g->local, g->bins.data(), al);
    SMCM_HIP(hipStreamSynchronize(al.stream));
    ctx->live++;
    *out = g.release();
  });
}

int smcm_graph_free(smcm_graph* g) {
  return guarded([&] {
    if (!g) return;
    DevAlloc& al = g->ctx->al;
    (void)hipSetDevice(g->ctx->device);
    for (auto& kv : g->cache) free_csr(al, kv.second);
    for (auto* b : g->bins) al.free(b);
    al.free(g->vtype);
    al.free(g->adj_ptr);
    al.free(g->adj_nbr);
    al.free(g->adj_et);
    al.free(g->local);
    (void)hipStreamSynchronize(al.stream);
    g->ctx->live--;
    delete g;
  });
}

int smcm_graph_dims(const smcm_graph* g, int32_t* dims, int32_t cap, int32_t* ntypes) {
  return guarded([&] {
    if (!g) fail(SMCM_EINVAL, "null graph");
    if (ntypes) *ntypes = g->ntypes;
    for (int32_t t = 0; dims && t < cap && t < g->ntypes; ++t) dims[t] = g->dims[t];
  });
}

int smcm_chain_run(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                   const int32_t* src_rows, int64_t n_src, smcm_result** out, smcm_times* times) {
  return run_chain(ctx, g, vtypes, etypes, L, src_rows, n_src, nullptr, out, times);
}

int smcm_chain_run_plan(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                        const int32_t* src_rows, int64_t n_src, const int32_t* pairs, smcm_result** out,
                        smcm_times* times) {
  if (!pairs && L > 1) return set_err(SMCM_EINVAL, "null plan");
  return run_chain(ctx, g, vtypes, etypes, L, src_rows, n_src, pairs, out, times);
}

int smcm_chain_run_blocked(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes,
                           int32_t L, const int32_t* src_rows, int64_t n_src, uint64_t block_bytes, int64_t* nnz,
                           uint64_t* row_digest, smcm_times* times) {
  return guarded([&] {
    check_ctx(ctx);
    if (block_bytes < 12) fail(SMCM_EINVAL, "block_bytes must hold at least one entry");
    Chain ch{ctx, g};
    ch.prepare(vtypes, etypes, L, src_rows, n_src);
    DevAlloc& al = ctx->al;
    const int32_t nrows = ch.mats[0].nrows;
    DevBuf<uint64_t> dig(al, row_digest ? nrows : 1);
    int64_t total = 0;
    smcm::BlockConsumer consume = [&](const Csr& blk, int32_t b, int32_t e) {
      total += blk.nnz;
      if (row_digest) smcm::row_digest(blk, dig.p + b, al.stream);
    };
    run_core(ch, nullptr, nullptr, &consume, (int64_t)block_bytes, times);
    ctx->last.nnz = total;
    if (L > 1) ctx->last.ref_nnz = total;
    if (nnz) *nnz = total;
    if (row_digest && nrows)
      SMCM_HIP(hipMemcpyAsync(row_digest, dig.p, sizeof(uint64_t) * nrows, hipMemcpyDeviceToHost, al.stream));
    SMCM_HIP(hipStreamSynchronize(al.stream));
  });
}

int smcm_kernel_stats(smcm_ctx* ctx, int32_t idx, char* name, int32_t cap, double* total_ms, int64_t* count,
                      double* bytes) {
  return guarded([&] {
    if (!ctx) fail(SMCM_EINVAL, "null context");
    if (idx < 0 || idx >= (int32_t)ctx->timer.agg.size()) fail(SMCM_EINVAL, "index out of range");
    auto it = ctx->timer.agg.begin();
    std::advance(it, idx);
    if (name && cap > 0) {
      std::strncpy(name, it->first.c_str(), (size_t)cap - 1);
      name[cap - 1] = 0;
    }
    if (total_ms) *total_ms = it->second.ms;
    if (count) *count = it->second.count;
    if (bytes) *bytes = it->second.bytes;
  });
}

int smcm_copy_bandwidth(smcm_ctx* ctx, uint64_t bytes, int32_t iters, double* gbps) {
  return guarded([&] {
    check_ctx(ctx);
    if (bytes < 16 || iters < 1 || !gbps) fail(SMCM_EINVAL, "bad arguments");
    DevAlloc& al = ctx->al;
    bytes = bytes / 16 * 16;
    DevBuf<unsigned char> a(al, (int64_t)bytes), b(al, (int64_t)bytes);
    smcm::fill_i32((int32_t*)a.p, (int64_t)bytes / 4, 1, al.stream);
    smcm::device_copy(a.p, b.p, (int64_t)bytes, al.stream);  // warm-up
    hipEvent_t e0, e1;
    SMCM_HIP(hipEventCreate(&e0));
    SMCM_HIP(hipEventCreate(&e1));
    SMCM_HIP(hipEventRecord(e0, al.stream));
    for (int32_t i = 0; i < iters; ++i) {
      const int tok = ctx->timer.start("copy", al.stream);
      smcm::device_copy(a.p, b.p, (int64_t)bytes, al.stream);
      ctx->timer.stop(tok, al.stream, 2.0 * (double)bytes);
    }
    SMCM_HIP(hipEventRecord(e1, al.stream));
    SMCM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SMCM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->timer.flush();
    *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;  // read + write
  });
}

int smcm_kernel_stats_reset(smcm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) fail(SMCM_EINVAL, "null context");
    ctx->timer.reset();
  });
}

int smcm_chain_plan(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                    const int32_t* src_rows, int64_t n_src, double* c_table, double* m_table, int32_t* s_table,
                    int32_t* pairs) {
  return guarded([&] {
    check_ctx(ctx);
    Chain ch{ctx, g};
    ch.prepare(vtypes, etypes, L, src_rows, n_src);
    ch.estimate();
    ch.dp();
    const size_t LL = (size_t)L * L;
    if (c_table) std::memcpy(c_table, ch.c.data(), sizeof(double) * LL);
    if (m_table) std::memcpy(m_table, ch.mtab.data(), sizeof(double) * LL);
    if (s_table) std::memcpy(s_table, ch.stab.data(), sizeof(int32_t) * LL);
    if (pairs && !ch.pairs.empty()) std::memcpy(pairs, ch.pairs.data(), sizeof(int32_t) * ch.pairs.size());
  });
}

int smcm_chain_products_estimate(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes,
                                 int32_t L, const int32_t* src_rows, int64_t n_src, double* p_table) {
  return guarded([&] {
    check_ctx(ctx);
    if (!p_table) fail(SMCM_EINVAL, "null output");
    Chain ch{ctx, g};
    ch.prepare(vtypes, etypes, L, src_rows, n_src);
    ch.want_products = true;
    ch.estimate();
    const size_t n3 = (size_t)L * L * L;
    if (L > 1) std::memcpy(p_table, ch.ptab.data(), sizeof(double) * n3);
    else p_table[0] = 0.0;
  });
}

int smcm_estimate(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                  const int32_t* src_rows, int64_t n_src, double* est_nnz) {
  return guarded([&] {
    check_ctx(ctx);
    if (!est_nnz) fail(SMCM_EINVAL, "null output");
    Chain ch{ctx, g};
    ch.prepare(vtypes, etypes, L, src_rows, n_src);
    ch.estimate();
    *est_nnz = ch.c[L - 1];
  });
}

int smcm_estimator_rows(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                        const int32_t* src_rows, int64_t n_src, int32_t i, double* out, int64_t cap) {
  return guarded([&] {
    check_ctx(ctx);
    Chain ch{ctx, g};
    ch.prepare(vtypes, etypes, L, src_rows, n_src);
    if (i < 0 || i >= L) fail(SMCM_EINVAL, "position out of range");
    ch.estimate();
    const int64_t need = (int64_t)ch.mats[i].nrows * (L - i);
    if (!out || cap < need) fail(SMCM_EINVAL, "output too small");
    if (need)
      SMCM_HIP(hipMemcpyAsync(out, ch.R[i].p, sizeof(double) * need, hipMemcpyDeviceToHost, ctx->al.stream));
    SMCM_HIP(hipStreamSynchronize(ctx->al.stream));
  });
}

int32_t smcm_result_nrows(const smcm_result* r) { return r ? r->C.nrows : -1; }
int64_t smcm_result_nnz(const smcm_result* r) { return r ? r->C.nnz : -1; }

int smcm_result_copy_rows(const smcm_result* r, int32_t b, int32_t e, int64_t* rowptr, int32_t* cols,
                          uint64_t* vals) {
  return guarded([&] {
    if (!r) fail(SMCM_EINVAL, "null result");
    if (b < 0 || e < b || e > r->C.nrows) fail(SMCM_EINVAL, "row range out of bounds");
    smcm_ctx* ctx = r->ctx;
    check_ctx(ctx);
    DevAlloc& al = ctx->al;
    std::vector<int64_t> ptr(e - b + 1);
    SMCM_HIP(hipMemcpyAsync(ptr.data(), r->C.ptr + b, sizeof(int64_t) * ptr.size(), hipMemcpyDeviceToHost, al.stream));
    SMCM_HIP(hipStreamSynchronize(al.stream));
    const int64_t p0 = ptr[0], cnt = ptr.back() - ptr[0];
    if (rowptr)
      for (size_t q = 0; q < ptr.size(); ++q) rowptr[q] = ptr[q] - p0;
    if (cols && cnt) {
      if (r->col_map) {
        const int64_t chunk = (int64_t)1 << 26;
        DevBuf<int32_t> tmp(al, std::min(cnt, chunk));
        for (int64_t x = 0; x < cnt; x += chunk) {
          const int64_t c = std::min(chunk, cnt - x);
          smcm::map_columns(r->C.col + p0 + x, c, r->col_map, tmp.p, al.stream);
          SMCM_HIP(hipMemcpyAsync(cols + x, tmp.p, sizeof(int32_t) * c, hipMemcpyDeviceToHost, al.stream));
          SMCM_HIP(hipStreamSynchronize(al.stream));
        }
      } else {
        SMCM_HIP(hipMemcpyAsync(cols, r->C.col + p0, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost, al.stream));
      }
    }
    if (vals && cnt)
      SMCM_HIP(hipMemcpyAsync(vals, r->C.val + p0, sizeof(uint64_t) * cnt, hipMemcpyDeviceToHost, al.stream));
    SMCM_HIP(hipStreamSynchronize(al.stream));
  });
}

int smcm_result_copy(const smcm_result* r, int64_t* rowptr, int32_t* cols, uint64_t* vals) {
  if (!r) return set_err(SMCM_EINVAL, "null result");
  return smcm_result_copy_rows(r, 0, r->C.nrows, rowptr, cols, vals);
}

int smcm_result_row_ids(const smcm_result* r, int32_t* rows) {
  return guarded([&] {
    if (!r || !rows) fail(SMCM_EINVAL, "null argument");
    std::copy(r->row_ids.begin(), r->row_ids.end(), rows);
  });
}

int smcm_result_row_digest(const smcm_result* r, uint64_t* out) {
  return guarded([&] {
    if (!r || !out) fail(SMCM_EINVAL, "null argument");
    check_ctx(r->ctx);
    DevAlloc& al = r->ctx->al;
    DevBuf<uint64_t> d(al, r->C.nrows);
    smcm::row_digest(r->C, d.p, al.stream);
    if (r->C.nrows)
      SMCM_HIP(hipMemcpyAsync(out, d.p, sizeof(uint64_t) * r->C.nrows, hipMemcpyDeviceToHost, al.stream));
    SMCM_HIP(hipStreamSynchronize(al.stream));
  });
}

int smcm_result_free(smcm_result* r) {
  return guarded([&] {
    if (!r) return;
    (void)hipSetDevice(r->ctx->device);
    free_csr(r->ctx->al, r->C);
    r->ctx->al.free(r->col_map);
    r->ctx->live--;
    delete r;
  });
}

int smcm_spgemm_csr(smcm_ctx* ctx, int32_t m, int32_t k, int32_t n, const int64_t* a_ptr, const int32_t* a_col,
                    const uint64_t* a_val, const int64_t* b_ptr, const int32_t* b_col, const uint64_t* b_val,
                    smcm_result** out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out || !a_ptr || !b_ptr) fail(SMCM_EINVAL, "null argument");
    *out = nullptr;
    if (m < 0 || k < 0 || n < 0) fail(SMCM_EINVAL, "negative dimension");
    const int64_t annz = a_ptr[m], bnnz = b_ptr[k];
    for (int64_t x = 0; x < annz; ++x)
      if (a_col[x] < 0 || a_col[x] >= k) fail(SMCM_EINVAL, "A column out of range");
    for (int32_t r = 0; r < k; ++r)
      for (int64_t x = b_ptr[r]; x < b_ptr[r + 1]; ++x) {
        if (b_col[x] < 0 || b_col[x] >= n) fail(SMCM_EINVAL, "B column out of range");
        if (x > b_ptr[r] && b_col[x] < b_col[x - 1]) fail(SMCM_EINVAL, "B rows must be sorted");
      }
    DevAlloc& al = ctx->al;
    auto up = [&](const void* h, size_t bytes) {
      void* d = al.alloc(bytes);
      if (bytes) SMCM_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, al.stream));
      return d;
    };
    Csr A, B, C;
    A.nrows = m;
    A.ncols = k;
    A.nnz = annz;
    A.ptr = (int64_t*)up(a_ptr, sizeof(int64_t) * (m + 1));
    A.col = (int32_t*)up(a_col, sizeof(int32_t) * annz);
    A.val = a_val ? (uint64_t*)up(a_val, sizeof(uint64_t) * annz) : nullptr;
    B.nrows = k;
    B.ncols = n;
    B.nnz = bnnz;
    B.ptr = (int64_t*)up(b_ptr, sizeof(int64_t) * (k + 1));
    B.col = (int32_t*)up(b_col, sizeof(int32_t) * bnnz);
    B.val = b_val ? (uint64_t*)up(b_val, sizeof(uint64_t) * bnnz) : nullptr;
    smcm::SpgemmStats ss;
    try {
      smcm::spgemm(A, B, nullptr, 0, C, ctx->opts, &ss, al, true);
    } catch (...) {
      free_csr(al, A);
      free_csr(al, B);
      throw;
    }
    free_csr(al, A);
    free_csr(al, B);
    smcm_stats st{};
    st.products = ss.products;
    st.nnz = C.nnz;
    for (int t = 0; t < 4; ++t) st.rows_tier[t] = ss.rows[t];
    st.n_spgemm = 1;
    ctx->last = st;
    auto* r = new smcm_result;
    r->ctx = ctx;
    r->C = C;
    ctx->live++;
    r->row_ids.resize(m);
    for (int32_t i = 0; i < m; ++i) r->row_ids[i] = i;
    SMCM_HIP(hipStreamSynchronize(al.stream));
    *out = r;
  });
}

}  // extern "C"
