// api_internal.h - host-side state behind the C ABI (contexts, device graphs, results) and
// the per-chain executor, shared by the C-ABI translation units.
#pragma once

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "kernels.h"
#include "providers.h"
#include "smcm.h"

using smcm::Csr;
using smcm::DevAlloc;
using smcm::DevBuf;
using smcm::HipError;

namespace smcm_host {

inline thread_local std::string g_err;

// Device bytes per estimated entry in the product-aware DP's memory bound: an entry of an
// operand or output takes 12 (column and count); a left operand's entry also needs 16 of SpGEMM
// scratch (its B row length and their prefix).
constexpr double kPlanBytesLeft = 28.0, kPlanBytesEntry = 12.0;

inline int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct ApiError {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw ApiError{code, msg}; }

template <class F>
inline int guarded(F&& f) {
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

// device address of smcm_estimate_noalloc's scratch array on `device` (the current device),
// looked up with hipGetSymbolAddress on first use per device (noalloc.hip)
void* noalloc_scratch(int device);

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline void free_csr(DevAlloc& al, Csr& c) {
  al.free(c.ptr);
  al.free(c.col);
  al.free(c.val);
  c = Csr{};
}

}  // namespace smcm_host

struct smcm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  DevAlloc al;
  uint64_t budget = 0;
  smcm::SpgemmOpts opts;
  bool exact_cost_sum = false;
  int plan_cost = 0;  // 0: the reference DP over estimated sizes; 1: estimated products + sizes;
                      // 2-5: the reference's other cost providers (MetaAC, MNC, DMap, LGraph)
  uint64_t lgraph_seed = 0x5EED0000ull;  // LGraph ranks
  struct {
    int64_t small_products = 1 << 25;      // chains with fewer estimated products use the executor
    int threads = 512, sort_cap = 1024;    // executor workgroup shape
    int tile_rows = 32;
    int log2w = 13;                        // executor window width (columns)
    int64_t arena_entries = (int64_t)1 << 31;  // executor output arena (12 bytes per entry)
    int streams = 2;                       // jobs of the tiered path run concurrently on this many
                                           // worker contexts (own stream and allocator each)
    size_t worker_cache = SIZE_MAX;        // freed blocks a worker keeps for its next jobs; the rest
                                           // go back to the device pool, shared by all streams
  } batch;
  struct {
    size_t worker_cache = (size_t)4 << 30; // the same for each worker of a server created from this
                                           // context
  } server;
  std::vector<smcm_ctx*> batch_workers;  // created on first use, destroyed with the context
  // set while the batch workers run jobs: their allocators are then in use by their own threads
  // and an allocation failing on this context must not release their caches
  std::atomic<bool> batch_running{false};
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
  std::vector<std::vector<int32_t>> h_bins;  // host: ascending global ids of every type
  std::vector<int32_t*> bins;
  std::map<std::tuple<int32_t, int32_t, int32_t>, Csr> cache;
  std::recursive_mutex cache_mu;  // typed-matrix cache, shared by the threads of a server
};

struct smcm_result {
  smcm_ctx* ctx = nullptr;
  Csr C;
  int32_t* col_map = nullptr;  // device: local column id -> global id (null = identity)
  std::vector<int32_t> row_ids;
};

namespace smcm_host {

inline void check_ctx(smcm_ctx* c) {
  if (!c) fail(SMCM_EINVAL, "null context");
  SMCM_HIP(hipSetDevice(c->device));
}

// host-side fallback that sorts every row of a (base) matrix whose adjacency order is not
// ascending; used only for the right operand of a product, never for the estimator
inline Csr sorted_copy(DevAlloc& al, const Csr& a) {
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
inline const Csr& typed(smcm_graph* g, int32_t ts, int32_t et, int32_t td) {
  std::lock_guard<std::recursive_mutex> lk(g->cache_mu);
  auto key = std::make_tuple(ts, et, td);
  auto it = g->cache.find(key);
  if (it != g->cache.end()) return it->second;
  DevAlloc& al = g->ctx->al;
  Csr c;
  smcm::build_typed_csr(g->vtype, g->adj_ptr, g->adj_nbr, g->adj_et, g->local, g->bins[ts], nullptr, g->dims[ts],
                        et, td, g->dims[td], c, al);
  // complete before other streams use it
  SMCM_HIP(hipStreamSynchronize(al.stream));
  return g->cache.emplace(key, c).first->second;
}

// sorted version of a cached typed matrix (for right operands); cached under et + 1<<20
inline const Csr& typed_sorted(smcm_graph* g, int32_t ts, int32_t et, int32_t td) {
  std::lock_guard<std::recursive_mutex> lk(g->cache_mu);
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
  int provider = 0;                // 2-5: plan with one of the reference's other cost providers
  std::vector<double> mtab;
  std::vector<int32_t> stab, pairs;
  double es_time = 0, dp_time = 0, spgemm_time = 0, total = 0;
  // called with every product of positions i..j that execute() materialises (not a streamed root)
  std::function<void(int32_t i, int32_t j, const Csr& C)> on_product;

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

  // Estimator passes (and, with want_products, the product-cost table).  By default the c
  // table (and ptab) are copied to the host and the stream is synchronised.  With c_out (device,
  // L*L doubles, and p_out, L*L*L doubles, when want_products) the tables are written there
  // instead and nothing waits: the caller copies them later.
  // Positions 0..k of the meta-path replaced by P = M_0 ... M_k over the chosen sources
  // (e.g. a cached product); the chain is then P * M_{k+1} ... M_{len-1}.
  void prepare_prefix(const int32_t* vtypes, const int32_t* etypes, int32_t len, const Csr& P, int32_t k) {
    if (!g) fail(SMCM_EINVAL, "null graph");
    if (k < 0 || k >= len) fail(SMCM_EINVAL, "bad prefix length");
    L = len - k;
    want_products = ctx->plan_cost == 1;
    vts.assign(1, vtypes[0]);
    vts.insert(vts.end(), vtypes + k + 1, vtypes + len + 1);
    ets.assign(1, -1);
    ets.insert(ets.end(), etypes + k + 1, etypes + len);
    mats.resize(L);
    mats[0] = P;
    for (int32_t i = 1; i < L; ++i) mats[i] = typed(g, vts[i], ets[i], vts[i + 1]);
    subset = false;
  }

  // The reference's other cost providers (efficient/{meta,mnc,dm,lg}.cpp; accuracy = the
  // estimate-only programs accuracy/*.cpp): their c, m and s tables over the global index space.
  void provider_tables(int provider, bool accuracy, smcm::ProviderOut& out) {
    DevAlloc& al = ctx->al;
    smcm::ProviderChain pc;
    pc.L = L;
    pc.mats = mats.data();
    pc.n = g->n;
    pc.seed = ctx->lgraph_seed;
    DevBuf<int32_t> src_g;
    for (int32_t i = 0; i < L; ++i) {
      pc.rowg.push_back(g->bins[vts[i]]);
      pc.colg.push_back(g->bins[vts[i + 1]]);
      pc.same_type.push_back(vts[i] == vts[i + 1]);
    }
    if (subset) {
      // rows of M_0 are the sources: their global ids
      std::vector<int32_t> h(src_local.size());
      const std::vector<int32_t>& ids = g->h_bins[vts[0]];
      for (size_t r = 0; r < h.size(); ++r) h[r] = ids[src_local[r]];
      src_g = DevBuf<int32_t>(al, (int64_t)h.size());
      if (!h.empty())
        SMCM_HIP(hipMemcpyAsync(src_g.p, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice, al.stream));
      pc.rowg[0] = src_g.p;
    }
    std::vector<int32_t> lo(g->ntypes, INT32_MAX), hi(g->ntypes, -1);
    for (int32_t v = 0; v < g->n; ++v) {
      const int32_t t = g->h_vtype[v];
      lo[t] = std::min(lo[t], v);
      hi[t] = std::max(hi[t], v);
    }
    for (int32_t q = 0; q <= L; ++q) {
      const int32_t t = vts[q];
      pc.blo.push_back(hi[t] < 0 ? 0 : lo[t] / 256);
      pc.bhi.push_back(hi[t] < 0 ? -1 : hi[t] / 256);
    }
    if (provider == smcm::kProviderLGraph) {
      for (int32_t i = 0; i < L; ++i) pc.col_type_ids.push_back(g->h_bins[vts[i + 1]]);
    }
    smcm::run_provider(provider, accuracy, pc, out, al);
  }

  void estimate(double* c_out = nullptr, double* p_out = nullptr) {
    DevAlloc& al = ctx->al;
    const double t0 = now_s();
    if (provider >= smcm::kProviderMetaAC) {
      if (c_out || p_out) fail(SMCM_EINVAL, "cost providers 2-5 run on the host-planned path only");
      smcm::ProviderOut po;
      provider_tables(provider, false, po);
      c = po.c;
      mtab = po.m;
      stab = po.s;
      es_time = now_s() - t0;
      return;
    }
    c.assign((size_t)L * L, 0.0);
    R.clear();
    R.resize(L);
    DevBuf<double> cbuf;
    double* cdev = c_out;
    if (!cdev) {
      cbuf = DevBuf<double>(al, (int64_t)L * L);
      cdev = cbuf.p;
    }
    smcm::fill_i64((int64_t*)cdev, (int64_t)L * L, 0, al.stream);  // +0.0 bit pattern
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
      p.c_dev = cdev + (size_t)i * L + i + 1;
      p.exact_sum = ctx->exact_cost_sum;
      p.tm = ctx->timer.enabled ? &ctx->timer : nullptr;
      smcm::estimator_pass(p, al);
    }
    DevBuf<double> pbuf;
    if (want_products && L > 1) {
      const int64_t n3 = (int64_t)L * L * L;
      double* pdev = p_out;
      if (!pdev) {
        pbuf = DevBuf<double>(al, n3);
        pdev = pbuf.p;
      }
      smcm::fill_i64((int64_t*)pdev, n3, 0, al.stream);
      std::vector<const double*> rp(L, nullptr);
      for (int32_t q = 1; q < L; ++q) rp[q] = R[q].p;
      std::vector<double> dq(L + 1);
      for (int32_t q = 0; q <= L; ++q) dq[q] = (double)g->dims[vts[q]];
      smcm::ProductEstimate pe;
      pe.mats = mats.data();
      pe.L = L;
      pe.R = rp.data();
      pe.dims = dq.data();
      pe.p_dev = pdev;
      pe.tm = ctx->timer.enabled ? &ctx->timer : nullptr;
      smcm::product_estimate(pe, al);
      if (!p_out) {
        ptab.assign((size_t)n3, 0.0);
        SMCM_HIP(hipMemcpyAsync(ptab.data(), pdev, sizeof(double) * n3, hipMemcpyDeviceToHost, al.stream));
      }
    }
    if (c_out) {
      es_time = now_s() - t0;
      return;
    }
    SMCM_HIP(hipMemcpyAsync(c.data(), cdev, sizeof(double) * c.size(), hipMemcpyDeviceToHost, al.stream));
    SMCM_HIP(hipStreamSynchronize(al.stream));
    es_time = now_s() - t0;
  }

  // resmcm.cpp:519-538 and :563-573
  void dp() {
    const double t0 = now_s();
    if (provider >= smcm::kProviderMetaAC) {
      // the provider's DP ran with its estimates (provider_tables)
      pairs.clear();
      if (L > 1) plan_rec(0, L - 1);
      dp_time = now_s() - t0;
      return;
    }
    mtab.assign((size_t)L * L, 0.0);
    stab.assign((size_t)L * L, 0);
    if (want_products && !ptab.empty()) {
      // cost(i,j) = min_k cost(i,k) + cost(k+1,j) + P(i,k,j), plus the estimated size of every
      // intermediate (not of the root); same evaluation order and lowest-k tie rule as the
      // reference DP.  Any plan gives the same product (exact integer arithmetic), so this only
      // changes the amount of work.  A product whose estimated operands and output (the root's
      // output aside: it may be streamed) exceed the context's budget is excluded, unless every
      // plan needs one.
      if (!product_dp((double)ctx->budget)) product_dp(HUGE_VAL);
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

  // the product DP without the splits whose estimated device bytes exceed budget (left operand
  // kPlanBytesLeft per entry, right operand and a non-root output kPlanBytesEntry; base
  // relations are resident and count 0); false when every plan is excluded
  bool product_dp(double budget) {
    for (int32_t l = 1; l < L; ++l) {
      for (int32_t i = 0; i < L - l; ++i) {
        const int32_t j = i + l;
        double best = HUGE_VAL;
        for (int32_t k = i; k < j; ++k) {
          const double left = k > i ? c[i * L + k] : 0.0, right = j > k + 1 ? c[(k + 1) * L + j] : 0.0;
          const double out = l < L - 1 ? c[i * L + j] : 0.0;
          if (kPlanBytesLeft * left + kPlanBytesEntry * right + kPlanBytesEntry * out > budget) continue;
          const double cost = mtab[i * L + k] + mtab[(k + 1) * L + j] + ptab[((size_t)i * L + k) * L + j];
          if (cost < best) {
            best = cost;
            stab[i * L + j] = k;
          }
        }
        mtab[i * L + j] = best + (l < L - 1 ? c[i * L + j] : 0.0);
      }
    }
    return mtab[L - 1] < HUGE_VAL;
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
  // set by a server's cancel: checked before every product
  const std::atomic<bool>* cancel = nullptr;

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
      // the estimator's row densities guide row binning (absent when a cost provider planned)
      const double* est = i < (int32_t)R.size() && R[i].p ? R[i].p + (j - i) : nullptr;
      const bool root = ++done_products == nprod;
      // the operands have left the stack: on an error they and the partial product are freed here
      try {
        if (cancel && cancel->load()) fail(SMCM_ECANCELED, "query cancelled");
        if (root && consumer)
          smcm::spgemm_blocked(a.m, b.m, est, L - i, block_bytes, *consumer, ctx->opts, &ss, al, tm);
        else
          smcm::spgemm(a.m, b.m, est, L - i, C, ctx->opts, &ss, al, false, tm);
      } catch (...) {
        free_csr(al, C);  // stream-ordered: blocks are reused only after the queued kernels
        if (a.owned) free_csr(al, a.m);
        if (b.owned) free_csr(al, b.m);
        throw;
      }
      if (on_product && !(root && consumer)) on_product(i, j, C);
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
    SMCM_HIP(hipStreamSynchronize(al.stream));
    spgemm_time = now_s() - t0;
    return st.at(0).m;
  }
};

inline Csr identity_csr(DevAlloc& al, int32_t n) {
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

inline smcm_result* make_result(Chain& ch, Csr C) {
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
  const std::vector<int32_t>& b = ch.g->h_bins[t0];  // global id of local row x of type t0
  if (ch.subset) {
    r->row_ids.resize(ch.src_local.size());
    for (size_t q = 0; q < ch.src_local.size(); ++q) r->row_ids[q] = b[ch.src_local[q]];
  } else {
    r->row_ids = b;
  }
  SMCM_HIP(hipStreamSynchronize(ctx->al.stream));
  return r;
}

// Runs a chain: estimator, DP (or the forced plan) and the products.  Without a consumer the
// product is returned in *C; with one, the root product is streamed in row blocks.
inline void run_core(Chain& ch, const int32_t* forced_pairs, Csr* C, const smcm::BlockConsumer* consumer,
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
        smcm::spgemm(ch.mats[0], I, nullptr, 0, out, ctx->opts, &ss, ctx->al, false, tm);
    } catch (...) {
      free_csr(ctx->al, I);
      throw;
    }
    free_csr(ctx->al, I);
    SMCM_HIP(hipStreamSynchronize(ctx->al.stream));
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

}  // namespace smcm_host
