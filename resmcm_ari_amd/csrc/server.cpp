// server.cpp - serving form: a request loop over device-resident graphs.
//
// Queries (meta-path, source rows, graph) are submitted at any time and run by `slots` worker
// threads, each with its own context (HIP stream, stream-ordered allocator), so the products of
// different queries overlap on the device.  A worker plans its query (estimator, DP), then asks
// for admission: the query's estimator-predicted bytes (12 per estimated entry of every product
// it materialises; a root above block_bytes is streamed in row blocks and counts as one block)
// must fit the budget next to the queries already running, else the worker waits until running
// queries finish; a query that alone exceeds the budget runs when nothing else does.  Waiting
// queries never hold others back: other workers keep taking new arrivals.
//
// Prefix cache: every product of positions 0..j that a query materialises (sources included in
// the key) is copied into a device cache (LRU, capacity cache_bytes, entries up to
// cache_entry_max).  A later query with the same graph, sources and leading matrices starts
// from the longest cached prefix, P * M_{k+1} ... M_{L-1}; integer products are associative, so
// its result is the same bit for bit.
//
// Results are reduced on the device to per-row digests (as smcm_result_row_digest) and the nnz.
#include <condition_variable>
#include <deque>
#include <list>
#include <set>
#include <thread>

#include "api_internal.h"
#include "server_sched.h"
#include "smcm.h"

using namespace smcm_host;

namespace {

double mono_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A query: its meta-path, sources and graph, and what the executor reports (scheduling state,
// status and times live in the Job base)
struct Query : smcm_sched::Job {
  int32_t graph = 0;
  int32_t L = 0;
  std::vector<int32_t> vts, ets, src;  // src: global ids (empty + whole: all rows)
  bool whole = false;
  bool alone = false;          // admitted by itself (a rerun after running out of device memory)
  uint64_t* digest = nullptr;  // caller's buffer (rows of the result), may be null
  smcm_query_out out{};        // nnz, predicted bytes, cache use, blocks, reruns
};

struct CacheEntry {
  std::vector<int32_t> key;
  Csr C;  // device memory owned by the cache (hipMalloc)
  uint64_t bytes = 0;
  int pins = 0;
};

struct PrefixCache {
  std::mutex mu;
  std::list<CacheEntry> lru;  // front = most recently used
  uint64_t bytes = 0, cap = 0, entry_max = 0;
  int64_t hits = 0, inserts = 0;

  static std::vector<int32_t> key_of(const Query& q, int32_t j) {
    std::vector<int32_t> k;
    k.reserve(4 + 2 * j + q.src.size());
    k.push_back(q.graph);
    k.push_back(j);
    k.insert(k.end(), q.vts.begin(), q.vts.begin() + j + 2);
    k.insert(k.end(), q.ets.begin(), q.ets.begin() + j + 1);
    k.push_back(q.whole ? -1 : (int32_t)q.src.size());
    k.insert(k.end(), q.src.begin(), q.src.end());
    return k;
  }

  // longest cached prefix 0..j (j >= 1); pins it
  CacheEntry* lookup(const Query& q, int32_t& jout) {
    if (cap == 0) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    for (int32_t j = q.L - 1; j >= 1; --j) {
      const std::vector<int32_t> k = key_of(q, j);
      for (auto it = lru.begin(); it != lru.end(); ++it) {
        if (it->key == k) {
          lru.splice(lru.begin(), lru, it);
          lru.front().pins++;
          hits++;
          jout = j;
          return &lru.front();
        }
      }
    }
    return nullptr;
  }

  void unpin(CacheEntry* e) {
    if (!e) return;
    std::lock_guard<std::mutex> lk(mu);
    e->pins--;
  }

  // stream-ordered free on the evicting worker's stream: an unpinned entry has no reader left
  // (a reader unpins after its stream is synchronised)
  static void free_entry(CacheEntry& e, hipStream_t st) {
    if (e.C.ptr) (void)hipFreeAsync(e.C.ptr, st);
    if (e.C.col) (void)hipFreeAsync(e.C.col, st);
    if (e.C.val) (void)hipFreeAsync(e.C.val, st);
  }

  // copy C (positions 0..j of q) into the cache; st is the producer's stream
  void insert(const Query& q, int32_t j, const Csr& C, hipStream_t st) {
    if (cap == 0) return;
    const uint64_t bytes = 8ull * (C.nrows + 1) + 12ull * (uint64_t)C.nnz;
    if (bytes > entry_max || bytes > cap) return;
    std::vector<int32_t> k = key_of(q, j);
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& e : lru)
        if (e.key == k) return;
      // make room: evict unpinned entries from the back
      auto it = lru.end();
      while (bytes + this->bytes > cap && it != lru.begin()) {
        --it;
        if (it->pins == 0) {
          this->bytes -= it->bytes;
          free_entry(*it, st);
          it = lru.erase(it);
        }
      }
      if (bytes + this->bytes > cap) return;
      this->bytes += bytes;  // reserved
    }
    CacheEntry e;
    e.key = std::move(k);
    e.bytes = bytes;
    e.C = C;
    e.C.ptr = nullptr;
    e.C.col = nullptr;
    e.C.val = nullptr;
    // products always carry values (base relations never reach the cache)
    bool ok = C.val != nullptr || C.nnz == 0;
    ok = ok && hipMallocAsync((void**)&e.C.ptr, sizeof(int64_t) * (C.nrows + 1), st) == hipSuccess &&
         hipMallocAsync((void**)&e.C.col, sizeof(int32_t) * std::max<int64_t>(C.nnz, 1), st) == hipSuccess &&
         hipMallocAsync((void**)&e.C.val, sizeof(uint64_t) * std::max<int64_t>(C.nnz, 1), st) == hipSuccess;
    ok = ok && hipMemcpyAsync(e.C.ptr, C.ptr, sizeof(int64_t) * (C.nrows + 1), hipMemcpyDeviceToDevice, st) == hipSuccess;
    if (ok && C.nnz) {
      ok = hipMemcpyAsync(e.C.col, C.col, sizeof(int32_t) * C.nnz, hipMemcpyDeviceToDevice, st) == hipSuccess &&
           hipMemcpyAsync(e.C.val, C.val, sizeof(uint64_t) * C.nnz, hipMemcpyDeviceToDevice, st) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(st) == hipSuccess;
    std::lock_guard<std::mutex> lk(mu);
    if (!ok) {
      (void)hipGetLastError();
      free_entry(e, st);
      this->bytes -= bytes;
      return;
    }
    inserts++;
    lru.push_front(std::move(e));
  }

  void clear() {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& e : lru) {
      (void)hipFree(e.C.ptr);
      (void)hipFree(e.C.col);
      (void)hipFree(e.C.val);
    }
    lru.clear();
    bytes = 0;
  }
};

}  // namespace

struct smcm_server {
  smcm_ctx* main = nullptr;
  std::vector<smcm_graph*> graphs;
  smcm_server_opts opts{};
  std::vector<smcm_ctx*> wctx;
  std::unique_ptr<smcm_sched::Scheduler> sched;  // queues, lanes, admission, shutdown
  PrefixCache cache;
};

namespace {

// estimator-predicted bytes of the products the plan materialises (12 per estimated entry); a
// root above block_bytes is streamed and counts one block.  root receives the root's estimate.
double predicted_bytes(const Chain& ch, uint64_t block_bytes, double& root) {
  const int32_t L = ch.L;
  root = 12.0 * std::max(0.0, ch.c[L - 1]);
  double pred = 0;
  std::vector<std::pair<int32_t, int32_t>> stk;
  for (size_t x = 0; 2 * x < ch.pairs.size(); ++x) {
    const int32_t k = ch.pairs[2 * x], l = ch.pairs[2 * x + 1];
    std::pair<int32_t, int32_t> A, B;
    if (k >= 0 && l >= 0) {
      A = {k, k};
      B = {l, l};
    } else if (k < 0 && l >= 0) {
      A = stk.back();
      stk.pop_back();
      B = {l, l};
    } else if (k >= 0 && l < 0) {
      B = stk.back();
      stk.pop_back();
      A = {k, k};
    } else {
      B = stk.back();
      stk.pop_back();
      A = stk.back();
      stk.pop_back();
    }
    if (!(A.first == 0 && B.second == L - 1)) pred += 12.0 * std::max(0.0, ch.c[A.first * L + B.second]);
    stk.push_back({A.first, B.second});
  }
  return pred + (root > (double)block_bytes ? (double)block_bytes : root);
}

// runs the planned chain; per-row digests to dig_out (host, nrows), returns the nnz
int64_t execute_digest(Chain& ch, bool blocked, uint64_t block_bytes, DevAlloc& al, uint64_t* dig_out,
                       int32_t nrows) {
  hipStream_t st = al.stream;
  smcm_stats sst{};
  if (blocked) {
    DevBuf<uint64_t> dig(al, std::max(nrows, 1));
    int64_t total = 0;
    smcm::BlockConsumer consume = [&](const Csr& blk, int32_t b, int32_t e) {
      if (ch.cancel && ch.cancel->load()) fail(SMCM_ECANCELED, "query cancelled");
      total += blk.nnz;
      if (dig_out) smcm::row_digest(blk, dig.p + b, st);
    };
    (void)ch.execute(ch.pairs, &sst, &consume, (int64_t)block_bytes);
    if (dig_out && nrows)
      SMCM_HIP(hipMemcpyAsync(dig_out, dig.p, sizeof(uint64_t) * nrows, hipMemcpyDeviceToHost, st));
    SMCM_HIP(hipStreamSynchronize(st));
    return total;
  }
  Csr C = ch.execute(ch.pairs, &sst);
  try {
    if (dig_out && C.nrows) {
      DevBuf<uint64_t> d(al, C.nrows);
      smcm::row_digest(C, d.p, st);
      SMCM_HIP(hipMemcpyAsync(dig_out, d.p, sizeof(uint64_t) * C.nrows, hipMemcpyDeviceToHost, st));
    }
    SMCM_HIP(hipStreamSynchronize(st));
  } catch (...) {
    free_csr(al, C);
    throw;
  }
  const int64_t nnz = C.nnz;
  free_csr(al, C);
  return nnz;
}

// Plans and runs one query.  Returns false (query not finished) when a small-lane worker finds
// it large: the query is then queued for the large-lane workers.
bool run_query(smcm_server* s, smcm_ctx* w, Query& q, bool small_lane) {
  auto check_cancel = [&q] {
    if (q.cancelled.load()) fail(SMCM_ECANCELED, "query cancelled");
  };
  check_cancel();
  smcm_graph* g = s->graphs[q.graph];
  DevAlloc& al = w->al;
  hipStream_t st = al.stream;
  const int32_t nrows = q.whole ? g->dims[q.vts[0]] : (int32_t)q.src.size();
  int32_t kk = 0;
  CacheEntry* hit = s->cache.lookup(q, kk);
  struct Unpin {
    PrefixCache& c;
    CacheEntry*& e;
    ~Unpin() { c.unpin(e); }
  } unpin{s->cache, hit};
  q.out.cache_k = hit ? kk : 0;
  auto finish_digest = [&](const Csr& C) {
    q.out.nnz = C.nnz;
    if (q.digest && C.nrows) {
      DevBuf<uint64_t> d(al, C.nrows);
      smcm::row_digest(C, d.p, st);
      SMCM_HIP(hipMemcpyAsync(q.digest, d.p, sizeof(uint64_t) * C.nrows, hipMemcpyDeviceToHost, st));
    }
    SMCM_HIP(hipStreamSynchronize(st));
  };
  // admission: held from admit() until the query's device work is done (RAII, also on errors);
  // a query cancelled while it waits is not admitted
  struct Admission {
    smcm_sched::Scheduler& sc;
    Query* q = nullptr;
    void admit(uint64_t b, Query& qq) {
      if (!(qq.alone ? sc.admit_alone(qq) : sc.admit(qq, b))) fail(SMCM_ECANCELED, "query cancelled");
      q = &qq;
      qq.out.predicted_bytes = b;
    }
    ~Admission() {
      if (q) sc.release(*q);  // the query's device work ends before its budget is freed
    }
  } adm{*s->sched};
  if (hit && kk == q.L - 1) {  // the whole chain is cached
    adm.admit(0, q);
    finish_digest(hit->C);
    return true;
  }
  Chain ch{w, g};
  if (hit) {
    ch.prepare_prefix(q.vts.data(), q.ets.data(), q.L, hit->C, kk);
  } else {
    ch.prepare(q.vts.data(), q.ets.data(), q.L, q.whole ? nullptr : q.src.data(), q.whole ? 0 : q.src.size());
  }
  ch.want_products = s->opts.plan_cost == 1 && ch.L > 1;
  ch.cancel = &q.cancelled;
  if (ch.L == 1) {
    // a single relation: merge duplicates through the per-chain path
    adm.admit(12ull * (uint64_t)ch.mats[0].nnz, q);
    Csr C;
    run_core(ch, nullptr, &C, nullptr, 0, nullptr);
    try {
      finish_digest(C);
    } catch (...) {
      free_csr(al, C);
      throw;
    }
    free_csr(al, C);
    return true;
  }
  ch.estimate();
  ch.dp();
  double root = 0;
  const double pred = predicted_bytes(ch, s->opts.block_bytes, root);
  if (small_lane && pred > (double)s->opts.large_bytes) {
    s->sched->hand_over(q);
    return false;
  }
  const uint64_t quarter = std::max<uint64_t>(s->opts.budget_bytes / 4, 1);
  if (!hit && pred > (double)(s->opts.budget_bytes / 2) && nrows > 1) {
    // Too large to run whole next to others: independent blocks of source rows, one after the
    // other (row u of the result depends only on source row u), each planned on its own.
    const int64_t K = std::min<int64_t>(nrows, (int64_t)std::ceil(pred / (double)quarter));
    std::vector<int32_t> rows = q.src;
    if (q.whole) rows.insert(rows.end(), g->h_bins[q.vts[0]].begin(), g->h_bins[q.vts[0]].end());
    adm.admit(quarter, q);
    int64_t total = 0;
    for (int64_t b = 0; b < K; ++b) {
      const int64_t r0 = b * nrows / K, r1 = (b + 1) * nrows / K;
      if (r1 == r0) continue;
      check_cancel();
      Chain cb{w, g};
      cb.prepare(q.vts.data(), q.ets.data(), q.L, rows.data() + r0, r1 - r0);
      cb.want_products = s->opts.plan_cost == 1;
      cb.cancel = &q.cancelled;
      cb.estimate();
      cb.dp();
      double rb = 0;
      (void)predicted_bytes(cb, s->opts.block_bytes, rb);
      total += execute_digest(cb, rb > (double)s->opts.block_bytes, s->opts.block_bytes, al,
                              q.digest ? q.digest + r0 : nullptr, (int32_t)(r1 - r0));
    }
    q.out.nnz = total;
    q.out.blocks = (int32_t)K;
    return true;
  }
  adm.admit((uint64_t)std::min(pred, 1.8e19), q);
  const int32_t off = hit ? kk : 0;
  if (s->cache.cap) {
    ch.on_product = [&](int32_t i, int32_t j, const Csr& C) {
      if (i == 0) s->cache.insert(q, j + off, C, st);
    };
  }
  q.out.nnz = execute_digest(ch, root > (double)s->opts.block_bytes, s->opts.block_bytes, al, q.digest, nrows);
  return true;
}

// the scheduler's executor: runs one query on worker wi's context and maps failures to a status
bool run_job(smcm_server* s, smcm_sched::Job& job, int wi, bool small_lane) {
  Query& q = static_cast<Query&>(job);
  smcm_ctx* w = s->wctx[wi];
  int status = SMCM_OK;
  std::string msg;
  bool finished = true;
  try {
    SMCM_HIP(hipSetDevice(w->device));
    try {
      finished = run_query(s, w, q, small_lane);
    } catch (const HipError& e) {
      // Admission counts predicted bytes, so queries admitted together can still outgrow the
      // device.  A query that ran out of memory next to others runs once more by itself, after
      // its worker's cached blocks went back to the device pool (its buffers were freed while
      // the exception unwound).
      if (e.code != (int)hipErrorOutOfMemory || q.alone || q.cancelled.load()) throw;
      (void)hipGetLastError();
      q.alone = true;
      q.out.reruns++;
      w->al.release_cache();
      SMCM_HIP(hipStreamSynchronize(w->stream));
      smcm::DevAlloc::trim_pool();
      finished = run_query(s, w, q, small_lane);
    }
  } catch (const ApiError& e) {
    status = e.code;
    msg = e.msg;
  } catch (const HipError& e) {
    status = e.code == (int)hipErrorOutOfMemory ? SMCM_ENOMEM : SMCM_EDEVICE;
    msg = e.what();
    (void)hipGetLastError();
  } catch (const std::exception& e) {
    status = SMCM_EDEVICE;
    msg = e.what();
  }
  if (!finished) return false;
  q.status = status;
  q.error = msg;
  if (status != SMCM_OK && std::getenv("SMCM_TRACE")) fprintf(stderr, "[smcm] query failed: %s\n", msg.c_str());
  return true;
}

// checks of a query that smcm_server_submit and smcm_server_replay make before anything is queued
void check_query(smcm_server* s, int32_t graph, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                 const int32_t* src_rows, int64_t n_src) {
  if (!vtypes || !etypes) fail(SMCM_EINVAL, "null meta-path");
  if (graph < 0 || graph >= (int32_t)s->graphs.size()) fail(SMCM_EINVAL, "graph index out of range");
  if (L < 1 || L > 16) fail(SMCM_EINVAL, "the meta-path is incorrect: L must be in [1, 16]");
  smcm_graph* g = s->graphs[graph];
  for (int32_t q = 0; q <= L; ++q)
    if (vtypes[q] < 0 || vtypes[q] >= g->ntypes) fail(SMCM_EINVAL, "the meta-path is incorrect: vertex type");
  for (int32_t q = 0; q < L; ++q)
    if (etypes[q] < 0) fail(SMCM_EINVAL, "the meta-path is incorrect: negative edge type");
  if (src_rows) {
    if (n_src < 0 || n_src > INT32_MAX) fail(SMCM_EINVAL, "bad source row count");
    for (int64_t r = 0; r < n_src; ++r) {
      const int32_t v = src_rows[r];
      if (v < 0 || v >= g->n || g->h_vtype[v] != vtypes[0])
        fail(SMCM_EINVAL, "source row " + std::to_string(v) + " is not a vertex of the start type");
    }
  }
}

}  // namespace

extern "C" {

int smcm_server_create(smcm_ctx* ctx, smcm_graph* const* graphs, int32_t n_graphs, const smcm_server_opts* opts,
                       smcm_server** out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out || !graphs || n_graphs < 1) fail(SMCM_EINVAL, "bad arguments");
    *out = nullptr;
    auto s = std::make_unique<smcm_server>();
    s->main = ctx;
    for (int32_t i = 0; i < n_graphs; ++i) {
      if (!graphs[i] || graphs[i]->ctx != ctx) fail(SMCM_EINVAL, "graphs must belong to the server's context");
      s->graphs.push_back(graphs[i]);
    }
    smcm_server_opts o{};
    o.slots = 6;
    o.plan_cost = 1;
    o.budget_bytes = 200ull << 30;
    o.cache_bytes = 16ull << 30;
    o.cache_entry_max = 2ull << 30;
    o.block_bytes = 32ull << 30;
    o.large_slots = 2;
    o.large_bytes = 4ull << 30;
    o.large_cu_percent = 50;
    if (opts) o = *opts;
    if (o.large_slots < 0 || o.large_slots > o.slots) fail(SMCM_EINVAL, "large_slots must be in [0, slots]");
    if (o.large_cu_percent < 0 || o.large_cu_percent > 100) fail(SMCM_EINVAL, "large_cu_percent must be in [0, 100]");
    if (o.slots < 1 || o.slots > 16) fail(SMCM_EINVAL, "slots must be in [1, 16]");
    if (o.plan_cost != 0 && o.plan_cost != 1) fail(SMCM_EINVAL, "plan_cost must be 0 or 1");
    if (o.block_bytes < (1ull << 20)) fail(SMCM_EINVAL, "block_bytes too small");
    s->opts = o;
    // the workers allocate on their own streams: blocks the main context and its batch workers
    // keep cached for reuse (they stay idle while the server runs) go back to the device pool, or
    // they would be out of the workers' reach
    if (ctx->al.on_oom) ctx->al.on_oom();  // the batch workers' caches
    ctx->al.release_cache();
    SMCM_HIP(hipStreamSynchronize(ctx->stream));
    smcm::DevAlloc::trim_pool();
    s->cache.cap = o.cache_bytes;
    s->cache.entry_max = o.cache_entry_max;
    // every typed matrix of the graphs, built now so workers only read the cache
    for (smcm_graph* g : s->graphs) {
      std::vector<std::tuple<int32_t, int32_t, int32_t>> rel;
      std::vector<int64_t> ptr(g->n + 1);
      std::vector<int32_t> nbr(g->m), et(g->m);
      SMCM_HIP(hipMemcpy(ptr.data(), g->adj_ptr, sizeof(int64_t) * ptr.size(), hipMemcpyDeviceToHost));
      if (g->m) {
        SMCM_HIP(hipMemcpy(nbr.data(), g->adj_nbr, sizeof(int32_t) * g->m, hipMemcpyDeviceToHost));
        SMCM_HIP(hipMemcpy(et.data(), g->adj_et, sizeof(int32_t) * g->m, hipMemcpyDeviceToHost));
      }
      std::set<std::tuple<int32_t, int32_t, int32_t>> seen;
      for (int32_t v = 0; v < g->n; ++v)
        for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e) seen.insert({g->h_vtype[v], et[e], g->h_vtype[nbr[e]]});
      for (auto& t : seen) typed_sorted(g, std::get<0>(t), std::get<1>(t), std::get<2>(t));
    }
    for (int32_t i = 0; i < o.slots; ++i) {
      smcm_ctx* w = nullptr;
      const int rc = smcm_ctx_create(ctx->device, ctx->budget, &w);
      if (rc != SMCM_OK) fail(rc, "worker context: " + std::string(smcm_last_error()));
      w->opts = ctx->opts;
      w->exact_cost_sum = ctx->exact_cost_sum;
      w->plan_cost = o.plan_cost;
      w->al.cache_cap = ctx->server.worker_cache;  // the rest of the freed memory goes to the pool
      // small blocks too (at most 1 GiB of them): a short query's allocations then wait for no
      // device-pool call, which another worker's mapping of a large block can hold for seconds
      w->al.small_cache_cap = std::min<size_t>(ctx->server.worker_cache, (size_t)1 << 30);
      w->al.headroom = ctx->al.headroom;
      s->wctx.push_back(w);
      const bool large_lane = o.large_slots > 0 && o.large_slots < o.slots && i >= o.slots - o.large_slots;
      if (large_lane && o.large_cu_percent > 0 && o.large_cu_percent < 100) {
        // the persistent (queue-driven) kernels of large queries are sized for this share of
        // the CUs, so their workgroups leave room for the kernels of short queries
        w->opts.cus = std::max<int32_t>(1, ctx->opts.cus * o.large_cu_percent / 100);
        // one window-tier workgroup per CU of the share: the largest window and workgroup
        w->opts.w_threads = 1024;
        w->opts.window_cols = 16384;
      }
    }
    // lanes only when both kinds of workers exist (the scheduler's small_lanes)
    s->sched = std::make_unique<smcm_sched::Scheduler>(o.slots, o.large_slots, o.budget_bytes);
    smcm_server* raw = s.get();
    s->sched->start([raw](smcm_sched::Job& j, int wi, bool small) { return run_job(raw, j, wi, small); });
    ctx->live++;
    *out = s.release();
  });
}

int smcm_server_destroy(smcm_server* s) {
  return guarded([&] {
    if (!s) return;
    s->sched->stop();  // every submitted query finishes first
    s->cache.clear();
    for (smcm_ctx* w : s->wctx) smcm_ctx_destroy(w);
    s->main->live--;
    delete s;
  });
}

int smcm_server_submit(smcm_server* s, int32_t graph, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                       const int32_t* src_rows, int64_t n_src, uint64_t* row_digest, int64_t* ticket) {
  return guarded([&] {
    if (!s || !ticket) fail(SMCM_EINVAL, "null argument");
    check_query(s, graph, vtypes, etypes, L, src_rows, n_src);
    auto q = std::make_unique<Query>();
    q->graph = graph;
    q->L = L;
    q->vts.assign(vtypes, vtypes + L + 1);
    q->ets.assign(etypes, etypes + L);
    q->whole = src_rows == nullptr;
    if (!q->whole) q->src.assign(src_rows, src_rows + n_src);
    q->digest = row_digest;
    *ticket = s->sched->submit(std::move(q));
  });
}

int smcm_server_poll(smcm_server* s, int64_t ticket, int32_t wait, smcm_query_out* out) {
  if (!s || !out) return set_err(SMCM_EINVAL, "null argument");
  std::unique_ptr<smcm_sched::Job> j;
  const int rc = s->sched->poll(ticket, wait != 0, &j);
  if (rc < 0) return set_err(SMCM_EINVAL, "unknown ticket");
  if (rc > 0) return SMCM_EAGAIN;
  const Query& q = static_cast<const Query&>(*j);
  *out = q.out;
  out->status = q.status;
  out->t_submit = q.t_submit;
  out->t_start = q.t_start;
  out->t_done = q.t_done;
  std::memset(out->error, 0, sizeof(out->error));
  if (q.status != SMCM_OK) std::strncpy(out->error, q.error.c_str(), sizeof(out->error) - 1);
  return SMCM_OK;
}

int smcm_server_cancel(smcm_server* s, int64_t ticket) {
  if (!s) return set_err(SMCM_EINVAL, "null server");
  const int rc = s->sched->cancel(ticket, SMCM_ECANCELED, "query cancelled before it started");
  if (rc < 0) return set_err(SMCM_EINVAL, "unknown ticket");
  return SMCM_OK;
}

int smcm_server_stats(smcm_server* s, int64_t* cache_hits, int64_t* cache_inserts, uint64_t* cache_bytes) {
  if (!s) return set_err(SMCM_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->cache.mu);
  if (cache_hits) *cache_hits = s->cache.hits;
  if (cache_inserts) *cache_inserts = s->cache.inserts;
  if (cache_bytes) *cache_bytes = s->cache.bytes;
  return SMCM_OK;
}

int smcm_server_replay(smcm_server* s, int32_t n, const int32_t* graph, const int32_t* path_len,
                       const int32_t* vtypes, const int32_t* etypes, const int64_t* src_off, const int32_t* src_rows,
                       const double* arrival, uint64_t* row_digest, smcm_query_out* out) {
  return guarded([&] {
    if (!s || n < 0 || (n > 0 && (!graph || !path_len || !vtypes || !etypes || !arrival || !out)))
      fail(SMCM_EINVAL, "bad arguments");
    if (src_rows && !src_off) fail(SMCM_EINVAL, "src_rows without src_off");
    // every query is checked before the first is submitted: a submitted query writes into
    // row_digest until it is polled, so the call must not return while one still runs
    {
      int64_t vo = 0, eo = 0;
      for (int32_t i = 0; i < n; ++i) {
        const int32_t L = path_len[i];
        if (L < 1 || L > 16) fail(SMCM_EINVAL, "query " + std::to_string(i) + ": the meta-path is incorrect");
        if (src_rows && src_off[i + 1] < src_off[i]) fail(SMCM_EINVAL, "src_off must not decrease");
        try {
          check_query(s, graph[i], vtypes + vo, etypes + eo, L, src_rows ? src_rows + src_off[i] : nullptr,
                      src_rows ? src_off[i + 1] - src_off[i] : 0);
        } catch (const ApiError& e) {
          fail(e.code, "query " + std::to_string(i) + ": " + e.msg);
        }
        vo += L + 1;
        eo += L;
      }
    }
    std::vector<int64_t> tickets;
    tickets.reserve(n);
    int64_t vo = 0, eo = 0, dig = 0;
    const double base = mono_s();
    const double shift = base - s->sched->t0();  // scheduler times -> times relative to this call
    int rc = SMCM_OK;
    std::string err;
    for (int32_t i = 0; i < n && rc == SMCM_OK; ++i) {
      const double due = base + arrival[i];
      for (double now = mono_s(); now < due; now = mono_s())
        std::this_thread::sleep_for(std::chrono::duration<double>(std::min(due - now, 0.002)));
      const int32_t L = path_len[i];
      const int32_t t0 = vtypes[vo];
      const int32_t* src = src_rows ? src_rows + src_off[i] : nullptr;
      const int64_t ns = src_rows ? src_off[i + 1] - src_off[i] : 0;
      uint64_t* d = row_digest ? row_digest + dig : nullptr;
      int64_t t = 0;
      rc = smcm_server_submit(s, graph[i], vtypes + vo, etypes + eo, L, src, ns, d, &t);
      if (rc != SMCM_OK) {
        err = smcm_last_error();
        break;
      }
      tickets.push_back(t);
      vo += L + 1;
      eo += L;
      dig += src_rows ? ns : s->graphs[graph[i]]->dims[t0];
    }
    // wait for every submitted query, also after a failed submission (they write row_digest)
    for (size_t i = 0; i < tickets.size(); ++i) {
      const int prc = smcm_server_poll(s, tickets[i], 1, &out[i]);
      if (prc != SMCM_OK) {
        if (rc == SMCM_OK) {
          rc = prc;
          err = smcm_last_error();
        }
        continue;
      }
      out[i].t_submit -= shift;
      out[i].t_start -= shift;
      out[i].t_done -= shift;
    }
    if (rc != SMCM_OK) fail(rc, err);
  });
}

}  // extern "C"
