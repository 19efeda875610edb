// batch_api.cpp - smcm_batch_run: many chains in one call.
//
// 0. Chains with the same meta-path become one job over the union of their source rows: row u
//    of a chain's result depends only on row u of its first relation, so each chain's result is
//    the job result's rows at its sources (digests and counts are picked, handles gathered).
// 1. Planning, per job on the device: the row estimator and the product-cost table
//    (estimator.hip), all chains enqueued back to back and the tables copied once; then the
//    product-aware DP on the host.  Every plan yields the same exact product, so the choice only
//    changes the work.
// 2. Chains whose estimated products are small go to the persistent executor (batch.hip): all of
//    their products are nodes of one launch, the subset of the first relation is a row map (no
//    gather), intermediate rows stay in the executor's arena.  Their results are digested (and,
//    if asked for, compacted into result handles) straight from the arena.
// 3. The other chains, and any small chain whose products outgrew the arena, run through the
//    per-chain tiered path with the same plan, several jobs at a time on worker contexts (own
//    stream and allocator each), largest first, each admitted when its estimator-predicted bytes
//    fit the context's budget next to the running ones.  A job that still runs out of device
//    memory is run again alone on the calling context.
#include "smcm.h"

#include "api_internal.h"
#include "server_sched.h"

using namespace smcm_host;

namespace {

struct BatchChain {
  int32_t L = 0;
  const int32_t* vts = nullptr;
  const int32_t* ets = nullptr;
  const int32_t* src = nullptr;  // global ids (null = every vertex of the start type)
  int64_t nsrc = 0;
  int32_t nrows = 0;             // rows of the result
  int64_t dig_off = 0;           // first row in the caller's digest array
  std::vector<int32_t> pairs;
  double est_products = 0;
  bool small = false;
  int32_t root = -1;             // executor node of the result
  DevBuf<int32_t> rowmap;        // local ids of the source rows (executor row map)
  std::vector<int32_t> src_local;
  std::vector<int32_t> src_global;  // the job's rows (owned here; src points into it)
  // outputs per row of the job: digest, length; the result itself when handles are wanted
  std::vector<uint64_t> h_dig;
  std::vector<int32_t> h_len;
  Csr C;
  bool have_C = false;
  smcm_ctx* c_owner = nullptr;   // the context whose allocator holds C
  std::vector<double> ctab;      // estimated sizes c_ij (L*L)
};

// one chain of the call: its job and its rows' positions in the job's rows
struct Member {
  int32_t job = -1;
  bool whole = false;
  std::vector<int32_t> local;  // local ids of its source rows
  int64_t dig_off = 0;
};

template <class T>
T* upload(DevAlloc& al, const std::vector<T>& h) {
  T* d = al.alloc_n<T>((int64_t)h.size());
  if (!h.empty()) SMCM_HIP(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, al.stream));
  return d;
}

// global ids of the rows of a result (as make_result)
std::vector<int32_t> result_rows(smcm_graph* g, int32_t t0, const std::vector<int32_t>& src_local, bool whole) {
  const std::vector<int32_t>& b = g->h_bins[t0];
  if (whole) return b;
  std::vector<int32_t> rows;
  rows.resize(src_local.size());
  for (size_t q = 0; q < src_local.size(); ++q) rows[q] = b[src_local[q]];
  return rows;
}

smcm_result* wrap_result(smcm_ctx* ctx, smcm_graph* g, const int32_t* vts, int32_t L,
                         const std::vector<int32_t>& src_local, bool whole, Csr C) {
  auto* r = new smcm_result;
  r->ctx = ctx;
  r->C = C;
  ctx->live++;
  const int32_t tcol = vts[L];
  const int32_t dcol = g->dims[tcol];
  r->col_map = ctx->al.alloc_n<int32_t>(dcol);
  if (dcol)
    SMCM_HIP(hipMemcpyAsync(r->col_map, g->bins[tcol], sizeof(int32_t) * dcol, hipMemcpyDeviceToDevice,
                            ctx->al.stream));
  r->row_ids = result_rows(g, vts[0], src_local, whole);
  return r;
}

struct Executor {
  smcm_ctx* ctx;
  smcm_graph* g;
  std::vector<smcm::BatchNode> nodes;
  std::vector<std::vector<int32_t>> consumers;
  std::vector<int64_t> row_off;  // node -> first row in the row arrays

  // post-order plan -> nodes; returns the root node
  int32_t add_chain(BatchChain& bc) {
    struct Op {
      int32_t node;  // -1: leaf
      int32_t i, j;
    };
    std::vector<Op> st;
    auto leaf = [&](int32_t x) { return Op{-1, x, x}; };
    auto make = [&](Op a, Op b) {
      smcm::BatchNode n;
      const int32_t id = (int32_t)nodes.size();
      if (a.node >= 0) {
        n.a_node = a.node;
        n.nrows = nodes[a.node].nrows;
        consumers[a.node].push_back(id);
      } else {
        const Csr& m = typed(g, bc.vts[a.i], bc.ets[a.i], bc.vts[a.i + 1]);
        n.a_ptr = m.ptr;
        n.a_col = m.col;
        n.a_val = m.val;
        n.nrows = m.nrows;
        if (a.i == 0 && bc.src) {
          n.a_rows = bc.rowmap.p;
          n.nrows = (int32_t)bc.nsrc;
        }
      }
      if (b.node >= 0) {
        n.b_node = b.node;
        consumers[b.node].push_back(id);
      } else {
        const Csr& m = typed(g, bc.vts[b.i], bc.ets[b.i], bc.vts[b.i + 1]);
        n.b_ptr = m.ptr;
        n.b_col = m.col;
        n.b_val = m.val;
      }
      n.ncols = g->dims[bc.vts[b.j + 1]];
      n.ninputs = (a.node >= 0) + (b.node >= 0);
      nodes.push_back(n);
      consumers.emplace_back();
      return Op{id, a.i, b.j};
    };
    for (size_t q = 0; 2 * q < bc.pairs.size(); ++q) {
      const int32_t k = bc.pairs[2 * q], l = bc.pairs[2 * q + 1];
      if (k >= 0 && l >= 0) {
        st.push_back(make(leaf(k), leaf(l)));
      } else if (k < 0 && l >= 0) {
        Op top = st.back();
        st.pop_back();
        st.push_back(make(top, leaf(l)));
      } else if (k >= 0 && l < 0) {
        Op top = st.back();
        st.pop_back();
        st.push_back(make(leaf(k), top));
      } else {
        Op r = st.back();
        st.pop_back();
        Op lft = st.back();
        st.pop_back();
        st.push_back(make(lft, r));
      }
    }
    return st.at(0).node;
  }
};

// per-chain tiered path with the job's plan: row digests and lengths, and the result itself
// when keep is set
void run_large(smcm_ctx* ctx, smcm_graph* g, BatchChain& bc, bool keep) {
  DevAlloc& al = ctx->al;
  Chain ch{ctx, g};
  ch.prepare(bc.vts, bc.ets, bc.L, bc.src, bc.src ? bc.nsrc : 0);
  ch.want_products = false;
  Csr C;
  if (bc.L == 1) {
    run_core(ch, nullptr, &C, nullptr, 0, nullptr);
  } else {
    ch.estimate();  // the estimator's densities drive row binning
    smcm_stats st{};
    C = ch.execute(bc.pairs, &st);
  }
  try {
    bc.h_dig.assign(C.nrows, 0);
    bc.h_len.assign(C.nrows, 0);
    if (C.nrows) {
      DevBuf<uint64_t> d(al, C.nrows);
      smcm::row_digest(C, d.p, al.stream);
      std::vector<int64_t> ptr(C.nrows + 1);
      SMCM_HIP(hipMemcpyAsync(bc.h_dig.data(), d.p, sizeof(uint64_t) * C.nrows, hipMemcpyDeviceToHost, al.stream));
      SMCM_HIP(hipMemcpyAsync(ptr.data(), C.ptr, sizeof(int64_t) * ptr.size(), hipMemcpyDeviceToHost, al.stream));
      SMCM_HIP(hipStreamSynchronize(al.stream));
      for (int32_t u = 0; u < C.nrows; ++u) bc.h_len[u] = (int32_t)(ptr[u + 1] - ptr[u]);
    }
  } catch (...) {
    free_csr(al, C);
    throw;
  }
  if (keep) {
    bc.C = C;
    bc.have_C = true;
    bc.c_owner = ctx;
  } else {
    free_csr(al, C);
  }
}

// estimator-predicted device bytes of a job on the tiered path: 12 per estimated entry of every
// product its plan materialises, the root included (it is kept or digested whole)
double job_bytes(const BatchChain& b, smcm_graph* g) {
  const int32_t L = b.L;
  if (L == 1) return 12.0 * (double)typed(g, b.vts[0], b.ets[0], b.vts[1]).nnz;
  double pred = 0;
  std::vector<std::pair<int32_t, int32_t>> stk;
  for (size_t x = 0; 2 * x < b.pairs.size(); ++x) {
    const int32_t k = b.pairs[2 * x], l = b.pairs[2 * x + 1];
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
    pred += 12.0 * std::max(0.0, b.ctab[(size_t)A.first * L + B.second]);
    stk.push_back({A.first, B.second});
  }
  return pred;
}

// the worker contexts of the batched form (created on first use, options of the caller's context)
std::vector<smcm_ctx*>& batch_workers(smcm_ctx* ctx, int n) {
  while ((int)ctx->batch_workers.size() < n) {
    smcm_ctx* w = nullptr;
    const int rc = smcm_ctx_create(ctx->device, ctx->budget, &w);
    if (rc != SMCM_OK) fail(rc, "batch worker context: " + std::string(smcm_last_error()));
    ctx->batch_workers.push_back(w);
  }
  // the workers keep their cached blocks between calls (they are reused by the next batch); an
  // allocation of the calling context that fails takes them back first, unless they are running
  // jobs (a typed matrix built for one of their jobs allocates on the calling context)
  ctx->al.on_oom = [ctx] {
    if (ctx->batch_running.load()) return;
    for (smcm_ctx* w : ctx->batch_workers) {
      (void)hipStreamSynchronize(w->stream);
      w->al.release_cache();
      (void)hipStreamSynchronize(w->stream);
    }
  };
  for (smcm_ctx* w : ctx->batch_workers) {
    w->opts = ctx->opts;
    w->exact_cost_sum = ctx->exact_cost_sum;
    w->budget = ctx->budget;
    w->al.headroom = ctx->al.headroom;
    w->al.cache_cap = ctx->batch.worker_cache;
  }
  return ctx->batch_workers;
}

struct LargeJob : smcm_sched::Job {
  int32_t idx = 0;
  uint64_t bytes = 0;
  bool oom = false;
};

// the tiered-path jobs `ids`, several at a time on worker contexts; jobs that ran out of device
// memory next to others run again alone on ctx (their number is returned)
int32_t run_large_jobs(smcm_ctx* ctx, smcm_graph* g, std::vector<BatchChain>& ch, std::vector<int32_t> ids,
                    bool keep) {
  const int nw = std::min<int>(ctx->batch.streams, (int)ids.size());
  // workers beyond this call's count stay idle (an autotuner trial with more streams created
  // them): their cached blocks go back to the device, where this call's allocations can map them
  bool freed = false;
  for (size_t i = (size_t)std::max(nw, 0); i < ctx->batch_workers.size(); ++i) {
    smcm_ctx* w = ctx->batch_workers[i];
    if (w->al.cached_bytes == 0 && !w->al.zero_block) continue;
    w->al.release_cache();
    SMCM_HIP(hipStreamSynchronize(w->stream));
    freed = true;
  }
  if (freed) smcm::DevAlloc::trim_pool();
  if (nw <= 1) {
    for (int32_t i : ids) run_large(ctx, g, ch[i], keep);
    return 0;
  }
  std::vector<smcm_ctx*>& workers = batch_workers(ctx, nw);
  std::stable_sort(ids.begin(), ids.end(), [&](int32_t a, int32_t b) { return ch[a].est_products > ch[b].est_products; });
  std::vector<int> status(ch.size(), SMCM_OK);
  std::vector<std::string> errors(ch.size());
  {
    struct Running {
      smcm_ctx* c;
      explicit Running(smcm_ctx* x) : c(x) { c->batch_running.store(true); }
      ~Running() { c->batch_running.store(false); }
    } running(ctx);
    // declared after `running`: the workers are joined before the flag is cleared
    smcm_sched::Scheduler sched(nw, 0, ctx->budget);
    sched.start([&](smcm_sched::Job& j, int wi, bool) {
      LargeJob& job = static_cast<LargeJob&>(j);
      smcm_ctx* w = workers[wi];
      try {
        SMCM_HIP(hipSetDevice(w->device));
        sched.admit(job, job.bytes);
        struct Release {
          smcm_sched::Scheduler& s;
          LargeJob& j;
          ~Release() { s.release(j); }
        } rel{sched, job};
        run_large(w, g, ch[job.idx], keep);
      } catch (const ApiError& e) {
        job.status = e.code;
        job.error = e.msg;
      } catch (const HipError& e) {
        job.oom = e.code == (int)hipErrorOutOfMemory;
        job.status = job.oom ? SMCM_ENOMEM : SMCM_EDEVICE;
        job.error = e.what();
        (void)hipGetLastError();
      } catch (const std::exception& e) {
        job.status = SMCM_EDEVICE;
        job.error = e.what();
      }
      return true;
    });
    std::vector<int64_t> tickets;
    for (int32_t i : ids) {
      auto j = std::make_unique<LargeJob>();
      j->idx = i;
      j->bytes = (uint64_t)std::min(job_bytes(ch[i], g), 1.8e19);
      tickets.push_back(sched.submit(std::move(j)));
    }
    for (int64_t t : tickets) {
      std::unique_ptr<smcm_sched::Job> out;
      if (sched.poll(t, true, &out) != 0 || !out) continue;
      const LargeJob& j = static_cast<const LargeJob&>(*out);
      status[j.idx] = j.oom ? 1 : j.status;  // 1: out of memory next to others
      errors[j.idx] = j.error;
    }
    sched.stop();
  }
  bool released = false;
  int32_t alone = 0;
  for (smcm_ctx* w : workers) SMCM_HIP(hipStreamSynchronize(w->stream));
  for (int32_t i : ids) {
    if (status[i] == 1) {
      alone++;
      if (!released) {  // the workers' cached blocks go back to the device pool first
        for (smcm_ctx* w : workers) {
          w->al.release_cache();
          SMCM_HIP(hipStreamSynchronize(w->stream));
        }
        smcm::DevAlloc::trim_pool();
        released = true;
      }
      // a partial result of the failed attempt is not kept (run_large frees C on errors)
      ch[i].have_C = false;
      run_large(ctx, g, ch[i], keep);
    } else if (status[i] != SMCM_OK) {
      fail(status[i], "batch job: " + errors[i]);
    }
  }
  return alone;
}

}  // namespace

extern "C" {

int smcm_batch_run(smcm_ctx* ctx, smcm_graph* g, int32_t n_chains, const int32_t* path_len, const int32_t* vtypes,
                   const int32_t* etypes, const int64_t* src_off, const int32_t* src_rows, int64_t* nnz,
                   uint64_t* row_digest, smcm_result** results, smcm_batch_stats* stats) {
  return guarded([&] {
    check_ctx(ctx);
    if (!g) fail(SMCM_EINVAL, "null graph");
    if (n_chains < 0 || (n_chains > 0 && (!path_len || !vtypes || !etypes)))
      fail(SMCM_EINVAL, "bad chain arguments");
    if (src_rows && !src_off) fail(SMCM_EINVAL, "src_rows without src_off");
    DevAlloc& al = ctx->al;
    hipStream_t st = al.stream;
    const double t_all = now_s();
    smcm_batch_stats bs{};
    bs.chains = n_chains;
    // allocator retries after an out-of-memory failure, counted over this call
    auto retries = [ctx] {
      int64_t r = ctx->al.oom_retries;
      for (smcm_ctx* w : ctx->batch_workers) r += w->al.oom_retries;
      return r;
    };
    const int64_t retries0 = retries();
    if (results)
      for (int32_t i = 0; i < n_chains; ++i) results[i] = nullptr;

    // ---- chains -> jobs --------------------------------------------------------------------
    std::vector<std::vector<int32_t>> global_of(g->ntypes);  // local id -> global id, per type
    for (int32_t t = 0; t < g->ntypes; ++t) global_of[t].resize(g->dims[t]);
    for (int32_t v = 0; v < g->n; ++v) global_of[g->h_vtype[v]][g->h_local[v]] = v;
    std::vector<Member> mem(n_chains);
    std::vector<BatchChain> ch;  // jobs
    std::map<std::vector<int32_t>, int32_t> job_of;
    std::vector<std::vector<int32_t>> job_rows;  // local ids, unsorted until the end
    std::vector<bool> job_whole;
    int64_t vo = 0, eo = 0, dig = 0;
    for (int32_t i = 0; i < n_chains; ++i) {
      const int32_t L = path_len[i];
      if (L < 1 || L > 16) fail(SMCM_EINVAL, "chain " + std::to_string(i) + ": L must be in [1, 16]");
      const int32_t* vts = vtypes + vo;
      const int32_t* ets = etypes + eo;
      vo += L + 1;
      eo += L;
      for (int32_t q = 0; q <= L; ++q)
        if (vts[q] < 0 || vts[q] >= g->ntypes)
          fail(SMCM_EINVAL, "chain " + std::to_string(i) + ": vertex type not in graph");
      for (int32_t q = 0; q < L; ++q)
        if (ets[q] < 0) fail(SMCM_EINVAL, "chain " + std::to_string(i) + ": negative edge type");
      Member& m = mem[i];
      m.whole = src_rows == nullptr;
      if (!m.whole) {
        if (src_off[i + 1] < src_off[i]) fail(SMCM_EINVAL, "src_off must be non-decreasing");
        m.local.resize(src_off[i + 1] - src_off[i]);
        for (int64_t r = src_off[i]; r < src_off[i + 1]; ++r) {
          const int32_t v = src_rows[r];
          if (v < 0 || v >= g->n || g->h_vtype[v] != vts[0])
            fail(SMCM_EINVAL, "chain " + std::to_string(i) + ": source row " + std::to_string(v) +
                                  " is not a vertex of the start type");
          m.local[r - src_off[i]] = g->h_local[v];
        }
      }
      std::vector<int32_t> key;
      key.push_back(L);
      key.insert(key.end(), vts, vts + L + 1);
      key.insert(key.end(), ets, ets + L);
      auto it = job_of.find(key);
      int32_t j;
      if (it == job_of.end()) {
        j = (int32_t)ch.size();
        job_of.emplace(key, j);
        ch.emplace_back();
        ch[j].L = L;
        ch[j].vts = vts;
        ch[j].ets = ets;
        job_rows.emplace_back();
        job_whole.push_back(false);
      } else {
        j = it->second;
      }
      m.job = j;
      if (m.whole) job_whole[j] = true;
      else job_rows[j].insert(job_rows[j].end(), m.local.begin(), m.local.end());
      m.dig_off = dig;
      dig += m.whole ? g->dims[vts[0]] : (int64_t)m.local.size();
    }
    const int32_t n_jobs = (int32_t)ch.size();
    for (int32_t j = 0; j < n_jobs; ++j) {
      BatchChain& b = ch[j];
      std::vector<int32_t>& r = job_rows[j];
      std::sort(r.begin(), r.end());
      r.erase(std::unique(r.begin(), r.end()), r.end());
      if (job_whole[j] || (int64_t)r.size() == g->dims[b.vts[0]]) {
        b.src = nullptr;
        b.nsrc = 0;
        b.nrows = g->dims[b.vts[0]];
        r.clear();
        job_whole[j] = true;
      } else {
        b.src_global.resize(r.size());
        for (size_t q = 0; q < r.size(); ++q) b.src_global[q] = global_of[b.vts[0]][r[q]];
        b.src = b.src_global.data();
        b.nsrc = (int64_t)r.size();
        b.nrows = (int32_t)r.size();
      }
    }
    bs.groups = 0;

    // ---- planning: estimator + product-cost tables, one copy back -------------------------
    const double t_plan = now_s();
    std::vector<int64_t> toff(n_jobs + 1, 0);
    for (int32_t i = 0; i < n_jobs; ++i) {
      const int64_t L = ch[i].L;
      toff[i + 1] = toff[i] + (L > 1 ? L * L + L * L * L : 0);
    }
    DevBuf<double> tables(al, std::max<int64_t>(toff[n_jobs], 1));
    for (int32_t i = 0; i < n_jobs; ++i) {
      BatchChain& b = ch[i];
      Chain c{ctx, g};
      c.prepare(b.vts, b.ets, b.L, b.src, b.src ? b.nsrc : 0);
      b.src_local = c.src_local;
      if (b.src) b.rowmap = std::move(c.d_src);
      if (b.L > 1) {
        c.want_products = true;
        const int64_t LL = (int64_t)b.L * b.L;
        c.estimate(tables.p + toff[i], tables.p + toff[i] + LL);
        c.R.clear();
      }
    }
    std::vector<double> htab(std::max<int64_t>(toff[n_jobs], 1));
    if (toff[n_jobs])
      SMCM_HIP(hipMemcpyAsync(htab.data(), tables.p, sizeof(double) * toff[n_jobs], hipMemcpyDeviceToHost, st));
    SMCM_HIP(hipStreamSynchronize(st));
    const int64_t small_max = ctx->batch.small_products;
    for (int32_t i = 0; i < n_jobs; ++i) {
      BatchChain& b = ch[i];
      if (b.L == 1) {
        b.small = false;
        continue;
      }
      Chain c{ctx, g};
      c.L = b.L;
      c.want_products = true;
      const int64_t LL = (int64_t)b.L * b.L;
      c.c.assign(htab.begin() + toff[i], htab.begin() + toff[i] + LL);
      c.ptab.assign(htab.begin() + toff[i] + LL, htab.begin() + toff[i] + LL + LL * b.L);
      c.dp();
      b.pairs = c.pairs;
      b.ctab = c.c;
      // estimated products of the chosen plan
      double tot = 0;
      std::vector<std::pair<int32_t, int32_t>> stk;
      for (size_t q = 0; 2 * q < b.pairs.size(); ++q) {
        const int32_t k = b.pairs[2 * q], l = b.pairs[2 * q + 1];
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
        tot += c.ptab[((size_t)A.first * b.L + A.second) * b.L + B.second];
        stk.push_back({A.first, B.second});
      }
      b.est_products = tot;
      b.small = tot <= (double)small_max;
    }
    bs.plan_s = now_s() - t_plan;

    // ---- small jobs: one persistent launch per arena-sized group ----------------------------
    const double t_exec = now_s();
    const int64_t arena_cap = ctx->batch.arena_entries;
    const bool keep = results != nullptr;
    std::vector<int32_t> small_ids, rerun;
    for (int32_t i = 0; i < n_jobs; ++i)
      if (ch[i].small) small_ids.push_back(i);
    size_t gi = 0;
    while (gi < small_ids.size()) {
      // group: jobs whose estimated products (x2 margin) fit the arena
      std::vector<int32_t> grp;
      double need = 0;
      while (gi < small_ids.size()) {
        const double e = 2.0 * ch[small_ids[gi]].est_products + 1024.0;
        if (!grp.empty() && need + e > (double)arena_cap) break;
        need += e;
        grp.push_back(small_ids[gi++]);
      }
      bs.groups++;
      Executor ex{ctx, g};
      for (int32_t i : grp) ch[i].root = ex.add_chain(ch[i]);
      const int32_t nn = (int32_t)ex.nodes.size();
      const int32_t tr = ctx->batch.tile_rows;
      int64_t rows_total = 0, tiles = 0;
      ex.row_off.resize(nn);
      for (int32_t n = 0; n < nn; ++n) {
        ex.row_off[n] = rows_total;
        rows_total += ex.nodes[n].nrows;
        ex.nodes[n].ntiles = std::max<int32_t>(1, (int32_t)smcm::ceil_div(ex.nodes[n].nrows, tr));
        tiles += ex.nodes[n].ntiles;
      }
      if (tiles > INT32_MAX) fail(SMCM_EINVAL, "batch too large");
      DevBuf<int64_t> ostart(al, std::max<int64_t>(rows_total, 1));
      DevBuf<int32_t> olen(al, std::max<int64_t>(rows_total, 1));
      for (int32_t n = 0; n < nn; ++n) {
        ex.nodes[n].out_start = ostart.p + ex.row_off[n];
        ex.nodes[n].out_len = olen.p + ex.row_off[n];
      }
      std::vector<int32_t> poff(nn + 1, 0), par;
      for (int32_t n = 0; n < nn; ++n) {
        poff[n + 1] = poff[n] + (int32_t)ex.consumers[n].size();
        par.insert(par.end(), ex.consumers[n].begin(), ex.consumers[n].end());
      }
      // initial tiles: nodes over base relations only, largest first
      std::vector<int32_t> first;
      for (int32_t n = 0; n < nn; ++n)
        if (ex.nodes[n].ninputs == 0) first.push_back(n);
      std::stable_sort(first.begin(), first.end(),
                       [&](int32_t a, int32_t b) { return ex.nodes[a].nrows > ex.nodes[b].nrows; });
      std::vector<unsigned long long> q0;
      for (int32_t n : first)
        for (int32_t t = 0; t < ex.nodes[n].ntiles; ++t) q0.push_back((((unsigned long long)n << 32) | (unsigned)t) + 1);
      DevBuf<smcm::BatchNode> dnodes(al, nn);
      SMCM_HIP(hipMemcpyAsync(dnodes.p, ex.nodes.data(), sizeof(smcm::BatchNode) * nn, hipMemcpyHostToDevice, st));
      DevBuf<int32_t> dpoff(al, nn + 1), dpar(al, std::max<int64_t>((int64_t)par.size(), 1));
      SMCM_HIP(hipMemcpyAsync(dpoff.p, poff.data(), sizeof(int32_t) * (nn + 1), hipMemcpyHostToDevice, st));
      if (!par.empty())
        SMCM_HIP(hipMemcpyAsync(dpar.p, par.data(), sizeof(int32_t) * par.size(), hipMemcpyHostToDevice, st));
      DevBuf<int32_t> counters(al, 3 * (int64_t)nn);
      smcm::fill_i32(counters.p, 3 * (int64_t)nn, 0, st);
      DevBuf<unsigned long long> queue(al, tiles);
      smcm::fill_i64((int64_t*)queue.p, tiles, 0, st);
      if (!q0.empty())
        SMCM_HIP(hipMemcpyAsync(queue.p, q0.data(), sizeof(unsigned long long) * q0.size(), hipMemcpyHostToDevice, st));
      smcm::BatchState hst{};
      hst.tail = (int32_t)q0.size();
      DevBuf<smcm::BatchState> dstate(al, 1);
      SMCM_HIP(hipMemcpyAsync(dstate.p, &hst, sizeof(hst), hipMemcpyHostToDevice, st));
      const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(arena_cap, (int64_t)need));
      DevBuf<int32_t> acol(al, cap);
      DevBuf<uint64_t> aval(al, cap);

      smcm::BatchLaunch bl;
      bl.nodes = dnodes.p;
      bl.parent_off = dpoff.p;
      bl.parents = dpar.p;
      bl.done_tiles = counters.p;
      bl.ready = counters.p + nn;
      bl.failed = counters.p + 2 * nn;
      bl.queue = queue.p;
      bl.state = dstate.p;
      bl.arena_col = acol.p;
      bl.arena_val = aval.p;
      bl.arena_cap = (uint64_t)cap;
      bl.spin_limit = 1000000000LL;  // 10 s at 100 MHz
      bl.total_tiles = (int32_t)tiles;
      bl.tile_rows = tr;
      bl.threads = ctx->batch.threads;
      bl.sort_cap = ctx->batch.sort_cap;
      bl.log2w = ctx->batch.log2w;
      const size_t lds = smcm::batch_lds_bytes(bl.threads, bl.sort_cap, bl.log2w);
      if ((int64_t)lds > ctx->opts.lds_bytes) fail(SMCM_EINVAL, "batch: LDS shape exceeds the device limit");
      const int64_t per_cu = std::max<int64_t>(
          1, std::min<int64_t>(ctx->opts.lds_bytes / (int64_t)(lds + 2048), (2048 / bl.threads)));
      bl.grid = (int)(ctx->opts.cus * per_cu);
      smcm::KernelTimer* tm = ctx->timer.enabled ? &ctx->timer : nullptr;
      const int tok = tm ? tm->start("batch.executor", st) : -1;
      smcm::batch_execute(bl, st);
      if (tm) tm->stop(tok, st, 0.0);

      std::vector<int32_t> hfail(nn);
      SMCM_HIP(hipMemcpyAsync(hfail.data(), bl.failed, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, st));
      SMCM_HIP(hipMemcpyAsync(&hst, dstate.p, sizeof(hst), hipMemcpyDeviceToHost, st));
      SMCM_HIP(hipStreamSynchronize(st));
      if (hst.error & 2) fail(SMCM_EDEVICE, "batch executor: a queue wait timed out");
      bs.nodes += nn;
      bs.tiles += tiles;
      DevBuf<uint64_t> ddig(al, std::max<int64_t>(rows_total, 1));
      for (int32_t i : grp) {
        BatchChain& b = ch[i];
        const smcm::BatchNode& R = ex.nodes[b.root];
        if (hfail[b.root]) {
          rerun.push_back(i);
          continue;
        }
        bs.small_jobs++;
        b.h_dig.assign(R.nrows, 0);
        b.h_len.assign(R.nrows, 0);
        if (R.nrows) {
          uint64_t* d = ddig.p + ex.row_off[b.root];
          smcm::padded_digest(R.out_start, R.out_len, acol.p, aval.p, nullptr, R.nrows, d, st);
          SMCM_HIP(hipMemcpyAsync(b.h_dig.data(), d, sizeof(uint64_t) * R.nrows, hipMemcpyDeviceToHost, st));
          SMCM_HIP(hipMemcpyAsync(b.h_len.data(), R.out_len, sizeof(int32_t) * R.nrows, hipMemcpyDeviceToHost, st));
        }
        if (keep) {
          smcm::padded_to_csr(R.out_start, R.out_len, acol.p, aval.p, R.nrows, R.ncols, b.C, al);
          b.have_C = true;
          b.c_owner = ctx;
        }
      }
      SMCM_HIP(hipStreamSynchronize(st));
      if (tm) tm->flush();
    }
    bs.exec_s = now_s() - t_exec;

    // ---- large jobs (and reruns) through the per-chain path -------------------------------
    const double t_large = now_s();
    std::vector<int32_t> large;
    for (int32_t i = 0; i < n_jobs; ++i)
      if (!ch[i].small) large.push_back(i);
    large.insert(large.end(), rerun.begin(), rerun.end());
    bs.rerun_jobs = (int32_t)rerun.size();
    bs.alone_jobs = run_large_jobs(ctx, g, ch, large, keep);
    bs.large_jobs = (int32_t)large.size();
    bs.large_s = now_s() - t_large;

    // ---- every chain: its rows of its job's result ----------------------------------------
    try {
      for (int32_t i = 0; i < n_chains; ++i) {
        const Member& m = mem[i];
        BatchChain& b = ch[m.job];
        const bool jw = b.src == nullptr;
        const int32_t nr = m.whole ? g->dims[b.vts[0]] : (int32_t)m.local.size();
        std::vector<int32_t> pos(nr);
        for (int32_t r = 0; r < nr; ++r) {
          const int32_t loc = m.whole ? r : m.local[r];
          pos[r] = jw ? loc
                      : (int32_t)(std::lower_bound(b.src_global.begin(), b.src_global.end(),
                                                   global_of[b.vts[0]][loc]) -
                                  b.src_global.begin());
        }
        int64_t tot = 0;
        for (int32_t r = 0; r < nr; ++r) tot += b.h_len[pos[r]];
        if (nnz) nnz[i] = tot;
        if (row_digest)
          for (int32_t r = 0; r < nr; ++r) row_digest[m.dig_off + r] = b.h_dig[pos[r]];
        if (results) {
          Csr C;
          DevBuf<int32_t> dpos(al, std::max<int32_t>(nr, 1));
          if (nr)
            SMCM_HIP(hipMemcpyAsync(dpos.p, pos.data(), sizeof(int32_t) * nr, hipMemcpyHostToDevice, st));
          smcm::csr_gather_rows(b.C, dpos.p, nr, C, al);
          results[i] = wrap_result(ctx, g, b.vts, b.L, m.local, m.whole, C);
          SMCM_HIP(hipStreamSynchronize(st));
        }
      }
    } catch (...) {
      SMCM_HIP(hipStreamSynchronize(st));
      for (auto& b : ch)
        if (b.have_C) free_csr(b.c_owner->al, b.C);
      throw;
    }
    // the gathers above ran on this context's stream and are complete (synchronised per chain)
    for (auto& b : ch)
      if (b.have_C) free_csr(b.c_owner->al, b.C);
    bs.jobs = n_jobs;
    for (int32_t i = 0; i < n_jobs; ++i) bs.products_est += (int64_t)ch[i].est_products;
    bs.total_s = now_s() - t_all;
    bs.oom_retries = (int32_t)std::min<int64_t>(retries() - retries0, INT32_MAX);
    if (stats) *stats = bs;
  });
}

}  // extern "C"

extern "C" {

// Autotuner of the batched form: times smcm_batch_run on the given (sample) chains for the
// small-job threshold and then the executor shapes, keeping the fastest setting of each stage.
int smcm_batch_autotune(smcm_ctx* ctx, smcm_graph* g, int32_t n_chains, const int32_t* path_len,
                        const int32_t* vtypes, const int32_t* etypes, const int64_t* src_off,
                        const int32_t* src_rows, int32_t reps, smcm_batch_tuning* out) {
  return guarded([&] {
    check_ctx(ctx);
    if (reps < 1) reps = 1;
    auto saved = ctx->batch;
    auto time_it = [&]() {
      double best = 1e300;
      for (int32_t r = 0; r < reps; ++r) {
        const double t0 = now_s();
        const int rc = smcm_batch_run(ctx, g, n_chains, path_len, vtypes, etypes, src_off, src_rows, nullptr, nullptr,
                                      nullptr, nullptr);
        if (rc != SMCM_OK) return -1.0;  // this setting does not run (e.g. LDS)
        best = std::min(best, now_s() - t0);
      }
      return best;
    };
    smcm_batch_tuning t{};
    t.trials = 0;
    // 0. how many tiered-path jobs run at once (worker streams)
    {
      const double base = time_it();
      double bs = base;
      int best_streams = ctx->batch.streams;
      for (int ns : {1, 2, 4, 8}) {
        ctx->batch.streams = ns;
        const double s = time_it();
        t.trials++;
        if (s >= 0 && (bs < 0 || s < bs)) {
          bs = s;
          best_streams = ns;
        }
      }
      ctx->batch.streams = best_streams;
    }
    // 1. which jobs go to the executor
    const int64_t thresholds[] = {0, 1 << 20, 1 << 22, 1 << 24, 1 << 25, 1 << 27};
    double best = time_it();
    t.baseline_s = best;
    int64_t best_thr = ctx->batch.small_products;
    for (int64_t thr : thresholds) {
      ctx->batch.small_products = thr;
      const double s = time_it();
      t.trials++;
      if (s >= 0 && s < best) {
        best = s;
        best_thr = thr;
      }
    }
    ctx->batch.small_products = best_thr;
    // 2. executor shape (only matters when jobs use the executor)
    if (best_thr > 0) {
      const int shapes[][2] = {{256, 512}, {256, 1024}, {256, 2048}, {512, 512}, {512, 1024}};
      const int tiles[] = {8, 16, 32, 64};
      auto bestb = ctx->batch;
      for (auto& sh : shapes) {
        for (int tr : tiles) {
          ctx->batch.threads = sh[0];
          ctx->batch.sort_cap = sh[1];
          ctx->batch.tile_rows = tr;
          const double s = time_it();
          t.trials++;
          if (s >= 0 && s < best) {
            best = s;
            bestb = ctx->batch;
          }
        }
      }
      ctx->batch = bestb;
    }
    (void)saved;
    // the trials' cached blocks are sized for the sample and its other settings: they go back to
    // the device, so the tuned batch does not meet full caches (each out-of-memory retry on the
    // way costs a synchronisation, a pool trim and new mappings)
    for (smcm_ctx* w : ctx->batch_workers) {
      w->al.release_cache();
      SMCM_HIP(hipStreamSynchronize(w->stream));
    }
    ctx->al.release_cache();
    SMCM_HIP(hipStreamSynchronize(ctx->stream));
    smcm::DevAlloc::trim_pool();
    t.best_s = best;
    t.small_products = ctx->batch.small_products;
    t.threads = ctx->batch.threads;
    t.sort_cap = ctx->batch.sort_cap;
    t.tile_rows = ctx->batch.tile_rows;
    t.streams = ctx->batch.streams;
    if (out) *out = t;
  });
}

}  // extern "C"
