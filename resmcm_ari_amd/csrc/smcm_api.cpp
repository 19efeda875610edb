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

#include "api_internal.h"

using namespace smcm_host;
using smcm::Csr;
using smcm::DevAlloc;
using smcm::DevBuf;
using smcm::HipError;

namespace {

int run_chain(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
              const int32_t* src_rows, int64_t n_src, const int32_t* forced_pairs, smcm_result** out,
              smcm_times* times) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out) fail(SMCM_EINVAL, "null output");
    *out = nullptr;
    Chain ch{ctx, g};
    ch.prepare(vtypes, etypes, L, src_rows, n_src);
    ch.provider = ctx->plan_cost >= smcm::kProviderMetaAC ? ctx->plan_cost : 0;
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
    ctx->opts.device_cus = cus;
    size_t freeb = 0, totalb = 0;
    SMCM_HIP(hipMemGetInfo(&freeb, &totalb));
    ctx->budget = hbm_budget ? hbm_budget : (uint64_t)(0.9 * (double)freeb);
    // the allocation-free entry point's scratch on this device: its address, looked up once
    (void)noalloc_scratch(device);
    *out = ctx.release();
  });
}

int smcm_ctx_destroy(smcm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    if (ctx->live != 0)
      fail(SMCM_EINVAL, std::to_string(ctx->live) + " graph/result handle(s) still alive; free them first");
    (void)hipSetDevice(ctx->device);
    for (smcm_ctx* w : ctx->batch_workers) smcm_ctx_destroy(w);
    ctx->batch_workers.clear();
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
    } else if (k == "window_major") {
      if (v != 0 && v != 1) fail(SMCM_EINVAL, "window_major must be 0 or 1");
      o.window_major = (int)v;
    } else if (k == "window_presence") {
      if (v < 0 || v > 2) fail(SMCM_EINVAL, "window_presence must be 0, 1 or 2");
      o.window_presence = (int)v;
    } else if (k == "sym_merge") {
      if (v < 0 || v > 2) fail(SMCM_EINVAL, "sym_merge must be 0, 1 or 2");
      o.sym_merge = (int)v;
    } else if (k == "copy_tier") {
      if (v != 0 && v != 1) fail(SMCM_EINVAL, "copy_tier must be 0 or 1");
      o.copy_tier = (int)v;
    } else if (k == "sym_window_cols") {
      if (v < 64) fail(SMCM_EINVAL, "sym_window_cols must be >= 64");
      o.sym_window_cols = v;
    } else if (k == "window_min_products") {
      if (v < 0) fail(SMCM_EINVAL, "window_min_products must be >= 0");
      o.window_min_products = v;
    } else if (k == "window_all_nwin") {
      if (v < 0) fail(SMCM_EINVAL, "window_all_nwin must be >= 0");
      o.window_all_nwin = v;
    } else if (k == "window_u32") {
      if (v != 0 && v != 1) fail(SMCM_EINVAL, "window_u32 must be 0 or 1");
      o.window_u32 = (int)v;
    } else if (k == "bbits_bytes") {
      if (v < 0) fail(SMCM_EINVAL, "bbits_bytes must be >= 0");
      o.bbits_bytes = v;
    } else if (k == "window_mixed_min") {
      if (v < 0) fail(SMCM_EINVAL, "window_mixed_min must be >= 0");
      o.window_mixed_min = v;
    } else if (k == "dense_row_products") {
      if (v < 0) fail(SMCM_EINVAL, "dense_row_products must be >= 0");
      o.dense_row_products = v;
    } else if (k == "dense_cells") {
      if (v < 0 || v > ((int64_t)1 << 28)) fail(SMCM_EINVAL, "dense_cells must be in [0, 2^28]");
      o.dense_cells = v;
    } else if (k == "bitmap_vcap") {
      if (v < -1) fail(SMCM_EINVAL, "bitmap_vcap must be -1 or >= 0");
      o.bitmap_vcap = v;
    } else if (k == "force_tier") {
      if (v < -1 || v > 2) fail(SMCM_EINVAL, "force_tier must be -1..2");
      o.force_tier = (int)v;
    } else if (k == "exact_cost_sum") {
      ctx->exact_cost_sum = v != 0;
    } else if (k == "wave_bitmap_cols") {
      if (v < 0) fail(SMCM_EINVAL, "wave_bitmap_cols must be >= 0");
      o.wave_bitmap_cols = v;
    } else if (k == "wave_bitmap_cap") {
      if (v < 64 || v > 65535) fail(SMCM_EINVAL, "wave_bitmap_cap must be in [64, 65535]");
      o.wave_bitmap_cap = v;
    } else if (k == "wave_bitmap_max_ub") {
      if (v < 64 || v > 65535) fail(SMCM_EINVAL, "wave_bitmap_max_ub must be in [64, 65535]");
      o.wave_bitmap_max_ub = v;
    } else if (k == "plan_cost") {
      if (v < 0 || v > 5)
        fail(SMCM_EINVAL, "plan_cost must be 0 (reference), 1 (products), 2 (MetaAC), 3 (MNC), 4 (DMap) or 5 (LGraph)");
      ctx->plan_cost = (int)v;
    } else if (k == "lgraph_seed") {
      ctx->lgraph_seed = (uint64_t)v;
    } else if (k == "batch_small_products") {
      if (v < 0) fail(SMCM_EINVAL, "batch_small_products must be >= 0");
      ctx->batch.small_products = v;
    } else if (k == "batch_threads") {
      if (v != 256 && v != 512) fail(SMCM_EINVAL, "batch_threads must be 256 or 512");
      ctx->batch.threads = (int)v;
    } else if (k == "batch_sort_cap") {
      if (v != 512 && v != 1024 && v != 2048) fail(SMCM_EINVAL, "batch_sort_cap must be 512, 1024 or 2048");
      ctx->batch.sort_cap = (int)v;
    } else if (k == "batch_tile_rows") {
      if (v < 1 || v > 64) fail(SMCM_EINVAL, "batch_tile_rows must be in [1, 64]");
      ctx->batch.tile_rows = (int)v;
    } else if (k == "batch_log2w") {
      if (v < 8 || v > 14) fail(SMCM_EINVAL, "batch_log2w must be in [8, 14]");
      ctx->batch.log2w = (int)v;
    } else if (k == "batch_arena_mb") {
      if (v < 1) fail(SMCM_EINVAL, "batch_arena_mb must be >= 1");
      ctx->batch.arena_entries = v * ((int64_t)1 << 20) / 12;
    } else if (k == "batch_streams") {
      if (v < 1 || v > 16) fail(SMCM_EINVAL, "batch_streams must be in [1, 16]");
      ctx->batch.streams = (int)v;
    } else if (k == "kernel_timing") {
      ctx->timer.enabled = v != 0;
    } else if (k == "batch_cache_mb") {
      if (v < 0) fail(SMCM_EINVAL, "batch_cache_mb must be >= 0");
      ctx->batch.worker_cache = (size_t)v << 20;
    } else if (k == "server_cache_mb") {
      if (v < 0) fail(SMCM_EINVAL, "server_cache_mb must be >= 0");
      ctx->server.worker_cache = (size_t)v << 20;
    } else if (k == "headroom_mb") {
      if (v < 0) fail(SMCM_EINVAL, "headroom_mb must be >= 0");
      ctx->al.headroom = (size_t)v << 20;
    } else if (k == "small_cache_mb") {
      if (v < 0) fail(SMCM_EINVAL, "small_cache_mb must be >= 0");
      ctx->al.small_cache_cap = (size_t)v << 20;
    } else if (k == "oom_inject") {
      if (v < 0) fail(SMCM_EINVAL, "oom_inject must be >= 0");
      ctx->al.oom_inject = v;
    } else if (k == "budget") {
      if (v <= 0) fail(SMCM_EINVAL, "budget must be > 0 bytes");
      ctx->budget = (uint64_t)v;
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
    else if (k == "window_major") *v = o.window_major;
    else if (k == "window_min_products") *v = o.window_min_products;
    else if (k == "window_all_nwin") *v = o.window_all_nwin;
    else if (k == "dense_cells") *v = o.dense_cells;
    else if (k == "dense_row_products") *v = o.dense_row_products;
    else if (k == "window_mixed_min") *v = o.window_mixed_min;
    else if (k == "bbits_bytes") *v = o.bbits_bytes;
    else if (k == "window_u32") *v = o.window_u32;
    else if (k == "sym_window_cols") *v = o.sym_window_cols;
    else if (k == "sym_merge") *v = o.sym_merge;
    else if (k == "copy_tier") *v = o.copy_tier;
    else if (k == "window_presence") *v = o.window_presence;
    else if (k == "force_tier") *v = o.force_tier;
    else if (k == "exact_cost_sum") *v = ctx->exact_cost_sum;
    else if (k == "plan_cost") *v = ctx->plan_cost;
    else if (k == "lgraph_seed") *v = (int64_t)ctx->lgraph_seed;
    else if (k == "wave_bitmap_cols") *v = o.wave_bitmap_cols;
    else if (k == "wave_bitmap_cap") *v = o.wave_bitmap_cap;
    else if (k == "wave_bitmap_max_ub") *v = o.wave_bitmap_max_ub;
    else if (k == "batch_small_products") *v = ctx->batch.small_products;
    else if (k == "batch_threads") *v = ctx->batch.threads;
    else if (k == "batch_sort_cap") *v = ctx->batch.sort_cap;
    else if (k == "batch_tile_rows") *v = ctx->batch.tile_rows;
    else if (k == "batch_log2w") *v = ctx->batch.log2w;
    else if (k == "batch_arena_mb") *v = ctx->batch.arena_entries * 12 / ((int64_t)1 << 20);
    else if (k == "batch_streams") *v = ctx->batch.streams;
    else if (k == "kernel_timing") *v = ctx->timer.enabled;
    else if (k == "oom_inject") *v = ctx->al.oom_inject;
    else if (k == "headroom_mb") *v = (int64_t)(ctx->al.headroom >> 20);
    else if (k == "small_cache_mb") *v = (int64_t)(ctx->al.small_cache_cap >> 20);
    else if (k == "server_cache_mb") *v = (int64_t)(ctx->server.worker_cache >> 20);
    else if (k == "batch_cache_mb")
      *v = ctx->batch.worker_cache == SIZE_MAX ? INT64_MAX : (int64_t)(ctx->batch.worker_cache >> 20);
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
    g->h_bins.resize(g->ntypes);
    for (int32_t t = 0; t < g->ntypes; ++t) g->h_bins[t].reserve(g->dims[t]);
    for (int32_t v = 0; v < n; ++v) g->h_bins[vtype[v]].push_back(v);
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
    ch.provider = ctx->plan_cost >= smcm::kProviderMetaAC ? ctx->plan_cost : 0;
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
    ch.provider = ctx->plan_cost >= smcm::kProviderMetaAC ? ctx->plan_cost : 0;
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

int smcm_cost_estimate(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                       int32_t provider, int32_t accuracy, double* estimate, double* c_table, double* m_table,
                       int32_t* s_table) {
  return guarded([&] {
    check_ctx(ctx);
    if (!estimate) fail(SMCM_EINVAL, "null output");
    if (provider < smcm::kProviderMetaAC || provider > smcm::kProviderLGraph)
      fail(SMCM_EINVAL, "provider must be 2 (MetaAC), 3 (MNC), 4 (DMap) or 5 (LGraph)");
    Chain ch{ctx, g};
    ch.prepare(vtypes, etypes, L, nullptr, 0);
    smcm::ProviderOut po;
    ch.provider_tables(provider, accuracy != 0, po);
    *estimate = po.estimate;
    const size_t LL = (size_t)L * L;
    if (c_table) std::memcpy(c_table, po.c.data(), sizeof(double) * LL);
    if (m_table) std::memcpy(m_table, po.m.data(), sizeof(double) * LL);
    if (s_table) std::memcpy(s_table, po.s.data(), sizeof(int32_t) * LL);
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

}  // extern "C"

namespace {
// Host CSR operand of smcm_spgemm_csr: ptr[0] == 0, non-decreasing row pointers, a column array
// whenever there are entries, columns in [0, ncols) and, for a right operand, ascending rows.
// returns whether every row is strictly ascending (no column repeated within a row)
bool check_host_csr(const char* name, int32_t rows, int32_t ncols, const int64_t* ptr, const int32_t* col,
                    bool sorted) {
  using smcm_host::fail;
  const std::string nm(name);
  if (ptr[0] != 0) fail(SMCM_EINVAL, nm + ": row pointer 0 must be 0");
  for (int32_t r = 0; r < rows; ++r)
    if (ptr[r + 1] < ptr[r]) fail(SMCM_EINVAL, nm + ": row pointers decrease at row " + std::to_string(r));
  if (ptr[rows] > 0 && !col) fail(SMCM_EINVAL, nm + ": null column array");
  bool strict = true;
  for (int32_t r = 0; r < rows; ++r)
    for (int64_t x = ptr[r]; x < ptr[r + 1]; ++x) {
      if (col[x] < 0 || col[x] >= ncols) fail(SMCM_EINVAL, nm + ": column out of range");
      if (x > ptr[r] && col[x] <= col[x - 1]) {
        strict = false;
        if (sorted && col[x] < col[x - 1]) fail(SMCM_EINVAL, nm + ": rows must be sorted");
      }
    }
  return strict;
}
}  // namespace

extern "C" {

int smcm_spgemm_csr(smcm_ctx* ctx, int32_t m, int32_t k, int32_t n, const int64_t* a_ptr, const int32_t* a_col,
                    const uint64_t* a_val, const int64_t* b_ptr, const int32_t* b_col, const uint64_t* b_val,
                    smcm_result** out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out || !a_ptr || !b_ptr) fail(SMCM_EINVAL, "null argument");
    *out = nullptr;
    if (m < 0 || k < 0 || n < 0) fail(SMCM_EINVAL, "negative dimension");
    // every check runs before the first upload, so a malformed operand launches no kernel
    check_host_csr("A", m, k, a_ptr, a_col, false);
    const bool b_distinct = check_host_csr("B", k, n, b_ptr, b_col, true);
    const int64_t annz = a_ptr[m], bnnz = b_ptr[k];
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
    B.nrows = k;
    B.ncols = n;
    B.nnz = bnnz;
    B.distinct = b_distinct;
    smcm::SpgemmStats ss;
    try {  // operands uploaded so far are freed when an allocation fails
      A.ptr = (int64_t*)up(a_ptr, sizeof(int64_t) * (m + 1));
      A.col = (int32_t*)up(a_col, sizeof(int32_t) * annz);
      A.val = a_val ? (uint64_t*)up(a_val, sizeof(uint64_t) * annz) : nullptr;
      B.ptr = (int64_t*)up(b_ptr, sizeof(int64_t) * (k + 1));
      B.col = (int32_t*)up(b_col, sizeof(int32_t) * bnnz);
      B.val = b_val ? (uint64_t*)up(b_val, sizeof(uint64_t) * bnnz) : nullptr;
      smcm::spgemm(A, B, nullptr, 0, C, ctx->opts, &ss, al, true, ctx->timer.enabled ? &ctx->timer : nullptr);
    } catch (...) {
      free_csr(al, A);
      free_csr(al, B);
      free_csr(al, C);  // what the product allocated before the failure
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
