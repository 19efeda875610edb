/*
 * smcm.h - C ABI of the MI355X sparse matrix chain multiplication library (libsmcm.so).
 *
 * The library keeps a heterogeneous graph resident in HBM, builds the typed relation matrices
 * of a meta-path on the device, runs the row-wise sparsity estimator, the chain-order DP and the
 * row-wise Gustavson SpGEMM (uint64 path counts), and returns the result matrix with the
 * reference's four timings.  Plain pointers and sizes only; no C++ or torch types.
 *
 * Result semantics (the reference result is Boolean and unsorted; this adds counts):
 *   rows    = the vertices of the start type in ascending global id (or the caller's source
 *             rows, in the caller's order);
 *   columns = global vertex ids of the end type, ascending within a row;
 *   values  = number of meta-path instances, mod 2^64.  The structure is the Boolean product:
 *             an entry whose count wraps to 0 mod 2^64 is still stored.
 *
 * Status codes: 0 ok; SMCM_EINVAL bad argument / malformed meta-path / type mismatch;
 * SMCM_ENOMEM device allocation failed; SMCM_EBUDGET predicted result exceeds the budget;
 * SMCM_EDEVICE any other HIP error.  Nothing throws across the ABI; the message of the last
 * failure on the calling thread is returned by smcm_last_error().
 *
 * Threading: one smcm_ctx per host thread (each owns a HIP stream); handles created from a
 * context are used with that context only.
 */
#ifndef SMCM_H
#define SMCM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMCM_OK 0
#define SMCM_EINVAL (-22)
#define SMCM_ENOMEM (-12)
#define SMCM_EBUDGET (-28)
#define SMCM_EDEVICE (-5)
#define SMCM_EAGAIN (-11) /* smcm_server_poll: the query has not finished */
#define SMCM_ECANCELED (-125) /* a query's status after smcm_server_cancel stopped it */

typedef struct smcm_ctx smcm_ctx;
typedef struct smcm_graph smcm_graph;
typedef struct smcm_result smcm_result;

/* The reference's globals es_time, dp_time, spgemm_time, smcm_time (efficient/resmcm.cpp:18-21),
 * in seconds: total covers estimator + DP + all products, as smcm_time (:679-771). */
typedef struct smcm_times {
  double total;
  double spgemm;
  double dp;
  double est;
} smcm_times;

/* Per-call statistics of the last smcm_chain_run* on a context. */
typedef struct smcm_stats {
  int64_t products;      /* intermediate products executed (sum over the plan's SpGEMMs) */
  int64_t nnz;           /* nnz of the result */
  int64_t rows_tier[4];  /* rows handled by tier: empty, wave, sort, bitmap (summed over SpGEMMs) */
  int32_t n_spgemm;      /* SpGEMMs executed */
  int32_t pairs[2 * 15]; /* executed plan (get_optimal_chain_order format) */
  double c_root;         /* estimated nnz of the whole chain (c_[0][L-1]) */
  int64_t ref_nnz;       /* nnz as the reference counts it: equals nnz for L >= 2; for L == 1 the
                            reference returns matrix 0 itself, duplicate entries included
                            (efficient/resmcm.cpp:670-677) */
} smcm_stats;

/* ---- context ---------------------------------------------------------------------------- */
/* The context (device, stream, allocator), options, error text, statistics and timing replace
 * the reference's process-global state: the timers es_time / dp_time / spgemm_time / smcm_time
 * (efficient/resmcm.cpp:18-21), the OpenMP thread count of main (:795-796) and its thrown
 * messages (:257, :295).  smcm_stats, smcm_ctx_sync, smcm_kernel_stats and smcm_copy_bandwidth
 * have no counterpart there.
 * hbm_budget: bytes the library may hold for results and intermediates (0 = 90% of free HBM). */
int smcm_ctx_create(int device, uint64_t hbm_budget, smcm_ctx** out);
int smcm_ctx_destroy(smcm_ctx* ctx);
/* Tuning / mode keys: "s_max", "m_max", "dense_div", "heavy_ub", "d_threads", "force_tier",
 * "exact_cost_sum" (1 = sum c_ij in row order on one lane: bit-exact with the reference at one
 * thread), "kernel_timing" (1 = time every kernel with HIP events on the context's stream),
 * "lds_bytes", "plan_cost" (the DP's cost: 0 the reference's estimated sizes, 1 estimated
 * products, 2-5 the reference's other providers MetaAC / MNC / DMap / LGraph, see
 * smcm_cost_estimate), "lgraph_seed", "window_major", "window_min_products", "window_cols",
 * "w_threads", "window_all_nwin", "window_presence", "sym_window_cols", "sym_merge", "dense_cells"
 * (products with rows x columns up to this many cells accumulate whole in a dense device array,
 * default 2^23, 0 = never), "dense_row_products" (products over more than 4096 columns whose rows
 * average this many products or more stay with the row tiers; default 10000, 0 = no limit),
 * "bbits_bytes" (largest right operand kept as a column bitmap for the
 * symbolic pass, default 64 MiB, 0 = never), "window_u32" (1 = 32-bit window accumulators when
 * every count of a product is bounded below 2^32, the default; 0 = always 64-bit),
 * "window_mixed_min" (products with at least this many products, default 2^26, whose whole bound
 * does not fit 32 bits bound each row: rows that fit take 32-bit windows, the rest a 64-bit
 * launch over the same windows), "oom_inject"
 * (tests of the out-of-memory recovery: the n-th device allocation from now reports out of
 * memory once, then the allocator releases its cache and tries again), "headroom_mb" (device
 * memory kept free for the HIP runtime, default 2048: the scratch of a kernel with a private
 * segment is allocated at its dispatch, and a failure there aborts the process, so an allocation
 * that would leave less free fails with SMCM_ENOMEM instead), "small_cache_mb" (freed device
 * blocks below 64 MiB the context keeps for reuse, default 256; blocks of 64 MiB and more are always
 * kept), "budget" (device
 * bytes the product-aware plan may assume: products whose estimated operands and output exceed it
 * are not planned when another plan avoids them). */
int smcm_set_option(smcm_ctx* ctx, const char* key, int64_t value);
int smcm_get_option(smcm_ctx* ctx, const char* key, int64_t* value);
const char* smcm_last_error(void);
int smcm_last_stats(smcm_ctx* ctx, smcm_stats* out);
/* Wait for all work queued on the context's stream. */
int smcm_ctx_sync(smcm_ctx* ctx);
/* Per-kernel totals collected while "kernel_timing" is on: entry idx (0-based; returns
 * SMCM_EINVAL past the end) gives the kernel name, summed event time, launch count and
 * algorithmic bytes of those launches. */
int smcm_kernel_stats(smcm_ctx* ctx, int32_t idx, char* name, int32_t cap, double* total_ms,
                      int64_t* count, double* bytes);
int smcm_kernel_stats_reset(smcm_ctx* ctx);
/* HBM copy bandwidth of a plain 16-byte-vector copy kernel over two buffers of `bytes` bytes,
 * `iters` launches: *gbps = (read + write bytes) / time.  No counterpart in the reference: the
 * measured bandwidth reference for the roofline figures. */
int smcm_copy_bandwidth(smcm_ctx* ctx, uint64_t bytes, int32_t iters, double* gbps);

/* ---- graph (replaces HeteroGraph + the per-chain BoolMatrix construction,
 *      efficient/resmcm.cpp:181-240, :374-396) ------------------------------------------- */
/* Copies the HIN to the device: vtype[n], adj_ptr[n+1], adj_nbr[m], adj_etype[m] (m = adj_ptr[n];
 * neighbour order = file order).  Builds the per-type index on the device. */
int smcm_graph_upload(smcm_ctx* ctx, int32_t n, const int32_t* vtype, const int64_t* adj_ptr,
                      const int32_t* adj_nbr, const int32_t* adj_etype, smcm_graph** out);
int smcm_graph_free(smcm_graph* g);
/* dims[t] for t < min(cap, ntypes); returns ntypes through *ntypes. */
int smcm_graph_dims(const smcm_graph* g, int32_t* dims, int32_t cap, int32_t* ntypes);

/* ---- chain (replaces build_pair / SSP_dynamic, efficient/resmcm.cpp:653-779) ------------ */
/* vtypes[L+1], etypes[L]; src_rows: global ids of start-type vertices (NULL = all of them). */
int smcm_chain_run(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes,
                   int32_t L, const int32_t* src_rows, int64_t n_src, smcm_result** out,
                   smcm_times* times);
/* Row-blocked run for results larger than device memory: the root product is materialised in
 * HBM in row blocks of at most block_bytes of output, each block is reduced to its row count
 * and (when row_digest != NULL) per-row digests (as smcm_result_row_digest), then its memory
 * is reused.  *nnz receives the result's nnz, row_digest[nrows] the digests.  The reference's
 * driver likewise only keeps the nnz of a result (efficient/resmcm.cpp:859-873). */
int smcm_chain_run_blocked(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes,
                           const int32_t* etypes, int32_t L, const int32_t* src_rows,
                           int64_t n_src, uint64_t block_bytes, int64_t* nnz, uint64_t* row_digest,
                           smcm_times* times);
/* Same, executing the given plan (L-1 pairs in get_optimal_chain_order format) instead of the
 * DP's; the estimator still runs (its densities drive row binning). */
int smcm_chain_run_plan(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes,
                        const int32_t* etypes, int32_t L, const int32_t* src_rows, int64_t n_src,
                        const int32_t* pairs, smcm_result** out, smcm_times* times);
/* Estimator + DP only (rows_sparse_optimal_matrix_chain_order, :488-541, and
 * get_optimal_chain_order, :563-573).  c_table, m_table: L*L doubles; s_table: L*L; pairs:
 * 2*(L-1).  Any output may be NULL. */
int smcm_chain_plan(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes,
                    int32_t L, const int32_t* src_rows, int64_t n_src, double* c_table,
                    double* m_table, int32_t* s_table, int32_t* pairs);
/* Estimated product counts of every split, p_table[(i*L + k)*L + j] for i <= k < j: the cost
 * of X_ik * X_{k+1,j} (X_ik = M_i..M_k) from column densities of X_ik and the estimator's row
 * densities of X_{k+1,j}; other entries are 0.  p_table: L*L*L doubles.  No counterpart in the
 * reference: the option plan_cost = 1 makes the DP minimise these products plus the estimated
 * intermediate sizes instead of the sizes alone (:519-538); every plan yields the same result. */
int smcm_chain_products_estimate(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes,
                                 const int32_t* etypes, int32_t L, const int32_t* src_rows,
                                 int64_t n_src, double* p_table);
/* Estimated nnz of the whole chain, c_[0][L-1] (accuracy/res.cpp:494-523, 564-600). */
int smcm_estimate(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes,
                  int32_t L, const int32_t* src_rows, int64_t n_src, double* est_nnz);
/* Allocation-free variant of smcm_estimate (no source subset).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Writes c_[0][L-1] to the
 * device double d_est_nnz; the work is queued on the context's stream and the call returns
 * without waiting (smcm_ctx_sync, or the caller's own stream order, sees the result).  It
 * allocates and frees nothing: the estimator's scratch comes from one 64 MiB device array per
 * GPU, so chains whose row densities and pass temporaries exceed it are rejected (SMCM_EINVAL)
 * before any kernel is queued for the excess; the meta-path's typed matrices must already be on
 * the device (any earlier smcm_estimate / smcm_chain_run over it on this graph builds them).
 * Calls that overlap on one GPU (from other contexts or threads) race on the shared array. */
int smcm_estimate_noalloc(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes,
                          int32_t L, double* d_est_nnz);
/* The reference's other DP cost providers over the whole chain (no source subset), selected by
 * `provider`: 2 MetaAC (efficient/meta.cpp:448-566), 3 MNC sketches (efficient/mnc.cpp:313-800),
 * 4 DMap density maps (efficient/dm.cpp:52-620), 5 LGraph rank propagation
 * (efficient/lg.cpp:52-740; ranks from the seeded generator of option lgraph_seed, where the
 * reference seeds mt19937 from random_device).  accuracy = 0: the planner's tables (the same
 * plan the option plan_cost = provider makes smcm_chain_run execute); accuracy = 1: the
 * estimate-only programs (accuracy/{meta,mnc,dm,lg}.cpp).  *estimate = the number those
 * programs print for the meta-path (MetaAC and DMap: density * n * n; MNC and LGraph: nnz).
 * c_table (the provider's density or size per interval), m_table (DP costs): L*L doubles;
 * s_table: L*L.  Any table may be NULL. */
int smcm_cost_estimate(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes,
                       int32_t L, int32_t provider, int32_t accuracy, double* estimate,
                       double* c_table, double* m_table, int32_t* s_table);
/* Row densities r_ij of chain position i for j = i..L-1: out[u*(L-i) + (j-i)], u over the rows
 * of M_i (cap = number of doubles available in out). */
int smcm_estimator_rows(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes,
                        const int32_t* etypes, int32_t L, const int32_t* src_rows, int64_t n_src,
                        int32_t i, double* out, int64_t cap);

/* ---- batched form (no counterpart in the reference) ------------------------------------- */
typedef struct smcm_batch_stats {
  int32_t chains, jobs;                    /* jobs: distinct meta-paths */
  int32_t small_jobs, large_jobs, rerun_jobs, groups;  /* large jobs run on "batch_streams"
                                              concurrent worker streams */
  int64_t nodes, tiles, products_est;
  double plan_s, exec_s, large_s, total_s; /* host wall-clock seconds */
  int32_t alone_jobs;                      /* large jobs that ran out of device memory next to
                                              others and ran again by themselves */
  int32_t oom_retries;                     /* device allocations of the call (all worker
                                              streams) that failed once and succeeded after the
                                              allocator released its cached blocks */
} smcm_batch_stats;
/* Runs n_chains chains in one call.  Chain i has path_len[i] = L_i matrices; its vertex types
 * are the next L_i+1 entries of vtypes and its edge types the next L_i entries of etypes.
 * src_rows == NULL: every chain starts from all vertices of its start type; otherwise chain i
 * starts from src_rows[src_off[i] .. src_off[i+1]) (global ids).  Outputs (any may be NULL):
 * nnz[i]; row_digest, the per-row digests (as smcm_result_row_digest) of every result, chain
 * after chain; results[i], result handles (freed with smcm_result_free).  Chains with the same
 * meta-path run as one job over the union of their source rows (each chain's result is the
 * job's rows at its sources).  Plans minimise estimated products (as option plan_cost = 1);
 * jobs with few estimated products run as nodes of one persistent device launch, the others
 * through the per-chain path.  Every result equals smcm_chain_run's for the same chain bit for
 * bit. */
int smcm_batch_run(smcm_ctx* ctx, smcm_graph* g, int32_t n_chains, const int32_t* path_len,
                   const int32_t* vtypes, const int32_t* etypes, const int64_t* src_off,
                   const int32_t* src_rows, int64_t* nnz, uint64_t* row_digest,
                   smcm_result** results, smcm_batch_stats* stats);

typedef struct smcm_batch_tuning {
  double baseline_s, best_s;     /* sample time with the settings on entry / the chosen ones */
  int64_t small_products;        /* chosen options (set on the context) */
  int32_t threads, sort_cap, tile_rows, trials;
  int32_t streams;               /* concurrent worker streams of the tiered-path jobs */
  int32_t pad;
} smcm_batch_tuning;
/* Autotuner of the batched form: runs smcm_batch_run on the given chains (a sample of the
 * workload), best of reps runs per setting: first the number of worker streams
 * (batch_streams), then the executor threshold (batch_small_products), then the executor shape
 * (batch_threads, batch_sort_cap, batch_tile_rows).  Leaves the fastest settings on ctx and
 * reports them. */
int smcm_batch_autotune(smcm_ctx* ctx, smcm_graph* g, int32_t n_chains, const int32_t* path_len,
                        const int32_t* vtypes, const int32_t* etypes, const int64_t* src_off,
                        const int32_t* src_rows, int32_t reps, smcm_batch_tuning* out);

/* ---- serving form (no counterpart in the reference) --------------------------------------
 * A request loop over graphs resident in HBM.  Worker threads (slots), each with its own HIP
 * stream and allocator, take submitted queries as they arrive; each query is planned (estimator
 * + DP), admitted when its estimator-predicted bytes fit the budget next to the running queries
 * (alone it always runs), and executed with results reduced to per-row digests and nnz.  A
 * query predicted above half the budget runs as independent blocks of its source rows, each
 * planned on its own and predicted at about a quarter of the budget.
 * Products of leading positions (sources included) are kept in a device prefix cache (LRU) and
 * reused by later queries with the same graph, sources and leading matrices.  Results are the
 * same bit for bit as smcm_chain_run's.  The context must not be used by other calls while a
 * server created from it is running. */
typedef struct smcm_server smcm_server;
typedef struct smcm_server_opts {
  int32_t slots;             /* worker threads / concurrent queries (1..16) */
  int32_t plan_cost;         /* 0: the reference DP; 1: estimated products + sizes */
  uint64_t budget_bytes;     /* admission budget of predicted bytes in flight */
  uint64_t cache_bytes;      /* prefix cache capacity (0 = no cache) */
  uint64_t cache_entry_max;  /* largest product kept in the cache */
  uint64_t block_bytes;      /* roots predicted larger than this are streamed in row blocks */
  int32_t large_slots;       /* workers reserved for large queries (0 = no lanes) */
  int32_t large_cu_percent;  /* share of the CUs the large-lane workers size their persistent
                                kernels for (0 or 100 = all) */
  uint64_t large_bytes;      /* queries predicted above this run on the large-lane workers */
} smcm_server_opts;
typedef struct smcm_query_out {
  int64_t nnz;
  double t_submit, t_start, t_done; /* seconds; t_start = admission */
  uint64_t predicted_bytes;
  int32_t status;                   /* SMCM_OK or the error of the query */
  int32_t cache_k;                  /* leading positions 0..cache_k served from the cache (0: none) */
  int32_t blocks;                   /* > 0: run as this many independent blocks of source rows */
  int32_t reruns;                   /* times the query ran again by itself after running out of
                                       device memory next to other queries */
  char error[160];                  /* message when status != SMCM_OK */
} smcm_query_out;
/* graphs: uploaded to ctx; opts NULL = 6 slots (2 of them for large queries, above 4 GiB
 * predicted, with kernels sized for 50% of the CUs), products plan, 200 GiB budget, 16 GiB cache (entries up to 2 GiB), 32 GiB row
 * blocks.  Small-lane workers take arrivals in order and hand a query they find large to the
 * large-lane workers, so short queries do not queue behind long ones. */
int smcm_server_create(smcm_ctx* ctx, smcm_graph* const* graphs, int32_t n_graphs,
                       const smcm_server_opts* opts, smcm_server** out);
int smcm_server_destroy(smcm_server* s);
/* Enqueues a query; row_digest (NULL allowed) receives the result's per-row digests and must
 * stay valid until the query is polled. */
int smcm_server_submit(smcm_server* s, int32_t graph, const int32_t* vtypes, const int32_t* etypes,
                       int32_t L, const int32_t* src_rows, int64_t n_src, uint64_t* row_digest,
                       int64_t* ticket);
/* wait != 0 blocks until the query is done; otherwise SMCM_EAGAIN while it runs.  A ticket is
 * released once its result has been returned. */
int smcm_server_poll(smcm_server* s, int64_t ticket, int32_t wait, smcm_query_out* out);
/* Cancels a submitted query (no counterpart in the reference, whose program runs to completion).
 * A query no worker has started finishes at once with status SMCM_ECANCELED; a running one stops
 * at its next check (before admission, before each product, between row blocks), its device
 * memory freed, with that status; a query past its last check finishes normally.  The ticket
 * stays valid: poll it as usual.  SMCM_EINVAL for an unknown or already polled ticket. */
int smcm_server_cancel(smcm_server* s, int64_t ticket);
int smcm_server_stats(smcm_server* s, int64_t* cache_hits, int64_t* cache_inserts, uint64_t* cache_bytes);
/* Submits n queries at their arrival times (seconds from the call) and waits for all; times in
 * out[] are relative to the call.  Arrays as smcm_batch_run, plus graph[i] and arrival[i]. */
int smcm_server_replay(smcm_server* s, int32_t n, const int32_t* graph, const int32_t* path_len,
                       const int32_t* vtypes, const int32_t* etypes, const int64_t* src_off,
                       const int32_t* src_rows, const double* arrival, uint64_t* row_digest,
                       smcm_query_out* out);

/* ---- result ----------------------------------------------------------------------------- */
int32_t smcm_result_nrows(const smcm_result* r);
int64_t smcm_result_nnz(const smcm_result* r);
/* Host copies; any pointer may be NULL.  rowptr[nrows+1], cols[nnz] (global ids), vals[nnz]. */
int smcm_result_copy(const smcm_result* r, int64_t* rowptr, int32_t* cols, uint64_t* vals);
/* Rows [row_begin, row_end): rowptr gets row_end-row_begin+1 offsets relative to row_begin. */
int smcm_result_copy_rows(const smcm_result* r, int32_t row_begin, int32_t row_end,
                          int64_t* rowptr, int32_t* cols, uint64_t* vals);
/* Global ids of the result rows. */
int smcm_result_row_ids(const smcm_result* r, int32_t* rows);
/* Per-row digest, computed on the device: mix(width) + sum over the row's entries of
 * mix(mix(local column) ^ value), mod 2^64, with mix the splitmix64 finaliser. */
int smcm_result_row_digest(const smcm_result* r, uint64_t* out);
int smcm_result_free(smcm_result* r);

/* ---- raw SpGEMM ------------------------------------------------------------------------- */
/* C = A * B for host CSR operands (A m x k, B k x n; NULL values = all ones; B rows sorted).
 * The result's columns are the operand's column indices. */
int smcm_spgemm_csr(smcm_ctx* ctx, int32_t m, int32_t k, int32_t n, const int64_t* a_ptr,
                    const int32_t* a_col, const uint64_t* a_val, const int64_t* b_ptr,
                    const int32_t* b_col, const uint64_t* b_val, smcm_result** out);

#ifdef __cplusplus
}
#endif
#endif
