// kernels.h - host launchers of the gfx950 kernels (all enqueue on DevAlloc::stream).
#pragma once

#include <functional>
#include <vector>

#include "alloc.h"
#include "common.h"
#include "ktimer.h"

namespace smcm {

// ---- scan.hip ---------------------------------------------------------------------------
// out[0] = 0, out[i+1] = out[i] + in[i] for i < n   (out has n+1 entries; in != out)
void scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, DevAlloc& al);

// ---- device_graph.hip -------------------------------------------------------------------
// local[v] = rank of v among vertices of its type (ascending id); bins[t][local] = v.
// bins[t] must hold dims[t] entries.
void build_type_index(const int32_t* vtype, int32_t n, int32_t ntypes, int32_t* local,
                      int32_t* const* bins, DevAlloc& al);

// Typed relation matrix: rows = bins_ts[rows_sel[r]] (rows_sel may be null = all rows 0..nrows-1),
// entries = neighbours with entry type et and vertex type td, in adjacency order, duplicates
// kept; columns are local ids.  Sets out.sorted.
void build_typed_csr(const int32_t* vtype, const int64_t* adj_ptr, const int32_t* adj_nbr,
                     const int32_t* adj_et, const int32_t* local, const int32_t* bins_ts,
                     const int32_t* rows_sel, int32_t nrows, int32_t et, int32_t td, int32_t ncols,
                     Csr& out, DevAlloc& al);

// out = rows rows_sel[0..nsel) of in (values copied when present)
void csr_gather_rows(const Csr& in, const int32_t* rows_sel, int32_t nsel, Csr& out, DevAlloc& al);

// bytes/16 16-byte copies (plain HBM copy: bandwidth reference and counter calibration)
void device_copy(const void* src, void* dst, int64_t bytes, hipStream_t st);

// dst[i] = map[src[i]]
void map_columns(const int32_t* src, int64_t n, const int32_t* map, int32_t* dst, hipStream_t st);

// per-row digest of the (column, value) sequence; same mixing as the oracle's orc_row_digest
void row_digest(const Csr& c, uint64_t* out, hipStream_t st);

// fill helpers (kernels, graph-capture safe)
void fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t st);
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);

// ---- estimator.hip ----------------------------------------------------------------------
struct EstimatorPass {
  const Csr* M = nullptr;        // M_i (rows of type t_i, file order)
  double leaf_dim = 1.0;         // dims[t_{i+1}]
  const double* Rn = nullptr;    // R_{i+1}: rows_{i+1} x W  (null when W == 0)
  int W = 0;                     // L-1-i
  const double* dimj = nullptr;  // host: dims[t_{j+1}] for j = i+1..L-1  (W values)
  double* Ri = nullptr;          // out: nrows x (W+1); Ri[u][0] = r_ii, Ri[u][1+q] = r_{i,i+1+q}
  double* c_dev = nullptr;       // out: W values c_{i,i+1+q}
  bool exact_sum = false;        // sequential sum in row order (bit-exact with one thread)
  KernelTimer* tm = nullptr;
};
void estimator_pass(const EstimatorPass& p, DevAlloc& al);

// Estimated product counts of every split: P[(i*L + k)*L + j] for i <= k < j (device table,
// L*L*L doubles, only those entries written) for X_ik * X_{k+1,j}.  Needs the estimator's
// row densities of every position (R[q] as written by estimator_pass, q = 1..L-1).
struct ProductEstimate {
  const Csr* mats = nullptr;        // the chain's L matrices (M_0 already restricted to the sources)
  int L = 0;
  const double* const* R = nullptr; // host array of L device pointers (R[0] unused)
  const double* dims = nullptr;     // host: dims[t_q] for q = 0..L
  double* p_dev = nullptr;
  KernelTimer* tm = nullptr;
};
void product_estimate(const ProductEstimate& pe, DevAlloc& al);

// ---- batch.hip --------------------------------------------------------------------------
// One product of a batch: A (a base CSR, optionally through a row map, or node a_node's output)
// times B (a base CSR or node b_node's output).  Node outputs are rows (start, length) in the
// batch's output arena.
struct BatchNode {
  const int64_t* a_ptr = nullptr;
  const int32_t* a_rows = nullptr;
  const int32_t* a_col = nullptr;
  const uint64_t* a_val = nullptr;
  const int64_t* b_ptr = nullptr;
  const int32_t* b_col = nullptr;
  const uint64_t* b_val = nullptr;
  int64_t* out_start = nullptr;
  int32_t* out_len = nullptr;
  int32_t a_node = -1, b_node = -1;
  int32_t nrows = 0, ncols = 0;
  int32_t ntiles = 0;
  int32_t ninputs = 0;  // operands that are node outputs
};
struct BatchState {
  int32_t head = 0, tail = 0;  // next queue slot to claim / to fill
  int32_t error = 0, pad = 0;  // bit 0: output arena exhausted; bit 1: a queue wait gave up
  unsigned long long arena_used = 0;
};
struct BatchLaunch {
  const BatchNode* nodes = nullptr;
  const int32_t* parent_off = nullptr;  // consumers of node n: parents[parent_off[n] .. parent_off[n+1])
  const int32_t* parents = nullptr;
  int32_t* done_tiles = nullptr;        // per node, zeroed
  int32_t* ready = nullptr;             // per node: completed node operands, zeroed
  int32_t* failed = nullptr;            // per node, zeroed
  unsigned long long* queue = nullptr;  // total_tiles slots: 0 = not yet filled, else ((node << 32) | tile) + 1
  BatchState* state = nullptr;
  int32_t* arena_col = nullptr;
  uint64_t* arena_val = nullptr;
  uint64_t arena_cap = 0;               // entries
  int64_t spin_limit = 0;               // wall-clock ticks (100 MHz) a queue wait may last
  int32_t total_tiles = 0;
  int32_t tile_rows = 32;
  int threads = 512, sort_cap = 1024, log2w = 13;
  int grid = 0;
};
size_t batch_lds_bytes(int threads, int sort_cap, int log2w);
bool batch_shape_supported(int threads, int sort_cap);
void batch_execute(const BatchLaunch& bl, hipStream_t st);
// digests (as row_digest) of rows given as (start, length) over col/val
void padded_digest(const int64_t* start, const int32_t* len, const int32_t* col, const uint64_t* val,
                   const int32_t* colmap, int32_t nrows, uint64_t* out, hipStream_t st);
// rows given as (start, length) -> a compact CSR (allocated from al)
void padded_to_csr(const int64_t* start, const int32_t* len, const int32_t* col, const uint64_t* val, int32_t nrows,
                   int32_t ncols, Csr& out, DevAlloc& al);

// ---- spgemm.hip -------------------------------------------------------------------------
// byte offset of the dense tier's readback in the thread's host stage (the row-binned path's
// sections lie below it)
constexpr size_t kHostStageDense = 576;

struct SpgemmOpts {
  int s_max = 32;          // rows with ub <= s_max: one wave per row, register sort
  int m_max = 1024;        // rows with ub <= m_max (and sparse estimate): workgroup LDS sort
  int dense_div = 64;      // estimated width >= ncols/dense_div sends the row to the bitmap tier
  int64_t heavy_ub = 1 << 16;  // bitmap-tier rows with ub above this are scheduled first
  int d_threads = 512;     // bitmap-tier workgroup size
  int w_threads = 512;     // window-tier workgroup size
  int64_t window_cols = 8192;  // window width cap (0 = largest that fits in LDS); 8192 columns
                               // (67 KB of LDS) fit two 512-thread workgroups per CU
  int64_t bitmap_vcap = -1; // cap of the bitmap tier's LDS value slots (-1 = as many as fit);
                            // rows wider than this use the window kernel
  int64_t lds_bytes = 65536;   // LDS available per workgroup
  int cus = 256;           // compute units (persistent grid size)
  int device_cus = 0;      // compute units of the device; cus below it reserves the rest for other
                           // streams, so the persistent window kernels then place one workgroup per CU
  int force_tier = -1;     // testing: 0 = wave, 1 = LDS sort, 2 = bitmap for every non-empty row
  int64_t wave_bitmap_cols = 32768;    // rows of products with at most this many columns and
  int64_t wave_bitmap_max_ub = 2048;   // 64 < ub <= wave_bitmap_max_ub: one wave per row over an LDS
                                       // column bitmap (cols 0 = never)
  int64_t wave_bitmap_cap = 2048;      // LDS value slots per wave; wider rows accumulate in HBM
  int window_major = 1;     // window tier in window-major task order (0 = one row at a time)
  int copy_tier = 1;        // rows of one entry over a right operand without repeated columns are
                            // copied (scaled) from B when they have more than s_max products
  int window_presence = 2;  // window-major tasks track occupied columns with presence bits from the
                            // start (1), take the nonzero values as the columns (0), or choose by
                            // the expected window fill (2: presence below 1/8)
  int64_t sym_window_cols = 262144;  // heavy rows' symbolic pass: columns per window-major task
  int sym_merge = 0;        // ... merges lanes sharing a bitmap word: 0 never (default: its bitmap
                            // layout spreads a dense run of columns over the banks instead), 1
                            // always, 2 when the heavy rows average at least 8 products per column
  int64_t window_min_products = 128;   // rows wider than the bitmap tier's LDS value slots use the
                                       // window kernel when they average at least this many
                                       // products per (row, window); otherwise the bitmap kernel
                                       // accumulates their values in HBM (device atomics)
  int64_t window_all_nwin = 4;  // products spanning 2..this many windows send every bitmap-tier row
                                // to the window kernel (LDS value slots are bypassed)
  int64_t dense_cells = (int64_t)1 << 23;  // products with rows(A) x cols(B) up to this many cells
                                           // accumulate whole in a dense device array (0 = never)
  int64_t dense_row_products = 10000;  // ... unless B has more columns than a workgroup combines in
                                       // LDS and the rows average at least this many products
                                       // (hub rows: device atomics on their columns serialise;
                                       // 0 = no such limit)
  int window_u32 = 1;       // window accumulators of 32 bits (twice the columns per window) when the
                            // operands' largest values bound every count below 2^32
  int64_t window_mixed_min = (int64_t)1 << 26;  // products with at least this many products, whose
                                                // whole bound does not fit 32 bits, bound each row:
                                                // rows that fit take 32 bits, the rest 64 (same
                                                // windows, a second launch)
  int64_t bbits_bytes = (int64_t)64 << 20;  // right operands whose column bitmap (rows x words) fits
                                            // this take the symbolic pass as ORs of bitmap rows,
                                            // when that reads fewer words than 4x the products (0 = never)
};

struct SpgemmStats {
  int64_t products = 0;     // intermediate products (sum over A entries of |B row|)
  int64_t nnz = 0;
  int64_t rows[4] = {0, 0, 0, 0};  // empty / wave / sort / bitmap
  int64_t windows = 1;
  // host milliseconds per phase; ms_numeric covers the numeric kernels only when spgemm()
  // synchronised (timing, a kernel timer or SMCM_TRACE), otherwise just their launches
  double ms_analysis = 0, ms_symbolic = 0, ms_numeric = 0;
};

// C = A * B (uint64 counts, columns sorted).  est_r (nullable, device) gives the estimator's
// row densities r[u * est_stride]; the predicted width is r * ncols(B).  B rows must be sorted.
// C is complete in stream order on al.stream; the host waits for it only when timing is set.
void spgemm(const Csr& A, const Csr& B, const double* est_r, int est_stride, Csr& C,
            const SpgemmOpts& o, SpgemmStats* st, DevAlloc& al, bool timing = false,
            KernelTimer* tm = nullptr);

// Dense tier (dense.hip): C = A * B accumulated whole in a dense all-zero device array when
// rows(A) x cols(B) <= o.dense_cells (and no tier is forced; products over wide B whose rows
// average o.dense_row_products or more stay with the row tiers).  Returns false, having done
// nothing, when the product does not qualify; C.ptr must be allocated (nrows + 1).
bool spgemm_dense(const Csr& A, const Csr& B, Csr& C, const SpgemmOpts& o, SpgemmStats* st, DevAlloc& al,
                  KernelTimer* tm);

// Row-blocked C = A * B for outputs larger than memory: the symbolic pass runs once, then
// rows are cut into blocks of at most block_bytes of output (a single row is always a block);
// each block is materialised in HBM and handed to consume(block, row_begin, row_end) before
// its memory is reused.  Block CSRs use block-relative row pointers.
using BlockConsumer = std::function<void(const Csr&, int32_t, int32_t)>;
void spgemm_blocked(const Csr& A, const Csr& B, const double* est_r, int est_stride,
                    int64_t block_bytes, const BlockConsumer& consume, const SpgemmOpts& o,
                    SpgemmStats* st, DevAlloc& al, KernelTimer* tm = nullptr);

}  // namespace smcm
