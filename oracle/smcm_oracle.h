/*
 * smcm_oracle.h - CPU restatement of RESMCM's sparse matrix chain path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP implementation: only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark may load it.  The
 * product path (libsmcm.so) never links, loads or falls back to it.
 *
 * Every function restates the reference algorithm it names (efficient/resmcm.cpp unless noted),
 * extended with uint64 path counts (the reference product is Boolean).  Arithmetic:
 *   - estimator: IEEE binary64, evaluated in the reference's order, built -ffp-contract=off;
 *   - SpGEMM values: uint64, mod 2^64; structure = Boolean product (an entry whose count wraps
 *     to 0 mod 2^64 stays a structural entry).
 */
#ifndef SMCM_ORACLE_H
#define SMCM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_csr {
  int32_t nrows, ncols;
  int64_t nnz;
  int64_t* ptr;   /* [nrows+1] */
  int32_t* col;   /* [nnz] */
  uint64_t* val;  /* [nnz] or NULL (every entry 1) */
} orc_csr;

void orc_csr_free(orc_csr* a);

/* Local (per-type) ids in ascending global order; dims[t] = number of vertices of type t.
 * HeteroGraph::set_dimensions / set_bins, resmcm.cpp:206-222. */
void orc_local_ids(int32_t n, const int32_t* vtype, int32_t ntypes, int32_t* local, int32_t* dims);

/* Typed adjacency matrix of meta-path edge (ts --e--> td): rows are the vertices of type ts
 * (or the given local rows, in the given order, when rows != NULL), columns the local ids of
 * the neighbours u with entry type e and vtype[u] == td, in file order, duplicates kept.
 * BoolMatrix(HIN, ...) constructor, resmcm.cpp:374-396. */
int orc_typed_matrix(int32_t n, const int32_t* vtype, const int64_t* adj_ptr,
                     const int32_t* adj_nbr, const int32_t* adj_et, int32_t ts, int32_t e,
                     int32_t td, const int32_t* rows, int64_t nrows_sel, orc_csr* out);

/* Row-wise sparsity estimator over a chain of L matrices mats[0..L-1] (mats[i] has
 * mats[i+1]->nrows == mats[i]->ncols).  coldim[i] = dims of the column type of mats[i].
 *   r_ii[u] = min(1, |row_u(M_i)| / coldim[i])                 getSparsities, :442-452
 *   r_ij[u] = 1 - prod_{v in row_u(M_i)} (1 - r_{i+1,j}[v])    compute_sparsities, :468-485
 *   c_ij    = sum_u r_ij[u] * coldim[j] (sequential, row order) :481 at one thread
 * schedule i = L-1..0, j = L-1..i+1                           :508-515
 * c (L*L, row-major) gets c_ij for j > i and 0 elsewhere.  When R != NULL, R[i] must hold
 * mats[i]->nrows * (L-i) doubles and receives r_ij[u] at R[i][u*(L-i) + (j-i)]. */
int orc_estimate(int32_t L, const orc_csr* const* mats, const int32_t* coldim, double* c,
                 double* const* R);

/* Chain-order DP (resmcm.cpp:519-538): m_ij = min_k m_ik + m_{k+1,j} + (c_ik + c_{k+1,j} + c_ij),
 * lowest k on ties.  m and s are L*L row-major. Returns m[0][L-1]. */
double orc_dp(int32_t L, const double* c, double* m, int32_t* s);

/* Post-order plan (get_optimal_chain_order, resmcm.cpp:563-573): pairs[2p], pairs[2p+1] is the
 * p-th product's (left, right); a leaf is its matrix index, an earlier product is -1.
 * Writes L-1 pairs. */
void orc_plan(int32_t L, const int32_t* s, int32_t* pairs);

/* C = A * B with uint64 counts (Boolean row-wise Gustavson, resmcm.cpp:586-647, extended with
 * values; columns sorted ascending).  Uses OpenMP over rows. */
int orc_spgemm(const orc_csr* A, const orc_csr* B, orc_csr* C);

/* Execute a plan over mats (stack machine of SSP_dynamic, resmcm.cpp:709-760).  The consumed
 * base matrices are not freed; the result is written to out. */
int orc_execute_plan(int32_t L, const orc_csr* const* mats, const int32_t* pairs, orc_csr* out);

/* The result of a plan (L >= 2) reduced to per-row digests (as orc_row_digest) without
 * materialising it: the plan's products that do not involve position 0 are computed once,
 * those that do run on row blocks of M_0 of block_rows rows, and the root is digested row by
 * row.  digest[mats[0]->nrows]; *nnz_out = entries of the result; *products_out = the plan's
 * intermediate products over all rows.  Returns 0, or -1 for L < 2 or block_rows < 1. */
int orc_plan_digest(int32_t L, const orc_csr* const* mats, const int32_t* pairs, int32_t block_rows,
                    uint64_t* digest, int64_t* nnz_out, int64_t* products_out);

/* Per-row digest: mix(width) + sum over the row's entries of mix(mix(col) ^ val), mod 2^64
 * (mix = the splitmix64 finaliser; rows are sorted, so the order-free sum identifies a row). */
void orc_row_digest(const orc_csr* C, uint64_t* out);

/* Number of intermediate products of C = A*B: sum over A's entries of |row_k(B)|. */
int64_t orc_products(const orc_csr* A, const orc_csr* B);

int orc_num_threads(void);
void orc_set_threads(int32_t t);

#ifdef __cplusplus
}
#endif
#endif
