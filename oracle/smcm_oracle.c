/*
 * smcm_oracle.c - CPU restatement of RESMCM's sparse matrix chain path (test infrastructure).
 *
 * TEST INFRASTRUCTURE ONLY: the checker, never the product.  Loaded only by the tests, the
 * smoke check and the benchmark's CPU-baseline leg.  Build: see oracle/Makefile
 * (gcc -O2 -fopenmp -ffp-contract=off).
 *
 * Pinning of each function (what checks it besides its own restatement):
 *   orc_typed_matrix  : reference build fixtures (row sizes / sorted structure per row) and
 *                       duplicate/unsorted tiny graphs.
 *   orc_estimate      : reference build fixtures at one thread (every r_ij bit-exact, every c_ij
 *                       bit-exact), closed forms (single neighbour copies density, all-zero /
 *                       all-one right densities).
 *   orc_dp / orc_plan : reference build fixtures (chain order), brute-force enumeration of all
 *                       parenthesisations with the same cost and tie rule.
 *   orc_spgemm        : reference build fixtures (structure), dense uint64 matrix products on
 *                       tiny inputs, closed forms for X * X^T (diagonal = squared row sums),
 *                       order independence, row sums = integer SpMV chain.
 */
#include "smcm_oracle.h"

#include <float.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

void orc_csr_free(orc_csr* a) {
  if (!a) return;
  free(a->ptr);
  free(a->col);
  free(a->val);
  a->ptr = NULL;
  a->col = NULL;
  a->val = NULL;
  a->nrows = a->ncols = 0;
  a->nnz = 0;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int32_t t) {
  if (t > 0) omp_set_num_threads(t);
}

/* resmcm.cpp:206-222: dims[t] counts type t; bins[t] lists its vertices ascending, so the
 * local id of v is its rank inside bins[vtype[v]]. */
void orc_local_ids(int32_t n, const int32_t* vtype, int32_t ntypes, int32_t* local, int32_t* dims) {
  for (int32_t t = 0; t < ntypes; ++t) dims[t] = 0;
  for (int32_t v = 0; v < n; ++v) local[v] = dims[vtype[v]]++;
}

/* resmcm.cpp:374-396 */
int orc_typed_matrix(int32_t n, const int32_t* vtype, const int64_t* adj_ptr,
                     const int32_t* adj_nbr, const int32_t* adj_et, int32_t ts, int32_t e,
                     int32_t td, const int32_t* rows, int64_t nrows_sel, orc_csr* out) {
  int32_t* local = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* bins_ts;
  int32_t dts = 0, dtd = 0;
  for (int32_t v = 0; v < n; ++v) {
    if (vtype[v] == ts) local[v] = dts++;
    else if (vtype[v] == td) local[v] = 0;
  }
  /* separate pass for td so that ts == td gets the same numbering */
  for (int32_t v = 0; v < n; ++v)
    if (vtype[v] == td) local[v] = dtd++;
  bins_ts = (int32_t*)malloc(sizeof(int32_t) * (size_t)(dts > 0 ? dts : 1));
  {
    int32_t k = 0;
    for (int32_t v = 0; v < n; ++v)
      if (vtype[v] == ts) bins_ts[k++] = v;
  }
  int64_t nr = rows ? nrows_sel : dts;
  for (int64_t r = 0; rows && r < nr; ++r)
    if (rows[r] < 0 || rows[r] >= dts) {
      free(local);
      free(bins_ts);
      return -1;
    }
  out->nrows = (int32_t)nr;
  out->ncols = dtd;
  out->ptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nr + 1));
  out->ptr[0] = 0;
  for (int64_t r = 0; r < nr; ++r) {
    int32_t g = bins_ts[rows ? rows[r] : r];
    int64_t c = 0;
    for (int64_t x = adj_ptr[g]; x < adj_ptr[g + 1]; ++x)
      if (adj_et[x] == e && vtype[adj_nbr[x]] == td) ++c;
    out->ptr[r + 1] = out->ptr[r] + c;
  }
  out->nnz = out->ptr[nr];
  out->col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(out->nnz > 0 ? out->nnz : 1));
  out->val = NULL;
  for (int64_t r = 0; r < nr; ++r) {
    int32_t g = bins_ts[rows ? rows[r] : r];
    int64_t k = out->ptr[r];
    for (int64_t x = adj_ptr[g]; x < adj_ptr[g + 1]; ++x)
      if (adj_et[x] == e && vtype[adj_nbr[x]] == td) out->col[k++] = local[adj_nbr[x]];
  }
  free(local);
  free(bins_ts);
  return 0;
}

/* resmcm.cpp:442-452, 468-485, 506-517 (one thread: the reduction at :473 becomes a
 * sequential sum in row order). */
int orc_estimate(int32_t L, const orc_csr* const* mats, const int32_t* coldim, double* c,
                 double* const* R) {
  double** own = (double**)calloc((size_t)L, sizeof(double*));
  double** Rv = (double**)calloc((size_t)L, sizeof(double*));
  for (int32_t i = 0; i < L * L; ++i) c[i] = 0.0;
  for (int32_t i = 0; i < L; ++i) {
    size_t sz = (size_t)mats[i]->nrows * (size_t)(L - i);
    if (R && R[i]) Rv[i] = R[i];
    else Rv[i] = own[i] = (double*)malloc(sizeof(double) * (sz ? sz : 1));
  }
  for (int32_t i = L - 1; i >= 0; --i) {
    const orc_csr* M = mats[i];
    const int32_t W = L - i;
    double* Ri = Rv[i];
    /* leaf: getSparsities, :448-449 */
    for (int32_t u = 0; u < M->nrows; ++u) {
      double s = (double)(M->ptr[u + 1] - M->ptr[u]) / (double)coldim[i];
      Ri[(size_t)u * W] = s > 1 ? 1 : s;
    }
    /* propagation, :510-514 then :474-482 */
    for (int32_t j = L - 1; j > i; --j) {
      const double* Rn = Rv[i + 1];
      const int32_t Wn = L - i - 1;
      const double dim = (double)coldim[j];
      double nnz = 0.0;
      for (int32_t u = 0; u < M->nrows; ++u) {
        double rho = 1;
        for (int64_t x = M->ptr[u]; x < M->ptr[u + 1]; ++x)
          rho *= (1 - Rn[(size_t)M->col[x] * Wn + (size_t)(j - i - 1)]);
        double s = 1 - rho;
        Ri[(size_t)u * W + (size_t)(j - i)] = s;
        nnz += s * dim;
      }
      c[i * L + j] = nnz;
    }
  }
  for (int32_t i = 0; i < L; ++i) free(own[i]);
  free(own);
  free(Rv);
  return 0;
}

/* resmcm.cpp:519-538 */
double orc_dp(int32_t L, const double* c, double* m, int32_t* s) {
  for (int32_t i = 0; i < L * L; ++i) {
    m[i] = 0.0;
    s[i] = 0;
  }
  for (int32_t l = 1; l < L; ++l) {
    for (int32_t i = 0; i < L - l; ++i) {
      int32_t j = i + l;
      m[i * L + j] = DBL_MAX;
      for (int32_t k = i; k < j; ++k) {
        double tmp = c[i * L + k] + c[(k + 1) * L + j] + c[i * L + j];
        double cost = m[i * L + k] + m[(k + 1) * L + j] + tmp;
        if (cost < m[i * L + j]) {
          m[i * L + j] = cost;
          s[i * L + j] = k;
        }
      }
    }
  }
  return m[L - 1];
}

/* resmcm.cpp:563-573 */
static int32_t plan_rec(int32_t L, const int32_t* s, int32_t i, int32_t j, int32_t* pairs, int32_t* np) {
  if (i == j) return i;
  int32_t k = plan_rec(L, s, i, s[i * L + j], pairs, np);
  int32_t l = plan_rec(L, s, s[i * L + j] + 1, j, pairs, np);
  pairs[2 * (*np)] = k;
  pairs[2 * (*np) + 1] = l;
  ++(*np);
  return -1;
}

void orc_plan(int32_t L, const int32_t* s, int32_t* pairs) {
  int32_t np = 0;
  if (L > 1) plan_rec(L, s, 0, L - 1, pairs, &np);
}

/* resmcm.cpp:586-615 (per row) and :619-647 (row loop), with counts: the bitmap becomes a
 * dense uint64 accumulator plus a marker list; the discovered columns are sorted. */
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int orc_spgemm(const orc_csr* A, const orc_csr* B, orc_csr* C) {
  if (A->ncols != B->nrows) return -1;
  const int32_t m = A->nrows, n = B->ncols;
  C->nrows = m;
  C->ncols = n;
  C->ptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m + 1));
  int64_t* cnt = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
  int nt = omp_get_max_threads();
  unsigned char* marks = (unsigned char*)calloc((size_t)nt * (size_t)(n > 0 ? n : 1), 1);
  int32_t** lists = (int32_t**)calloc((size_t)nt, sizeof(int32_t*));
  for (int t = 0; t < nt; ++t) lists[t] = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t u = 0; u < m; ++u) {
    int tid = omp_get_thread_num();
    unsigned char* mk = marks + (size_t)tid * (size_t)(n > 0 ? n : 1);
    int32_t* ls = lists[tid];
    int64_t w = 0;
    for (int64_t x = A->ptr[u]; x < A->ptr[u + 1]; ++x) {
      int32_t k = A->col[x];
      for (int64_t y = B->ptr[k]; y < B->ptr[k + 1]; ++y) {
        int32_t v = B->col[y];
        if (!mk[v]) {
          mk[v] = 1;
          ls[w++] = v;
        }
      }
    }
    for (int64_t q = 0; q < w; ++q) mk[ls[q]] = 0;
    cnt[u] = w;
  }
  C->ptr[0] = 0;
  for (int32_t u = 0; u < m; ++u) C->ptr[u + 1] = C->ptr[u] + cnt[u];
  C->nnz = C->ptr[m];
  C->col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(C->nnz > 0 ? C->nnz : 1));
  C->val = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(C->nnz > 0 ? C->nnz : 1));
  uint64_t** accs = (uint64_t**)calloc((size_t)nt, sizeof(uint64_t*));
  for (int t = 0; t < nt; ++t) accs[t] = (uint64_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t u = 0; u < m; ++u) {
    int tid = omp_get_thread_num();
    unsigned char* mk = marks + (size_t)tid * (size_t)(n > 0 ? n : 1);
    uint64_t* acc = accs[tid];
    int32_t* out_c = C->col + C->ptr[u];
    uint64_t* out_v = C->val + C->ptr[u];
    int64_t w = 0;
    for (int64_t x = A->ptr[u]; x < A->ptr[u + 1]; ++x) {
      int32_t k = A->col[x];
      uint64_t a = A->val ? A->val[x] : 1u;
      for (int64_t y = B->ptr[k]; y < B->ptr[k + 1]; ++y) {
        int32_t v = B->col[y];
        uint64_t b = B->val ? B->val[y] : 1u;
        acc[v] += a * b;
        if (!mk[v]) {
          mk[v] = 1;
          out_c[w++] = v;
        }
      }
    }
    qsort(out_c, (size_t)w, sizeof(int32_t), cmp_i32);
    for (int64_t q = 0; q < w; ++q) {
      out_v[q] = acc[out_c[q]];
      acc[out_c[q]] = 0;
      mk[out_c[q]] = 0;
    }
  }
  for (int t = 0; t < nt; ++t) {
    free(lists[t]);
    free(accs[t]);
  }
  free(lists);
  free(accs);
  free(marks);
  free(cnt);
  return 0;
}

static void csr_copy(const orc_csr* a, orc_csr* out) {
  out->nrows = a->nrows;
  out->ncols = a->ncols;
  out->nnz = a->nnz;
  out->ptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(a->nrows + 1));
  memcpy(out->ptr, a->ptr, sizeof(int64_t) * (size_t)(a->nrows + 1));
  out->col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(a->nnz > 0 ? a->nnz : 1));
  if (a->nnz) memcpy(out->col, a->col, sizeof(int32_t) * (size_t)a->nnz);
  out->val = NULL;
  if (a->val) {
    out->val = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(a->nnz > 0 ? a->nnz : 1));
    if (a->nnz) memcpy(out->val, a->val, sizeof(uint64_t) * (size_t)a->nnz);
  }
}

/* resmcm.cpp:709-760: operands either come from the chain (index >= 0) or are the last
 * (right) / second-to-last (left) entries of the intermediate stack (index -1). */
int orc_execute_plan(int32_t L, const orc_csr* const* mats, const int32_t* pairs, orc_csr* out) {
  if (L == 1) {
    csr_copy(mats[0], out);
    return 0;
  }
  orc_csr* stack = (orc_csr*)calloc((size_t)L, sizeof(orc_csr));
  int32_t top = 0;
  for (int32_t p = 0; p < L - 1; ++p) {
    int32_t k = pairs[2 * p], l = pairs[2 * p + 1];
    orc_csr res;
    memset(&res, 0, sizeof(res));
    int rc;
    if (k >= 0 && l >= 0) {
      rc = orc_spgemm(mats[k], mats[l], &res);
      if (rc) return rc;
      stack[top++] = res;
    } else if (k < 0 && l >= 0) {
      rc = orc_spgemm(&stack[top - 1], mats[l], &res);
      if (rc) return rc;
      orc_csr_free(&stack[top - 1]);
      stack[top - 1] = res;
    } else if (k >= 0 && l < 0) {
      rc = orc_spgemm(mats[k], &stack[top - 1], &res);
      if (rc) return rc;
      orc_csr_free(&stack[top - 1]);
      stack[top - 1] = res;
    } else {
      rc = orc_spgemm(&stack[top - 2], &stack[top - 1], &res);
      if (rc) return rc;
      orc_csr_free(&stack[top - 1]);
      orc_csr_free(&stack[top - 2]);
      stack[top - 2] = res;
      --top;
    }
  }
  *out = stack[0];
  free(stack);
  return 0;
}

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* digest = mix(width) + sum over entries of mix(mix(col) ^ val), mod 2^64 (rows are sorted, so
 * the order-free sum still identifies the row) */
void orc_row_digest(const orc_csr* C, uint64_t* out) {
#pragma omp parallel for schedule(static)
  for (int32_t u = 0; u < C->nrows; ++u) {
    uint64_t h = 0;
    for (int64_t x = C->ptr[u]; x < C->ptr[u + 1]; ++x)
      h += mix64(mix64((uint64_t)(uint32_t)C->col[x] ^ 0x9E3779B97F4A7C15ull) ^ (C->val ? C->val[x] : 1u));
    out[u] = h + mix64((uint64_t)(C->ptr[u + 1] - C->ptr[u]) ^ 0x243F6A8885A308D3ull);
  }
}

/* Rows of A (a row range of a CSR, ptr may start above 0) times B: per-row digests of the
 * product as orc_row_digest computes them, without materialising it.  The digest is a sum over
 * the row's entries, so the columns need no sorting.  Returns the number of entries; products
 * receives sum over A's entries of |row_k(B)|. */
static int64_t spgemm_digest_rows(const orc_csr* A, const orc_csr* B, uint64_t* out, int64_t* products) {
  const int32_t m = A->nrows, n = B->ncols;
  int nt = omp_get_max_threads();
  unsigned char* marks = (unsigned char*)calloc((size_t)nt * (size_t)(n > 0 ? n : 1), 1);
  int32_t** lists = (int32_t**)calloc((size_t)nt, sizeof(int32_t*));
  uint64_t** accs = (uint64_t**)calloc((size_t)nt, sizeof(uint64_t*));
  for (int t = 0; t < nt; ++t) {
    lists[t] = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    accs[t] = (uint64_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(uint64_t));
  }
  int64_t nnz = 0, prods = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : nnz, prods)
  for (int32_t u = 0; u < m; ++u) {
    int tid = omp_get_thread_num();
    unsigned char* mk = marks + (size_t)tid * (size_t)(n > 0 ? n : 1);
    int32_t* ls = lists[tid];
    uint64_t* acc = accs[tid];
    int64_t w = 0;
    for (int64_t x = A->ptr[u]; x < A->ptr[u + 1]; ++x) {
      int32_t k = A->col[x];
      uint64_t a = A->val ? A->val[x] : 1u;
      prods += B->ptr[k + 1] - B->ptr[k];
      for (int64_t y = B->ptr[k]; y < B->ptr[k + 1]; ++y) {
        int32_t v = B->col[y];
        acc[v] += a * (B->val ? B->val[y] : 1u);
        if (!mk[v]) {
          mk[v] = 1;
          ls[w++] = v;
        }
      }
    }
    uint64_t h = 0;
    for (int64_t q = 0; q < w; ++q) {
      int32_t v = ls[q];
      h += mix64(mix64((uint64_t)(uint32_t)v ^ 0x9E3779B97F4A7C15ull) ^ acc[v]);
      acc[v] = 0;
      mk[v] = 0;
    }
    out[u] = h + mix64((uint64_t)w ^ 0x243F6A8885A308D3ull);
    nnz += w;
  }
  for (int t = 0; t < nt; ++t) {
    free(lists[t]);
    free(accs[t]);
  }
  free(lists);
  free(accs);
  free(marks);
  *products += prods;
  return nnz;
}

/* products of the rows of A (ptr may start above 0) against B */
static int64_t products_rows(const orc_csr* A, const orc_csr* B) {
  int64_t ip = 0;
  for (int32_t u = 0; u < A->nrows; ++u)
    for (int64_t x = A->ptr[u]; x < A->ptr[u + 1]; ++x) ip += B->ptr[A->col[x] + 1] - B->ptr[A->col[x]];
  return ip;
}

int orc_plan_digest(int32_t L, const orc_csr* const* mats, const int32_t* pairs, int32_t block_rows,
                    uint64_t* digest, int64_t* nnz_out, int64_t* products_out) {
  if (L < 2 || block_rows < 1) return -1;
  /* phase 1: the plan's products that do not involve position 0, once (the stack machine of
   * resmcm.cpp:709-760 over intervals; a step's result covers [lo, hi]) */
  orc_csr* pre = (orc_csr*)calloc((size_t)L, sizeof(orc_csr));
  int32_t* lo = (int32_t*)malloc(sizeof(int32_t) * (size_t)L);
  int32_t* src = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)L); /* operand: leaf x >= 0, step -(p+1) */
  int32_t* st_step = (int32_t*)malloc(sizeof(int32_t) * (size_t)L);  /* stack of step ids */
  int32_t top = 0;
  int64_t products = 0;
  int rc = 0;
  for (int32_t p = 0; p < L - 1; ++p) {
    int32_t k = pairs[2 * p], l = pairs[2 * p + 1];
    int32_t a, b;
    if (k >= 0 && l >= 0) {
      a = k;
      b = l;
    } else if (k < 0 && l >= 0) {
      a = -(st_step[--top] + 1);
      b = l;
    } else if (k >= 0 && l < 0) {
      a = k;
      b = -(st_step[--top] + 1);
    } else {
      b = -(st_step[--top] + 1);
      a = -(st_step[--top] + 1);
    }
    src[2 * p] = a;
    src[2 * p + 1] = b;
    lo[p] = a >= 0 ? a : lo[-a - 1];
    st_step[top++] = p;
    if (lo[p] > 0) {
      const orc_csr* A = a >= 0 ? mats[a] : &pre[-a - 1];
      const orc_csr* B = b >= 0 ? mats[b] : &pre[-b - 1];
      products += products_rows(A, B);
      if ((rc = orc_spgemm(A, B, &pre[p])) != 0) break;
    }
  }
  /* phase 2: row blocks of M_0 through the steps that involve position 0; the root (the last
   * step) is reduced to row digests */
  int64_t nnz = 0;
  orc_csr* blk = (orc_csr*)calloc((size_t)L, sizeof(orc_csr));
  const orc_csr* M0 = mats[0];
  for (int32_t r0 = 0; rc == 0 && r0 < M0->nrows; r0 += block_rows) {
    const int32_t r1 = M0->nrows - r0 < block_rows ? M0->nrows : r0 + block_rows;
    orc_csr view = *M0;
    view.nrows = r1 - r0;
    view.ptr = M0->ptr + r0;
    for (int32_t p = 0; p < L - 1 && rc == 0; ++p) {
      if (lo[p] > 0) continue;
      const int32_t a = src[2 * p], b = src[2 * p + 1];
      const orc_csr* A = a == 0 ? &view : (a > 0 ? mats[a] : &blk[-a - 1]);
      const orc_csr* B = b >= 0 ? mats[b] : (lo[-b - 1] > 0 ? &pre[-b - 1] : &blk[-b - 1]);
      if (p == L - 2) {
        nnz += spgemm_digest_rows(A, B, digest + r0, &products);
      } else {
        products += products_rows(A, B);
        rc = orc_spgemm(A, B, &blk[p]);
      }
      if (a < 0) orc_csr_free(&blk[-a - 1]);
    }
  }
  for (int32_t p = 0; p < L - 1; ++p) {
    if (pre[p].ptr) orc_csr_free(&pre[p]);
    if (blk[p].ptr) orc_csr_free(&blk[p]);
  }
  free(pre);
  free(blk);
  free(lo);
  free(src);
  free(st_step);
  if (nnz_out) *nnz_out = nnz;
  if (products_out) *products_out = products;
  return rc;
}

int64_t orc_products(const orc_csr* A, const orc_csr* B) {
  int64_t ip = 0;
  for (int64_t x = 0; x < A->nnz; ++x) ip += B->ptr[A->col[x] + 1] - B->ptr[A->col[x]];
  return ip;
}
