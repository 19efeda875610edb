// ref_harness.cpp - drives the reference's own resmcm implementation, compiled from its source
// where it lies (efficient/resmcm.cpp, included below via the include path set by
// oracle/Makefile), and dumps what the golden fixtures need.  TEST INFRASTRUCTURE ONLY: run in
// the build container to produce tests/golden/ fixtures; never shipped or linked by the product.
//
// For every meta-path in the list file it writes one binary record to <out>:
//   int32 L; int32 vts[L+1]; int32 ets[L];
//   double c[L*L];  int32 s[L*L];  int32 npairs; int32 pairs[2*npairs]   (chain order)
//   for i in 0..L-1, for j in i..L-1:  int64 nrows; double r_ij[nrows]     (rows of type t_i,
//                                                                          ascending global id)
//   int64 nrows; for each row: int64 len; int32 cols[len] (global ids, sorted ascending)
// The estimator tables come from DynamicOptimizer::rows_sparse_optimal_matrix_chain_order
// (resmcm.cpp:488-541) and get_optimal_chain_order (:563-573); the result rows from
// SSP_dynamic (:653-772), sorted here because the reference emits discovery order (:586-615).
#define main resmcm_reference_main
#include "efficient/resmcm.cpp"
#undef main

#include <cstdint>
#include <cstdio>

static void put(FILE* f, const void* p, size_t n) { fwrite(p, 1, n, f); }

int main(int argc, char** argv) {
  if (argc != 7) {
    fprintf(stderr, "usage: %s THREADS graph.txt vertex.txt edge.txt metapaths.txt out.bin\n", argv[0]);
    return 2;
  }
  omp_set_num_threads(atoi(argv[1]));
  HeteroGraph* HIN = new HeteroGraph(argv[2], argv[3], argv[4]);
  vector<MetaPath> paths = read_metapathesN(argv[5]);
  FILE* out = fopen(argv[6], "wb");
  if (!out) return 3;
  int tnum = omp_get_max_threads();
  t_rows = (long*)calloc(tnum, sizeof(long));
  t_times = (double*)calloc(tnum, sizeof(double));
  for (MetaPath& mp : paths) {
    int32_t L = mp.pathLen;
    put(out, &L, 4);
    for (int i = 0; i <= L; ++i) { int32_t v = mp.vertexTypes[i]; put(out, &v, 4); }
    for (int i = 0; i < L; ++i) { int32_t v = mp.edgeTypes[i]; put(out, &v, 4); }
    // estimator + DP, same calls as SSP_dynamic :664-687
    BoolMatrix** mats = (BoolMatrix**)malloc(L * sizeof(BoolMatrix*));
    for (int i = 0; i < L; ++i)
      mats[i] = new BoolMatrix(HIN->heteGraph, HIN->vertexTypes, mp.vertexTypes[i], mp.vertexTypes[i + 1],
                               mp.edgeTypes[i], HIN->n);
    DynamicOptimizer dy;
    vector<pair<int, int>> order;
    if (L > 1) {
      dy.rows_sparse_optimal_matrix_chain_order(L, mats, HIN->bins, HIN->dimensions, &mp.vertexTypes, HIN->n);
      dy.get_optimal_chain_order(0, L - 1, &order);
      put(out, dy.c_, sizeof(double) * L * L);
      for (int q = 0; q < L * L; ++q) { int32_t v = dy.s_[q]; put(out, &v, 4); }
    } else {
      vector<double> z(L * L, 0.0);
      put(out, z.data(), sizeof(double) * L * L);
      for (int q = 0; q < L * L; ++q) { int32_t v = 0; put(out, &v, 4); }
    }
    int32_t np = (int32_t)order.size();
    put(out, &np, 4);
    for (auto& p : order) { int32_t a = p.first, b = p.second; put(out, &a, 4); put(out, &b, 4); }
    for (int i = 0; i < L; ++i) {
      vector<int>* bin = HIN->bins[mp.vertexTypes[i]];
      for (int j = i; j < L; ++j) {
        int64_t nr = (int64_t)bin->size();
        put(out, &nr, 8);
        vector<double>* r = (L > 1) ? dy.r_[i * L + j] : nullptr;
        for (int u : *bin) { double x = r ? r->at(u) : 0.0; put(out, &x, 8); }
      }
    }
    if (L > 1) dy.free_rows_sparse_optimal_matrix_chain_order();
    for (int i = 0; i < L; ++i) delete mats[i];
    free(mats);
    // the product itself
    BoolMatrix* graph = nullptr;
    build_pair(HIN->heteGraph, graph, HIN->vertexTypes, HIN->edgeTypes, HIN->bins, HIN->dimensions, &mp,
               HIN->n, HIN->m);
    vector<int>* bin0 = HIN->bins[mp.vertexTypes[0]];
    int64_t nr = (int64_t)bin0->size();
    put(out, &nr, 8);
    for (int u : *bin0) {
      vector<int> row = *graph->rows[u];
      std::sort(row.begin(), row.end());
      int64_t len = (int64_t)row.size();
      put(out, &len, 8);
      put(out, row.data(), sizeof(int) * row.size());
    }
    graph->free_memory();
    delete graph;
  }
  fclose(out);
  return 0;
}
