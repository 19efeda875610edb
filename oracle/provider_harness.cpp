// provider_harness.cpp - drives one of the reference's alternative planners, compiled from its
// source where it lies (efficient/meta.cpp, efficient/mnc.cpp or efficient/dm.cpp, chosen by
// -DHARNESS_META / -DHARNESS_MNC / -DHARNESS_DM and included below via the include path set by
// oracle/Makefile), and dumps its DP tables.  TEST INFRASTRUCTURE ONLY: run in the build
// container to produce tests/golden/ fixtures; never shipped or linked by the product.
//
// For every meta-path in the list file it writes one binary record to <out>:
//   int32 L; double m[L*L]; int32 s[L*L]; double ret
// m and s are the optimizer's m_ and s_ after the planner ran (zeros when L = 1; MNC's long
// double costs are rounded to double) and ret is the planner's return value.
#define main provider_reference_main
#if defined(HARNESS_META)
#include "efficient/meta.cpp"
#elif defined(HARNESS_MNC)
#include "efficient/mnc.cpp"
#elif defined(HARNESS_DM)
#include "efficient/dm.cpp"
#else
#error "define HARNESS_META, HARNESS_MNC or HARNESS_DM"
#endif
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
    // the matrices and the optimizer are left alive: the planners free parts of them themselves
    BoolMatrix** mats = (BoolMatrix**)malloc(L * sizeof(BoolMatrix*));
    for (int i = 0; i < L; ++i)
      mats[i] = new BoolMatrix(HIN->heteGraph, HIN->vertexTypes, mp.vertexTypes[i], mp.vertexTypes[i + 1],
                               mp.edgeTypes[i], HIN->n);
    DynamicOptimizer* dy = new DynamicOptimizer();
    double ret = 0;
    std::vector<double> m((size_t)L * L, 0.0);
    std::vector<int32_t> s((size_t)L * L, 0);
    if (L > 1) {
#if defined(HARNESS_META)
      ret = (double)dy->sparse_optimal_matrix_chain_order(L, mats, HIN->bins, HIN->dimensions, &mp.vertexTypes,
                                                          HIN->n);
#elif defined(HARNESS_MNC)
      ret = (double)dy->mnc_optimal_matrix_chain_order(L, mats, HIN->n);
#else
      ret = (double)dy->dense_map_sparse_optimal_matrix_chain_order(L, mats, HIN->bins, HIN->dimensions,
                                                                    &mp.vertexTypes, HIN->n);
#endif
      for (int q = 0; q < L * L; ++q) {
        m[q] = (double)dy->m_[q];
        s[q] = dy->s_[q];
      }
    }
    put(out, m.data(), sizeof(double) * m.size());
    put(out, s.data(), sizeof(int32_t) * s.size());
    put(out, &ret, 8);
  }
  fclose(out);
  return 0;
}
