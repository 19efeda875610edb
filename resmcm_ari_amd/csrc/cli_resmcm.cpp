// smcm_resmcm - drop-in replacement of the reference's efficient/resmcm program
// (main, efficient/resmcm.cpp:782-895): same positional arguments
//   THREADS ITERATIONS graph.txt vertex.txt edge.txt metapaths.txt output.txt
// and the same appended output line per meta-path (:883):
//   "<meta-path> <nnz> <total> <spgemm> <dp> <est>"
// with times averaged over the iterations.  THREADS is accepted for compatibility (the
// products run on the GPU).  An optional 8th argument selects the device.  The environment
// variable SMCM_PLAN_COST (meta, mnc, dm, lg or 0-5) selects the DP's cost provider, which makes
// the program a drop-in for the reference's efficient/{meta,mnc,dm,lg} programs as well (same
// arguments and output line).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cli_common.h"
#include "smcm.h"
#include "smcm_hin.h"

int main(int argc, char** argv) {
  if (argc != 8 && argc != 9) {
    std::cerr << "usage: " << argv[0]
              << " THREADS ITERATIONS graph.txt vertex.txt edge.txt metapaths.txt output.txt [device]" << std::endl;
    return 2;
  }
  const int threads = std::atoi(argv[1]);
  const double iterations = std::atof(argv[2]);
  const int device = argc == 9 ? std::atoi(argv[8]) : 0;
  smcm_cli::banner(threads, iterations, argv);

  hin_graph* hg = nullptr;
  if (hin_read_text(argv[3], argv[4], argv[5], &hg) != 0) {
    std::cerr << hin_last_error() << std::endl;
    return 1;
  }
  std::cout << "*\tthe number of vertex is: " << hg->n << "." << std::endl;
  std::cout << "*\tthe number of edge is: " << hg->m << "." << std::endl;

  smcm_ctx* ctx = nullptr;
  smcm_graph* g = nullptr;
  if (smcm_ctx_create(device, 0, &ctx) != 0 ||
      smcm_graph_upload(ctx, hg->n, hg->vtype, hg->adj_ptr, hg->adj_nbr, hg->adj_et, &g) != 0) {
    std::cerr << smcm_last_error() << std::endl;
    return 1;
  }

  if (const char* pc = std::getenv("SMCM_PLAN_COST")) {
    const std::string v = pc;
    const int64_t code = v == "meta" ? 2 : v == "mnc" ? 3 : v == "dm" ? 4 : v == "lg" ? 5 : std::atoll(pc);
    if (smcm_set_option(ctx, "plan_cost", code) != 0) {
      std::cerr << smcm_last_error() << std::endl;
      return 1;
    }
  }

  // read_metapathesN (:339-357): stop at the first malformed line
  std::vector<std::vector<int32_t>> vts, ets;
  smcm_cli::read_metapaths(argv[6], vts, ets);
  std::cout << "*\tsuccessfully get metapathesN." << std::endl;

  int rc = 0;
  for (size_t p = 0; p < vts.size(); ++p) {
    const int32_t L = (int32_t)ets[p].size();
    double cost = 0, spg = 0, dp = 0, es = 0;
    int64_t nnz = 0;
    for (int it = 0; it < iterations; ++it) {
      smcm_result* r = nullptr;
      smcm_times t{};
      if (smcm_chain_run(ctx, g, vts[p].data(), ets[p].data(), L, nullptr, 0, &r, &t) != 0) {
        std::cerr << smcm_last_error() << std::endl;
        rc = 1;
        break;
      }
      cost += t.total / iterations;
      spg += t.spgemm / iterations;
      dp += t.dp / iterations;
      es += t.est / iterations;
      smcm_stats st;
      smcm_last_stats(ctx, &st);
      nnz = st.ref_nnz;  // nnz as the reference counts it (:859-868)
      smcm_result_free(r);
    }
    char mp[1 << 14];
    hin_metapath_to_string(vts[p].data(), ets[p].data(), L, mp, sizeof(mp));
    std::cout << "time cost of spgemm: " << spg << "s" << std::endl;
    std::cout << "time cost:" << cost << "s" << std::endl;
    std::ofstream out(argv[7], std::ios::out | std::ios::app);
    out << mp << " " << nnz << " " << cost << " " << spg << " " << dp << " " << es << std::endl;
  }
  smcm_graph_free(g);
  smcm_ctx_destroy(ctx);
  hin_free(hg);
  return rc;
}
