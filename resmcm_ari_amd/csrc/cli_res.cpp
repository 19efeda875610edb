// smcm_res - drop-in replacement of the reference's estimate-only program accuracy/res
// (main, accuracy/res.cpp:612-709): same positional arguments
//   THREADS ITERATIONS graph.txt vertex.txt edge.txt metapaths.txt output.txt
// and the same appended output line per meta-path (:697):
//   "<meta-path> <estimated nnz>"
// with the estimate c_[0][L-1] of final_result_estimate (:564-600) in the stream's default
// format (6 significant digits).  The estimator runs on the GPU (smcm_estimate); its cost sums
// are taken in row order on one lane, as the reference's sum at one thread.  THREADS is accepted
// for compatibility; an optional 8th argument selects the device.
#include <chrono>
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
  smcm_ctx* ctx = nullptr;
  smcm_graph* g = nullptr;
  if (smcm_ctx_create(device, 0, &ctx) != 0 || smcm_set_option(ctx, "exact_cost_sum", 1) != 0 ||
      smcm_graph_upload(ctx, hg->n, hg->vtype, hg->adj_ptr, hg->adj_nbr, hg->adj_et, &g) != 0) {
    std::cerr << smcm_last_error() << std::endl;
    return 1;
  }
  std::vector<std::vector<int32_t>> vts, ets;
  smcm_cli::read_metapaths(argv[6], vts, ets);

  int rc = 0;
  for (size_t p = 0; p < vts.size() && rc == 0; ++p) {
    const int32_t L = (int32_t)ets[p].size();
    double est = 0, est_time = 0;
    for (int it = 0; it < iterations; ++it) {
      const auto t0 = std::chrono::steady_clock::now();
      if (smcm_estimate(ctx, g, vts[p].data(), ets[p].data(), L, nullptr, 0, &est) != 0) {
        std::cerr << smcm_last_error() << std::endl;
        rc = 1;
        break;
      }
      est_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / iterations;
    }
    if (rc) break;
    std::cout << "time cost of estimate: " << est_time << "s" << std::endl;
    char mp[1 << 14];
    hin_metapath_to_string(vts[p].data(), ets[p].data(), L, mp, sizeof(mp));
    std::ofstream out(argv[7], std::ios::out | std::ios::app);
    out << mp << " " << est << std::endl;
  }
  smcm_graph_free(g);
  smcm_ctx_destroy(ctx);
  hin_free(hg);
  return rc;
}
