// smcm_metapaths - meta-path enumeration as tool/generate_DBLP.cpp (and generate_FourSquare.cpp,
// generate_IMDB.cpp) do it (:286-304, :308-368): every meta-path with L edges of a dataset's
// schema, printed one per line and written to a file.
//   smcm_metapaths L [DBLP|FourSquare|IMDB] [output.txt]
// The reference writes ../materials/input/metapath/<DS>/<DS>-l<L>.txt; here the output file is
// an argument (default <DS>-l<L>.txt in the working directory).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "smcm_hin.h"

namespace {

struct Schema {
  std::vector<int32_t> starts;
  std::vector<int32_t> src, et, dst;
};

bool schema_of(const std::string& ds, Schema& s) {
  auto rel = [&](int32_t a, int32_t e, int32_t b) {
    s.src.push_back(a);
    s.et.push_back(e);
    s.dst.push_back(b);
  };
  if (ds == "DBLP") {  // paper 0, author 1, venue 2, term 3
    s.starts = {0, 1, 2, 3};
    rel(0, 0, 1), rel(0, 1, 2), rel(0, 2, 3), rel(1, 3, 0), rel(2, 4, 0), rel(3, 5, 0);
  } else if (ds == "FourSquare") {  // venue 0, city 1, category 2, user 3, time 4
    s.starts = {0, 1, 2, 3, 4};
    rel(0, 0, 1), rel(0, 2, 2), rel(0, 4, 3), rel(0, 6, 4), rel(1, 1, 0), rel(2, 3, 0), rel(3, 5, 0), rel(4, 7, 0);
  } else if (ds == "IMDB") {  // movie 0, actor 1, director 2, writer 3
    s.starts = {0, 1, 2, 3};
    rel(0, 0, 1), rel(0, 2, 2), rel(0, 4, 3), rel(1, 1, 0), rel(2, 3, 0), rel(3, 5, 0);
  } else {
    return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2 || argc > 4) {
    std::cerr << "usage: " << argv[0] << " L [DBLP|FourSquare|IMDB] [output.txt]" << std::endl;
    return 2;
  }
  const int L = std::atoi(argv[1]);
  const std::string ds = argc >= 3 ? argv[2] : "DBLP";
  const std::string out = argc == 4 ? argv[3] : ds + "-l" + std::to_string(L) + ".txt";
  Schema s;
  if (L < 0 || !schema_of(ds, s)) {
    std::cerr << "unknown dataset or bad length" << std::endl;
    return 2;
  }
  int64_t n = 0;
  const int32_t nrel = (int32_t)s.src.size(), nst = (int32_t)s.starts.size();
  if (hin_enumerate_metapaths(s.src.data(), s.et.data(), s.dst.data(), nrel, s.starts.data(), nst, L, nullptr,
                              nullptr, 0, &n) != 0) {
    std::cerr << hin_last_error() << std::endl;
    return 1;
  }
  std::vector<int32_t> vts(n * (L + 1) + 1), ets(n * L + 1);
  hin_enumerate_metapaths(s.src.data(), s.et.data(), s.dst.data(), nrel, s.starts.data(), nst, L, vts.data(),
                          ets.data(), n, &n);
  std::ofstream f(out);
  char buf[4096];
  for (int64_t i = 0; i < n; ++i) {
    hin_metapath_to_string(vts.data() + i * (L + 1), ets.data() + i * L, L, buf, sizeof(buf));
    std::cout << buf << std::endl;
    f << buf << std::endl;
  }
  return 0;
}
