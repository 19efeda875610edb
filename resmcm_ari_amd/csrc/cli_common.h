// cli_common.h - pieces shared by the drop-in programs smcm_resmcm and smcm_res: the input
// banner of the reference's programs and their meta-path list reader.
#pragma once

#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "smcm_hin.h"

namespace smcm_cli {

// the banner of main (efficient/resmcm.cpp:807-815, accuracy/res.cpp:637-645)
inline void banner(int threads, double iterations, char** argv) {
  std::cout << "******************** input parameters ***************************" << std::endl;
  std::cout << "*\t# of thread:" << threads << std::endl;
  std::cout << "*\t# of iteration:" << iterations << std::endl;
  std::cout << "*\tthe input graph file:" << argv[3] << std::endl;
  std::cout << "*\tthe input vertex file:" << argv[4] << std::endl;
  std::cout << "*\tthe input edge file:" << argv[5] << std::endl;
  std::cout << "*\tthe input meta-path:" << argv[6] << std::endl;
  std::cout << "*\tthe output file:" << argv[7] << std::endl;
  std::cout << "*****************************************************************" << std::endl;
}

// read_metapathesN (efficient/resmcm.cpp:339-357): the first token of every line is a meta-path;
// reading stops at the first malformed line
inline void read_metapaths(const char* file, std::vector<std::vector<int32_t>>& vts,
                           std::vector<std::vector<int32_t>>& ets) {
  std::ifstream f(file);
  std::string line;
  while (std::getline(f, line)) {
    std::string tok;
    for (char ch : line) {
      if (ch == ' ' || ch == '\t' || ch == '\r') {
        if (!tok.empty()) break;
        continue;
      }
      tok.push_back(ch);
    }
    std::vector<int32_t> v(4096), e(4096);
    int32_t L = 0;
    if (tok.empty() || hin_parse_metapath(tok.c_str(), v.data(), e.data(), 4096, &L) != 0) {
      std::cerr << "the meta-path is incorrect" << std::endl;
      break;
    }
    v.resize(L + 1);
    e.resize(L);
    vts.push_back(v);
    ets.push_back(e);
  }
}

}  // namespace smcm_cli
