// providers.h - the reference's alternative DP cost providers (MetaAC, MNC, DMap, LGraph).
#pragma once

#include <cstdint>
#include <vector>

#include "alloc.h"
#include "common.h"

namespace smcm {

constexpr int kProviderMetaAC = 2;  // efficient/meta.cpp, accuracy/meta.cpp
constexpr int kProviderMNC = 3;     // efficient/mnc.cpp, accuracy/mnc.cpp
constexpr int kProviderDMap = 4;    // efficient/dm.cpp, accuracy/dm.cpp
constexpr int kProviderLGraph = 5;  // efficient/lg.cpp, accuracy/lg.cpp

// One meta-path's matrices in the reference's global index space.
struct ProviderChain {
  int32_t L = 0;
  const Csr* mats = nullptr;                      // L typed matrices (local rows and columns)
  int32_t n = 0;                                  // vertices of the graph
  std::vector<const int32_t*> rowg, colg;         // device, per matrix: local row / column -> global id
  std::vector<std::vector<int32_t>> col_type_ids; // host, per matrix: global ids of its column type
  std::vector<bool> same_type;                    // per matrix: row type == column type
  std::vector<int32_t> blo, bhi;                  // per position 0..L: 256-block range of its ids
  uint64_t seed = 0;                              // LGraph ranks
};

struct ProviderOut {
  std::vector<double> c;     // L*L: the provider's per-interval quantity (density or size)
  std::vector<double> m;     // L*L: DP costs
  std::vector<int32_t> s;    // L*L: DP split points
  double estimate = 0;       // the estimate the reference's accuracy program prints
};

// accuracy = the estimate-only variant (accuracy/*.cpp) instead of the planner (efficient/*.cpp)
void run_provider(int provider, bool accuracy, const ProviderChain& ch, ProviderOut& out, DevAlloc& al);

}  // namespace smcm
