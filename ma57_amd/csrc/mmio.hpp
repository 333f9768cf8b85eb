// mmio.hpp -- Matrix Market coordinate reader (see mmio.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace symla_hip {

struct MarketMatrix {
  int rows = 0, cols = 0;
  int symmetry = 0;  // 0 general, 1 symmetric, 2 skew-symmetric (entries as stored)
  std::vector<int64_t> colptr;
  std::vector<int> rowidx;
  std::vector<double> values;
};

MarketMatrix readMatrixMarket(const std::string& path);

}  // namespace symla_hip
