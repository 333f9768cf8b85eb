// mmio.cpp -- Matrix Market coordinate reader (the role of
// Eigen::loadMarket in the reference's real-matrix tests and benches,
// test/correctness/real_matrix_test.cpp:84, bench/tier1_timing_main.cpp:24).
//
// Reads "%%MatrixMarket matrix coordinate {real|integer|pattern}
// {general|symmetric|skew-symmetric}" files into CSC with the entries as
// stored (a symmetric file holds one triangle, conventionally the lower one,
// which is the triangle the solver reads), duplicates summed, rows sorted
// within each column.  Pattern files get the value 1.  Complex and array
// (dense) files are rejected.
#include "mmio.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

namespace symla_hip {

namespace {
std::string lower(std::string s) {
  for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}
}  // namespace

MarketMatrix readMatrixMarket(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::invalid_argument("symla: cannot open Matrix Market file " + path);
  std::string line;
  if (!std::getline(in, line)) throw std::invalid_argument("symla: empty Matrix Market file " + path);
  std::istringstream hs(line);
  std::string banner, object, format, field, symmetry;
  hs >> banner >> object >> format >> field >> symmetry;
  if (banner != "%%MatrixMarket") throw std::invalid_argument("symla: missing %%MatrixMarket banner in " + path);
  object = lower(object);
  format = lower(format);
  field = lower(field);
  symmetry = lower(symmetry);
  if (object != "matrix" || format != "coordinate")
    throw std::invalid_argument("symla: only coordinate matrices are supported (" + path + ")");
  const bool pattern = field == "pattern";
  if (!pattern && field != "real" && field != "integer" && field != "double")
    throw std::invalid_argument("symla: unsupported Matrix Market field '" + field + "' in " + path);
  MarketMatrix M;
  if (symmetry == "general")
    M.symmetry = 0;
  else if (symmetry == "symmetric")
    M.symmetry = 1;
  else if (symmetry == "skew-symmetric")
    M.symmetry = 2;
  else
    throw std::invalid_argument("symla: unsupported Matrix Market symmetry '" + symmetry + "' in " + path);
  // size line (after comments)
  int64_t rows = -1, cols = -1, nnz = -1;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '%') continue;
    std::istringstream ss(line);
    if (!(ss >> rows >> cols >> nnz)) throw std::invalid_argument("symla: bad size line in " + path);
    break;
  }
  if (rows < 0 || cols < 0 || nnz < 0) throw std::invalid_argument("symla: missing size line in " + path);
  if (rows > INT32_MAX || cols > INT32_MAX) throw std::invalid_argument("symla: matrix too large in " + path);
  M.rows = static_cast<int>(rows);
  M.cols = static_cast<int>(cols);
  struct T {
    int r, c;
    double v;
  };
  std::vector<T> ent;
  ent.reserve(static_cast<size_t>(nnz));
  for (int64_t k = 0; k < nnz; ++k) {
    if (!std::getline(in, line)) throw std::invalid_argument("symla: truncated entry list in " + path);
    if (line.empty() || line[0] == '%') {
      --k;
      continue;
    }
    const char* p = line.c_str();
    char* e = nullptr;
    const long i = std::strtol(p, &e, 10);
    if (e == p) throw std::invalid_argument("symla: bad entry line in " + path);
    p = e;
    const long j = std::strtol(p, &e, 10);
    if (e == p) throw std::invalid_argument("symla: bad entry line in " + path);
    p = e;
    double v = 1.0;
    if (!pattern) {
      v = std::strtod(p, &e);
      if (e == p) throw std::invalid_argument("symla: missing value in " + path);
    }
    if (i < 1 || i > rows || j < 1 || j > cols) throw std::invalid_argument("symla: entry index out of range in " + path);
    ent.push_back({static_cast<int>(i - 1), static_cast<int>(j - 1), v});
  }
  std::stable_sort(ent.begin(), ent.end(), [](const T& a, const T& b) { return a.c != b.c ? a.c < b.c : a.r < b.r; });
  M.colptr.assign(static_cast<size_t>(M.cols) + 1, 0);
  for (size_t t = 0; t < ent.size(); ++t) {
    if (!M.rowidx.empty() && t > 0 && ent[t].c == ent[t - 1].c && ent[t].r == ent[t - 1].r) {
      M.values.back() += ent[t].v;  // duplicates summed
      continue;
    }
    M.rowidx.push_back(ent[t].r);
    M.values.push_back(ent[t].v);
    ++M.colptr[ent[t].c + 1];
  }
  for (int c = 0; c < M.cols; ++c) M.colptr[c + 1] += M.colptr[c];
  return M;
}

}  // namespace symla_hip
