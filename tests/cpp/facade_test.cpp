// C++ facade test: the reference's smoke, multifrontal and static-pivoting
// assertions (test/unit/smoke_test.cpp, test/unit/multifrontal_test.cpp,
// test/correctness/kkt_static_pivoting_test.cpp) written against
// include/symla/solver.hpp and run on the device path, plus a flat
// reconstruction of the device factor read through numericFactor() (the
// reference's own check, test/unit/multifrontal_test_helpers.hpp:40-88).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <vector>

#include "symla/solver.hpp"

#define CHECK(cond)                                                            \
  do {                                                                         \
    if (!(cond)) {                                                             \
      std::fprintf(stderr, "CHECK failed: %s at %s:%d\n", #cond, __FILE__, __LINE__); \
      std::abort();                                                            \
    }                                                                          \
  } while (0)

using symla::DenseMatrix;
using symla::SparseMatrix;

// Elimination order, L and D of the whole factor assembled from the fronts
// (every finalized pivot in front order, which is a valid elimination order);
// returns max |P A P^T - L D L^T| over the dense reconstruction.
static double flattenResidual(const symla::SymbolicFactor& sf, const symla::NumericFactor& nf,
                              const std::vector<double>& Adense) {
  const int n = nf.n;
  std::vector<int> order;
  for (const auto& f : nf.fronts)
    for (int p = 0; p < f.nPivots; ++p) order.push_back(f.rowIndices[p]);
  CHECK(static_cast<int>(order.size()) == n);
  std::vector<int> pos(n, -1);
  for (int t = 0; t < n; ++t) pos[order[t]] = t;
  std::vector<double> L(static_cast<size_t>(n) * n, 0.0), D(static_cast<size_t>(n) * n, 0.0);
  auto at = [n](std::vector<double>& M, int i, int j) -> double& { return M[i + static_cast<size_t>(j) * n]; };
  for (int t = 0; t < n; ++t) at(L, t, t) = 1.0;
  for (const auto& f : nf.fronts) {
    const int m = static_cast<int>(f.rowIndices.size()), k = f.nPivots;
    std::vector<char> d21(k, 0);
    for (const auto& b : f.pivotBlocks)
      if (b.kind == symla::PivotKind::TwoByTwo) d21[b.start + 1] = 1;
    for (int p = 0; p < k; ++p) {
      for (int i = p + 1; i < m; ++i) {
        if (i < k && d21[i] && i == p + 1) continue;
        at(L, pos[f.rowIndices[i]], pos[f.rowIndices[p]]) = f.L(i, p);
      }
      for (int q = 0; q < k; ++q) at(D, pos[f.rowIndices[p]], pos[f.rowIndices[q]]) = f.D(p, q);
    }
  }
  double worst = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      double r = 0.0;
      for (int a = 0; a < n; ++a) {
        const double lia = at(L, i, a);
        if (lia == 0.0) continue;
        for (int b = 0; b < n; ++b) r += lia * at(D, a, b) * at(L, j, b);
      }
      const int oi = sf.perm[order[i]], oj = sf.perm[order[j]];
      worst = std::max(worst, std::abs(r - Adense[oi * n + oj]));
    }
  return worst;
}

static SparseMatrix fromDense(int n, const std::vector<double>& a) {
  std::vector<int> r, c;
  std::vector<double> v;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      if (a[i * n + j] != 0.0) {
        r.push_back(i);
        c.push_back(j);
        v.push_back(a[i * n + j]);
      }
  return SparseMatrix::fromTriplets(n, r, c, v);
}

static void smoke() {  // smoke_test.cpp:24-62
  symla::SymLDLT<double> s;
  s.setPivotThreshold(0.01);
  s.setMode(symla::Mode::ThresholdPivot);
  CHECK(!s.patternAnalyzed() && !s.factorized());
  SparseMatrix A = SparseMatrix::fromTriplets(2, {0, 1}, {0, 1}, {2.0, 3.0});
  s.analyzePattern(A);
  CHECK(s.patternAnalyzed() && !s.factorized());
  s.factorize(A);
  CHECK(s.factorized() && !s.isSingular());
  CHECK(s.inertia().n_pos == 2 && s.inertia().n_neg == 0 && s.inertia().n_zero == 0);
  DenseMatrix b(2, 1);
  b(0, 0) = 4.0;
  b(1, 0) = 9.0;
  DenseMatrix x = s.solve(b);
  CHECK(std::abs(x(0, 0) - 2.0) < 1e-12 && std::abs(x(1, 0) - 3.0) < 1e-12);
  symla::SymLDLT<double> unfactored;
  bool threw = false;
  try {
    unfactored.solve(DenseMatrix(2, 1));
  } catch (const std::logic_error&) {
    threw = true;
  }
  CHECK(threw);
}

static void arrowAndSingular() {  // multifrontal_test.cpp:102-196
  std::vector<double> a(25, 0.0);
  auto at = [&](int i, int j) -> double& { return a[i * 5 + j]; };
  at(2, 2) = 5;
  at(3, 3) = 7;
  at(4, 4) = 20;
  at(2, 3) = at(3, 2) = 1.1;
  at(3, 4) = at(4, 3) = 1.3;
  at(2, 4) = at(4, 2) = 0.9;
  at(0, 4) = at(4, 0) = 2.0;
  at(1, 3) = at(3, 1) = 1.7;
  SparseMatrix S = fromDense(5, a);
  symla::SymLDLT<double> s;
  s.setOrdering(symla::OrderingType::Natural);
  symla::SymbolicFactorOptions noAmalg;
  noAmalg.max_relax_size = 1;
  noAmalg.max_relax_fill_fraction = 0.0;
  s.analyzePattern(S, noAmalg);
  const symla::SymbolicFactor& sf = s.symbolicFactor();
  CHECK(sf.supernodes.size() == 3);
  CHECK(sf.supernodes[0].ncols == 1 && sf.supernodes[0].firstCol == 0);
  CHECK(sf.supernodes[1].ncols == 1 && sf.supernodes[1].firstCol == 1);
  CHECK(sf.supernodes[2].ncols == 3 && sf.supernodes[2].firstCol == 2);
  s.factorize(S);
  CHECK(s.factorized() && !s.isSingular());
  CHECK(s.inertia().n_pos + s.inertia().n_neg == 5 && s.inertia().n_zero == 0);
  // multifrontal_test.cpp:143-158: both leaves delay their column; the hub takes all 5 pivots
  const symla::NumericFactor& nf = s.numericFactor();
  CHECK(nf.fronts.size() == 3);
  CHECK(nf.fronts[0].nPivots == 0 && nf.fronts[1].nPivots == 0 && nf.fronts[2].nPivots == 5);
  CHECK(nf.inertia.n_zero == 0 && nf.parentSN[0] == 2 && nf.parentSN[1] == 2 && nf.rootsSN.size() == 1);
  CHECK(flattenResidual(sf, nf, a) < 1e-10 * 20.0);
  DenseMatrix b(5, 1);
  for (int i = 0; i < 5; ++i) b(i, 0) = i + 1;
  DenseMatrix x = s.solve(b);
  double rn = 0, xn = 0, bn = 0;
  for (int i = 0; i < 5; ++i) {
    double r = -b(i, 0);
    for (int j = 0; j < 5; ++j) r += a[i * 5 + j] * x(j, 0);
    rn += r * r;
    xn += x(i, 0) * x(i, 0);
    bn += b(i, 0) * b(i, 0);
  }
  CHECK(std::sqrt(rn) / (20.0 * std::sqrt(xn) + std::sqrt(bn)) < 1e-10);

  std::vector<double> g(9, 0.0);
  g[0] = 4.0;
  g[8] = 3.0;
  g[2] = g[6] = 0.5;
  symla::SymLDLT<double> t;
  t.compute(fromDense(3, g));
  CHECK(t.factorized() && t.isSingular());
  auto cols = t.singularColumns();
  CHECK(cols.size() == 1 && cols[0] == 1);
  bool threw = false;
  try {
    t.solve(DenseMatrix(3, 1));
  } catch (const std::logic_error&) {
    threw = true;
  }
  CHECK(threw);
}

static void staticRegularized() {  // kkt_static_pivoting_test.cpp: SQD and the throw path
  const int n1 = 30, n2 = 15, n = n1 + n2;
  std::mt19937 rng(11);
  std::uniform_real_distribution<double> u(-0.5, 0.5);
  std::vector<int> r, c;
  std::vector<double> v;
  for (int i = 0; i < n1; ++i) {
    r.push_back(i);
    c.push_back(i);
    v.push_back(-4.0);
    if (i + 1 < n1) {
      r.push_back(i + 1);
      c.push_back(i);
      v.push_back(-1.0);
    }
  }
  for (int i = 0; i < n2; ++i) {
    r.push_back(n1 + i);
    c.push_back(n1 + i);
    v.push_back(3.0);
    r.push_back(n1 + i);
    c.push_back(static_cast<int>(rng() % n1));
    v.push_back(u(rng));
  }
  SparseMatrix K = SparseMatrix::fromTriplets(n, r, c, v);
  symla::SymLDLT<double> s;
  s.setMode(symla::Mode::StaticRegularized);
  s.setKKTBlockSizes(n1, n2);
  s.analyzePattern(K);
  s.factorize(K);
  CHECK(s.inertia().n_pos == n2 && s.inertia().n_neg == n1 && s.inertia().n_zero == 0);
  CHECK(s.regularizationRetriesUsed() <= 1);
  DenseMatrix B(n, 2);
  for (int i = 0; i < n; ++i) {
    B(i, 0) = 1.0;
    B(i, 1) = i % 3;
  }
  auto ref = s.solveWithRefinement(K, B);
  CHECK(ref.diagnostics.backward_error < 1e-12);
  // wrong-signed diagonal with a zero retry budget must throw runtime_error
  v[0] = 0.05;
  SparseMatrix Kn = SparseMatrix::fromTriplets(n, r, c, v);
  symla::SymLDLT<double> t;
  t.setMode(symla::Mode::StaticRegularized);
  t.setKKTBlockSizes(n1, n2);
  t.setInitialRegularizationDelta(1e-6);
  t.setMaxRegularizationRetries(0);
  t.analyzePattern(Kn);
  bool threw = false;
  try {
    t.factorize(Kn);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  CHECK(threw && !t.factorized());
  bool inval = false;
  try {
    t.setKKTBlockSizes(-1, 2);
  } catch (const std::invalid_argument&) {
    inval = true;
  }
  CHECK(inval);
}

// A 2D 5-point Laplacian on 9 x 8 minus 0.3 I with random symmetric
// couplings, nested dissection, default amalgamation: the device factor read
// back through numericFactor() reconstructs P A P^T.
static void flattenReconstruction() {
  const int nx = 9, ny = 8, n = nx * ny;
  std::vector<double> a(static_cast<size_t>(n) * n, 0.0);
  auto at = [&](int i, int j) -> double& { return a[i * n + j]; };
  std::mt19937 rng(5);
  std::normal_distribution<double> nd(0.0, 0.3);
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y) {
      const int i = x * ny + y;
      at(i, i) = 4.0 - 0.3;
      if (x + 1 < nx) at(i, i + ny) = at(i + ny, i) = -1.0;
      if (y + 1 < ny) at(i, i + 1) = at(i + 1, i) = -1.0;
    }
  for (int t = 0; t < 40; ++t) {
    const int i = static_cast<int>(rng() % n), j = static_cast<int>(rng() % n);
    if (i != j) at(i, j) = at(j, i) = at(i, j) + nd(rng);
  }
  std::vector<int> r, c;
  std::vector<double> v;
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i)
      if (at(i, j) != 0.0) {
        r.push_back(i);
        c.push_back(j);
        v.push_back(at(i, j));
      }
  SparseMatrix A = SparseMatrix::fromTriplets(n, r, c, v);
  for (auto ord : {symla::OrderingType::Metis, symla::OrderingType::AMD}) {
    symla::SymLDLT<double> s;
    s.setOrdering(ord);
    s.compute(A);
    CHECK(s.factorized() && !s.isSingular());
    const symla::SymbolicFactor& sf = s.symbolicFactor();
    const symla::NumericFactor& nf = s.numericFactor();
    CHECK(nf.fronts.size() == sf.supernodes.size() && nf.n == n);
    int64_t cc = 0;
    for (int j = 0; j < n; ++j) cc += sf.etree.colCount[j];
    int64_t storage = 0;
    for (const auto& sn : sf.supernodes) storage += sn.storageNnz();
    CHECK(cc > 0 && cc <= storage);
    int pos = 0, neg = 0;
    for (const auto& f : nf.fronts)
      for (const auto& b : f.pivotBlocks) {  // dense_kernel.hpp:671-676, :753-768
        const int p = b.start;
        if (b.kind == symla::PivotKind::TwoByTwo) {
          const double det = f.D(p, p) * f.D(p + 1, p + 1) - f.D(p + 1, p) * f.D(p + 1, p);
          if (det < 0) {
            ++pos;
            ++neg;
          } else {
            (f.D(p, p) + f.D(p + 1, p + 1) > 0 ? pos : neg) += 2;
          }
        } else {
          (f.D(p, p) > 0 ? pos : neg) += 1;
        }
      }
    CHECK(pos == s.inertia().n_pos && neg == s.inertia().n_neg);
    CHECK(flattenResidual(sf, nf, a) < 1e-10 * 4.0);
  }
}

int main() {
  smoke();
  arrowAndSingular();
  flattenReconstruction();
  staticRegularized();
  std::printf("facade_test: all checks passed\n");
  return 0;
}
