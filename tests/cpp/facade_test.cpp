// C++ facade test: the reference's smoke, multifrontal and static-pivoting
// assertions (test/unit/smoke_test.cpp, test/unit/multifrontal_test.cpp,
// test/correctness/kkt_static_pivoting_test.cpp) written against
// include/symla/solver.hpp and run on the device path.
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
  s.factorize(S);
  CHECK(s.factorized() && !s.isSingular());
  CHECK(s.inertia().n_pos + s.inertia().n_neg == 5 && s.inertia().n_zero == 0);
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

int main() {
  smoke();
  arrowAndSingular();
  staticRegularized();
  std::printf("facade_test: all checks passed\n");
  return 0;
}
