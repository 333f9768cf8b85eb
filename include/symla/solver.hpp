// symla/solver.hpp -- drop-in C++ facade of symla::SymLDLT<double> over the
// MI355X numeric phase (libsymla_hip.so, C-ABI in symla_hip.h).
//
// Same class, method names, argument meaning and exceptions as the
// reference facade (include/symla/solver.hpp:29-418 of the symla
// repository): analyzePattern / factorize / compute / refactorize / solve /
// solveWithRefinement / inertia / isSingular / singularColumns, the
// Mode::StaticRegularized configuration and diagnostics.  Eigen is not a
// dependency: matrices are the small CSC / column-major types below; when
// Eigen is available, overloads taking Eigen::SparseMatrix<double> and
// Eigen::MatrixXd are added at the end of this header.
//
// Link with -lsymla_hip (ma57_amd/libsymla_hip.so).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../symla_hip.h"

namespace symla {

enum class Mode { ThresholdPivot, StaticRegularized };       // solver.hpp:20-23
enum class OrderingType { AMD, Metis, Natural };              // ordering.hpp:31-35
// Front storage precision (no counterpart in the reference): FP32 keeps the
// fronts, multipliers and Schur updates in FP32 with FP64 pivots, D and
// inertia; solveWithRefinement then restores FP64 accuracy.
enum class Precision { FP64, FP32 };

struct Inertia {  // inertia.hpp:14-18
  int n_pos = 0;
  int n_neg = 0;
  int n_zero = 0;
};

struct SymbolicFactorOptions {  // symbolic.hpp:161-184
  int max_relax_size = 64;
  double max_relax_fill_fraction = 0.25;
};

struct RefinementOptions {  // refine.hpp:43-59
  int max_iterations = 10;
  double tolerance = 1e-14;
};

struct RefinementResult {  // refine.hpp:61-65, plus the symmetric-norm backward error
  int iterations = 0;
  double initial_relative_residual = 0.0;
  double final_relative_residual = 0.0;
  double backward_error = 0.0;
};

// Compressed sparse column matrix (stands in for Eigen::SparseMatrix<double,
// ColMajor, int>).  Either triangle or both may be stored; the lower
// triangle is authoritative (multifrontal.hpp:190-198).
struct SparseMatrix {
  int nrows = 0, ncols = 0;
  std::vector<int64_t> colptr{0};
  std::vector<int> rowidx;
  std::vector<double> values;
  SparseMatrix() = default;
  SparseMatrix(int r, int c) : nrows(r), ncols(c), colptr(c + 1, 0) {}
  int rows() const { return nrows; }
  int cols() const { return ncols; }
  int64_t nonZeros() const { return colptr.empty() ? 0 : colptr.back(); }
  // (row, col, value) triplets, duplicates kept in the order given
  static SparseMatrix fromTriplets(int n, const std::vector<int>& r, const std::vector<int>& c,
                                   const std::vector<double>& v) {
    SparseMatrix A(n, n);
    for (int col : c) ++A.colptr[col + 1];
    for (int j = 0; j < n; ++j) A.colptr[j + 1] += A.colptr[j];
    A.rowidx.resize(r.size());
    A.values.resize(r.size());
    std::vector<int64_t> pos(A.colptr.begin(), A.colptr.end() - 1);
    for (size_t t = 0; t < r.size(); ++t) {
      A.rowidx[pos[c[t]]] = r[t];
      A.values[pos[c[t]]++] = v[t];
    }
    return A;
  }
};

// Column-major dense matrix (stands in for Eigen::MatrixXd).
struct DenseMatrix {
  int nrows = 0, ncols = 0;
  std::vector<double> data;
  DenseMatrix() = default;
  DenseMatrix(int r, int c) : nrows(r), ncols(c), data(static_cast<size_t>(r) * c, 0.0) {}
  int rows() const { return nrows; }
  int cols() const { return ncols; }
  double& operator()(int i, int j) { return data[i + static_cast<size_t>(j) * nrows]; }
  double operator()(int i, int j) const { return data[i + static_cast<size_t>(j) * nrows]; }
};

namespace detail {
inline void check(int rc, const char* msg) {
  if (rc == SYMLA_OK) return;
  const std::string m = msg ? msg : "";
  if (rc == SYMLA_ERR_INVALID) throw std::invalid_argument(m);
  if (rc == SYMLA_ERR_LOGIC) throw std::logic_error(m);
  throw std::runtime_error(m);
}
}  // namespace detail

template <typename Scalar_>
class SymLDLT;

// The numeric phase computes in FP64 (multifrontal.hpp:121-122 stores double
// regardless of Scalar); SymLDLT<double> is the supported instantiation.
template <>
class SymLDLT<double> {
 public:
  using Scalar = double;
  using SparseMatrix = symla::SparseMatrix;
  using DenseMatrix = symla::DenseMatrix;

  SymLDLT() = default;
  SymLDLT(const SymLDLT&) = delete;
  SymLDLT& operator=(const SymLDLT&) = delete;
  ~SymLDLT() { release(); }

  void setMode(Mode mode) { mode_ = mode; }
  void setPrecision(Precision p) { precision_ = p; }
  void setPivotThreshold(double u) { pivot_threshold_ = u; }
  void setOrdering(OrderingType ordering) { ordering_ = ordering; }
  void setKKTBlockSizes(int n1, int n2) {
    if (n1 < 0 || n2 < 0) throw std::invalid_argument("symla::SymLDLT::setKKTBlockSizes: sizes must be nonnegative");
    kkt_n1_ = n1;
    kkt_n2_ = n2;
    use_kkt_blocks_ = true;
    use_explicit_signs_ = false;
  }
  void setExpectedSignPattern(std::vector<int> signs) {
    expected_signs_orig_ = std::move(signs);
    use_explicit_signs_ = true;
    use_kkt_blocks_ = false;
  }
  void setInitialRegularizationDelta(double delta) { initial_regularization_delta_ = delta; }
  void setMaxRegularizationRetries(int retries) { max_regularization_retries_ = retries; }
  void setRegularizationGrowth(double growth) { regularization_growth_ = growth; }
  void setStaticPivotFloor(double relativeFloor, double absoluteFloor = 1e-300) {
    static_pivot_relative_floor_ = relativeFloor;
    static_pivot_absolute_floor_ = absoluteFloor;
  }
  // Accepted for source compatibility; the device path has no host threads.
  void setParallel(bool) {}
  void setNumThreads(int) {}

  double lastRegularizationDelta() const { return last_info_.last_delta; }
  int regularizationRetriesUsed() const { return last_info_.retries_used; }
  double totalPerturbation() const { return last_info_.total_perturbation; }
  int numPerturbedPivots() const { return last_info_.num_perturbed; }
  // Front precision of the last factorization (FP32 static-regularized
  // factorizations may fall back to FP64; no reference counterpart).
  Precision precisionUsed() const {
    return last_info_.precision_used == SYMLA_PREC_FP32 ? Precision::FP32 : Precision::FP64;
  }

  void analyzePattern(const SparseMatrix& A, const SymbolicFactorOptions& options = SymbolicFactorOptions()) {
    if (A.rows() != A.cols()) throw std::invalid_argument("symla::computeOrdering: matrix must be square");
    release();
    symla_symbolic* sym = nullptr;
    int ord = SYMLA_ORDER_AMD;
    if (ordering_ == OrderingType::Metis) ord = SYMLA_ORDER_METIS;
    if (ordering_ == OrderingType::Natural) ord = SYMLA_ORDER_NATURAL;
    detail::check(symla_analyze(A.rows(), A.colptr.data(), A.rowidx.data(), ord, nullptr, options.max_relax_size,
                                options.max_relax_fill_fraction, &sym),
                  symla_last_error());
    sym_ = sym;
    detail::check(symla_plan_create(sym_, A.colptr.data(), A.rowidx.data(), &plan_), symla_last_error());
    n_ = A.rows();
    nnz_ = A.nonZeros();
    pattern_analyzed_ = true;
    factorized_ = false;
  }

  void factorize(const SparseMatrix& A) {
    if (!pattern_analyzed_) throw std::logic_error("symla::SymLDLT::factorize: analyzePattern must be called first");
    if (A.rows() != n_ || A.nonZeros() != nnz_)
      throw std::invalid_argument("symla::SymLDLT::factorize: matrix pattern differs from the analyzed pattern");
    symla_factor_opts o{};
    o.mode = mode_ == Mode::StaticRegularized ? SYMLA_MODE_STATIC : SYMLA_MODE_THRESHOLD;
    o.pivot_threshold = pivot_threshold_;
    o.zero_tolerance = 1e-300;
    o.relative_pivot_floor = 1e-12;
    o.initial_delta = initial_regularization_delta_;
    o.max_retries = max_regularization_retries_;
    o.growth = regularization_growth_;
    o.static_rel_floor = static_pivot_relative_floor_;
    o.static_abs_floor = static_pivot_absolute_floor_;
    o.precision = precision_ == Precision::FP32 ? SYMLA_PREC_FP32 : SYMLA_PREC_FP64;
    std::vector<int> signs;
    if (mode_ == Mode::StaticRegularized) {
      signs = buildExpectedSignOriginal(n_);
      if (use_kkt_blocks_ || use_explicit_signs_) o.expected_sign = signs.data();
    }
    factorized_ = false;
    symla_factor_info info{};
    detail::check(symla_factorize(plan_, A.values.data(), &o, &info), symla_plan_last_error(plan_));
    last_info_ = info;
    inertia_ = Inertia{info.n_pos, info.n_neg, info.n_zero};
    singular_ = info.singular != 0;
    factorized_ = true;
  }

  void compute(const SparseMatrix& A) {
    analyzePattern(A);
    factorize(A);
  }
  void refactorize(const SparseMatrix& A) { factorize(A); }

  DenseMatrix solve(const DenseMatrix& B) const {
    if (!factorized_) throw std::logic_error("symla::SymLDLT::solve: factorize() must be called before solve()");
    if (singular_)
      throw std::logic_error(
          "symla::SymLDLT::solve: factorization is numerically singular (see isSingular()/singularColumns()); "
          "refusing to solve");
    if (B.rows() != n_)
      throw std::invalid_argument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size");
    DenseMatrix X(B.rows(), B.cols());
    if (B.cols() == 0 || n_ == 0) return X;
    detail::check(symla_solve(plan_, B.data.data(), n_, B.cols(), X.data.data(), n_), symla_plan_last_error(plan_));
    return X;
  }

  struct RefinedSolve {
    DenseMatrix x;
    RefinementResult diagnostics;
  };
  RefinedSolve solveWithRefinement(const SparseMatrix& A, const DenseMatrix& B,
                                   const RefinementOptions& options = RefinementOptions()) const {
    if (!factorized_) throw std::logic_error("symla::SymLDLT::solve: factorize() must be called before solve()");
    if (singular_)
      throw std::logic_error(
          "symla::SymLDLT::solve: factorization is numerically singular (see isSingular()/singularColumns()); "
          "refusing to solve");
    if (B.rows() != n_)
      throw std::invalid_argument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size");
    RefinedSolve out;
    out.x = DenseMatrix(B.rows(), B.cols());
    if (B.cols() == 0 || n_ == 0) return out;
    symla_refine_opts ro{options.max_iterations, options.tolerance};
    symla_refine_info ri{};
    detail::check(symla_solve_refined(plan_, A.values.data(), B.data.data(), n_, B.cols(), out.x.data.data(), n_, &ro,
                                      &ri),
                  symla_plan_last_error(plan_));
    out.diagnostics = RefinementResult{ri.iterations, ri.initial_relative_residual, ri.final_relative_residual,
                                       ri.backward_error};
    return out;
  }

  const Inertia& inertia() const { return inertia_; }
  bool patternAnalyzed() const { return pattern_analyzed_; }
  bool factorized() const { return factorized_; }
  bool isSingular() const { return singular_; }
  std::vector<int> singularColumns() const {
    if (!plan_) return {};
    const int cnt = symla_singular_cols(plan_, nullptr, 0);
    std::vector<int> out(cnt > 0 ? cnt : 0);
    if (cnt > 0) symla_singular_cols(plan_, out.data(), cnt);
    return out;
  }

  // Escape hatches for device-resident inputs (no reference counterpart).
  symla_plan* plan() const { return plan_; }

 private:
  Mode mode_ = Mode::ThresholdPivot;
  Precision precision_ = Precision::FP64;
  double pivot_threshold_ = 0.01;
  OrderingType ordering_ = OrderingType::AMD;
  bool pattern_analyzed_ = false;
  bool factorized_ = false;
  bool singular_ = false;
  Inertia inertia_;
  int n_ = 0;
  int64_t nnz_ = 0;
  symla_symbolic* sym_ = nullptr;
  symla_plan* plan_ = nullptr;
  symla_factor_info last_info_{};
  bool use_kkt_blocks_ = false;
  int kkt_n1_ = 0, kkt_n2_ = 0;
  bool use_explicit_signs_ = false;
  std::vector<int> expected_signs_orig_;
  double initial_regularization_delta_ = 0.0;
  int max_regularization_retries_ = 10;
  double regularization_growth_ = 10.0;
  double static_pivot_relative_floor_ = 1e-8;
  double static_pivot_absolute_floor_ = 1e-300;

  void release() {
    if (plan_) symla_plan_destroy(plan_);
    if (sym_) symla_symbolic_free(sym_);
    plan_ = nullptr;
    sym_ = nullptr;
  }

  // solver.hpp:299-317
  std::vector<int> buildExpectedSignOriginal(int n) const {
    if (use_explicit_signs_) {
      if (static_cast<int>(expected_signs_orig_.size()) != n)
        throw std::invalid_argument(
            "symla::SymLDLT::factorize: setExpectedSignPattern's vector size does not match the matrix size");
      return expected_signs_orig_;
    }
    std::vector<int> signs(n, 0);
    if (use_kkt_blocks_) {
      if (kkt_n1_ + kkt_n2_ != n)
        throw std::invalid_argument(
            "symla::SymLDLT::factorize: setKKTBlockSizes(n1, n2) sizes do not sum to the matrix size");
      for (int i = 0; i < kkt_n1_; ++i) signs[i] = -1;
      for (int i = kkt_n1_; i < n; ++i) signs[i] = +1;
    }
    return signs;
  }
};

}  // namespace symla

#if defined(__has_include)
#if __has_include(<Eigen/Sparse>)
#include <Eigen/Dense>
#include <Eigen/Sparse>
namespace symla {
// Input adaptors for callers holding Eigen types (same computation path).
inline SparseMatrix fromEigen(const Eigen::SparseMatrix<double, Eigen::ColMajor, int>& A) {
  SparseMatrix S(static_cast<int>(A.rows()), static_cast<int>(A.cols()));
  for (int c = 0; c < A.outerSize(); ++c) {
    for (Eigen::SparseMatrix<double, Eigen::ColMajor, int>::InnerIterator it(A, c); it; ++it) {
      S.rowidx.push_back(static_cast<int>(it.row()));
      S.values.push_back(it.value());
    }
    S.colptr[c + 1] = static_cast<int64_t>(S.rowidx.size());
  }
  return S;
}
inline DenseMatrix fromEigen(const Eigen::MatrixXd& B) {
  DenseMatrix D(static_cast<int>(B.rows()), static_cast<int>(B.cols()));
  for (int j = 0; j < B.cols(); ++j)
    for (int i = 0; i < B.rows(); ++i) D(i, j) = B(i, j);
  return D;
}
inline Eigen::MatrixXd toEigen(const DenseMatrix& D) {
  Eigen::MatrixXd B(D.rows(), D.cols());
  for (int j = 0; j < D.cols(); ++j)
    for (int i = 0; i < D.rows(); ++i) B(i, j) = D(i, j);
  return B;
}
}  // namespace symla
#endif
#endif
