// symla/solver.hpp -- drop-in C++ facade of symla::SymLDLT<double> over the
// MI355X numeric phase (libsymla_hip.so, C-ABI in symla_hip.h).
//
// Same class, method names, argument meaning and exceptions as the
// reference facade (include/symla/solver.hpp:29-418 of the symla
// repository): analyzePattern / factorize / compute / refactorize / solve /
// solveWithRefinement / inertia / isSingular / singularColumns /
// symbolicFactor / numericFactor, the Mode::StaticRegularized configuration
// and diagnostics.
//
// Matrix types.  When <Eigen/Sparse> is on the include path (and
// SYMLA_NO_EIGEN is not defined), SymLDLT<double>::SparseMatrix is
// Eigen::SparseMatrix<double, Eigen::ColMajor, int> and DenseMatrix is
// Eigen::MatrixXd, exactly as in the reference (solver.hpp:33-34), so a
// reference caller compiles unchanged.  Without Eigen they are the small CSC
// / column-major types below (symla::CscMatrix, symla::ColMatrix), which
// offer the subset of the Eigen interface the facade and its callers use.
// Either way the matrices are only read to hand plain pointers to the
// C-ABI: there is one computation path.
//
// Link with -lsymla_hip (ma57_amd/libsymla_hip.so).
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../symla_hip.h"

#if !defined(SYMLA_NO_EIGEN) && defined(__has_include)
#if __has_include(<Eigen/Sparse>)
#define SYMLA_FACADE_EIGEN 1
#include <Eigen/Dense>
#include <Eigen/Sparse>
#endif
#endif

namespace symla {

enum class Mode { ThresholdPivot, StaticRegularized };       // solver.hpp:20-23
enum class OrderingType { AMD, Metis, Natural };              // ordering.hpp:31-35
// Front storage precision (no counterpart in the reference): FP32 keeps the
// fronts, multipliers and Schur updates in FP32 with FP64 pivots, D and
// inertia; solveWithRefinement then restores FP64 accuracy.  BF16x3: FP32
// fronts whose deferred contribution-block updates run as BF16x3 split
// products accumulated in FP32.
enum class Precision { FP64, FP32, BF16x3 };

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
struct CscMatrix {
  int nrows = 0, ncols = 0;
  std::vector<int64_t> colptr{0};
  std::vector<int> rowidx;
  std::vector<double> values;
  CscMatrix() = default;
  CscMatrix(int r, int c) : nrows(r), ncols(c), colptr(c + 1, 0) {}
  int rows() const { return nrows; }
  int cols() const { return ncols; }
  int outerSize() const { return ncols; }
  int64_t nonZeros() const { return colptr.empty() ? 0 : colptr.back(); }
  // (row, col, value) triplets, duplicates kept in the order given
  static CscMatrix fromTriplets(int n, const std::vector<int>& r, const std::vector<int>& c,
                                const std::vector<double>& v) {
    CscMatrix A(n, n);
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
struct ColMatrix {
  int nrows = 0, ncols = 0;
  std::vector<double> values;
  ColMatrix() = default;
  ColMatrix(int r, int c) : nrows(r), ncols(c), values(static_cast<size_t>(r) * c, 0.0) {}
  int rows() const { return nrows; }
  int cols() const { return ncols; }
  double* data() { return values.data(); }
  const double* data() const { return values.data(); }
  void resize(int r, int c) {
    nrows = r;
    ncols = c;
    values.assign(static_cast<size_t>(r) * c, 0.0);
  }
  double& operator()(int i, int j) { return values[i + static_cast<size_t>(j) * nrows]; }
  double operator()(int i, int j) const { return values[i + static_cast<size_t>(j) * nrows]; }
};

// Dense vector (stands in for Eigen::VectorXi / VectorXd).
template <typename T>
struct ColVector {
  std::vector<T> values;
  ColVector() = default;
  explicit ColVector(int n) : values(n, T(0)) {}
  int size() const { return static_cast<int>(values.size()); }
  void resize(int n) { values.assign(n, T(0)); }
  T* data() { return values.data(); }
  const T* data() const { return values.data(); }
  T& operator[](int i) { return values[i]; }
  T operator[](int i) const { return values[i]; }
  T& operator()(int i) { return values[i]; }
  T operator()(int i) const { return values[i]; }
};

#ifdef SYMLA_FACADE_EIGEN
using SparseMatrix = Eigen::SparseMatrix<double, Eigen::ColMajor, int>;
using DenseMatrix = Eigen::MatrixXd;
using IndexVector = Eigen::VectorXi;
using RealVector = Eigen::VectorXd;
#else
using SparseMatrix = CscMatrix;
using DenseMatrix = ColMatrix;
using IndexVector = ColVector<int>;
using RealVector = ColVector<double>;
#endif

// ---- symbolic and numeric factor views (symbolic.hpp:130-159,
// elimination_tree.hpp:50-53, dense_kernel.hpp:137-207,
// multifrontal.hpp:118-186) ----

struct Supernode {
  int firstCol = 0;
  int ncols = 0;
  std::vector<int> rowPattern;  // sorted; the first ncols entries are the own columns
  std::int64_t storageNnz() const {
    const std::int64_t m = static_cast<std::int64_t>(rowPattern.size());
    const std::int64_t k = ncols;
    return k * m - k * (k - 1) / 2;
  }
};

struct EliminationTree {
  IndexVector parent;    // parent[j] in final order, -1 at a root
  IndexVector colCount;  // nnz(L(:, j)) including the diagonal
  int n = 0;
};

// perm[k] = original index of the k-th pivot.  fundamentalSupernodes is
// left empty: the analysis keeps only the amalgamated partition.
struct SymbolicFactor {
  IndexVector perm;
  EliminationTree etree;
  std::vector<Supernode> supernodes;
  std::vector<Supernode> fundamentalSupernodes;
};

template <typename Scalar>
struct BlockDiagonalD {
  RealVector diag;     // D(k, k)
  RealVector offdiag;  // D(k+1, k) for a 2x2 block starting at k, else 0
  int size() const { return static_cast<int>(diag.size()); }
  int rows() const { return size(); }
  int cols() const { return size(); }
  Scalar operator()(int i, int j) const {
    if (i == j) return diag(i);
    if (i == j + 1) return offdiag(j);
    if (j == i + 1) return offdiag(i);
    return Scalar(0);
  }
};

enum class PivotKind { OneByOne, TwoByTwo };
struct PivotBlock {
  int start;
  PivotKind kind;
};

// One front: rowIndices (final order, size m), nPivots, L (m x nPivots; the
// (k+1, k) slot of a 2x2 block holds d21, as in the reference), D, the pivot
// blocks and Ltop (L's leading nPivots x nPivots block with the d21 slots
// zeroed).
struct SupernodeFactor {
  std::vector<int> rowIndices;
  int nPivots = 0;
  DenseMatrix L;
  BlockDiagonalD<double> D;
  std::vector<PivotBlock> pivotBlocks;
  DenseMatrix Ltop;
};

struct NumericFactor {
  std::vector<SupernodeFactor> fronts;  // fronts[si]: supernode si (postorder)
  IndexVector perm;                     // SymbolicFactor::perm
  Inertia inertia;
  bool singular = false;
  std::vector<int> singularCols;        // final-order indices
  int n = 0;
  double totalPerturbation = 0.0;
  int numPerturbed = 0;
  std::vector<int> parentSN;
  std::vector<std::vector<int>> childrenSN;
  std::vector<int> rootsSN;
};

namespace detail {
inline void check(int rc, const char* msg) {
  if (rc == SYMLA_OK) return;
  const std::string m = msg ? msg : "";
  if (rc == SYMLA_ERR_INVALID) throw std::invalid_argument(m);
  if (rc == SYMLA_ERR_LOGIC) throw std::logic_error(m);
  throw std::runtime_error(m);
}

// Plain CSC arrays of a sparse matrix for the C-ABI (64-bit column starts).
struct CscArrays {
  int n = 0;
  std::vector<int64_t> colptr;
  const int* rowidx = nullptr;
  const double* values = nullptr;
  int64_t nnz = 0;
#ifdef SYMLA_FACADE_EIGEN
  SparseMatrix compressed;  // a compressed copy when the caller's matrix is not
#endif
};
inline void cscArrays(const CscMatrix& A, CscArrays& out) {
  out.n = A.cols();
  out.colptr = A.colptr;
  out.rowidx = A.rowidx.data();
  out.values = A.values.data();
  out.nnz = A.nonZeros();
}
#ifdef SYMLA_FACADE_EIGEN
inline void cscArrays(const SparseMatrix& A, CscArrays& out) {
  const SparseMatrix* src = &A;
  if (!A.isCompressed()) {
    out.compressed = A;
    out.compressed.makeCompressed();
    src = &out.compressed;
  }
  out.n = static_cast<int>(src->cols());
  out.colptr.assign(src->outerIndexPtr(), src->outerIndexPtr() + src->cols() + 1);
  out.rowidx = src->innerIndexPtr();
  out.values = src->valuePtr();
  out.nnz = src->nonZeros();
}
#endif
template <typename V>
inline void toVector(const std::vector<int>& in, V& out) {
  out.resize(static_cast<int>(in.size()));
  for (size_t i = 0; i < in.size(); ++i) out[static_cast<int>(i)] = in[i];
}
}  // namespace detail

// Eigen::loadMarket's role in the reference's real-matrix tests and benches
// (test/correctness/real_matrix_test.cpp:84): reads a coordinate Matrix
// Market file, entries as stored, duplicates summed.  Returns false when
// the file cannot be read (as Eigen::loadMarket does).
inline bool loadMarket(SparseMatrix& A, const std::string& path) {
  int nr = 0, nc = 0, sym = 0;
  int64_t nnz = 0;
  if (symla_read_matrix_market(path.c_str(), &nr, &nc, &nnz, &sym, nullptr, nullptr, nullptr) != SYMLA_OK)
    return false;
  std::vector<int64_t> colptr(static_cast<size_t>(nc) + 1);
  std::vector<int> rowidx(static_cast<size_t>(nnz));
  std::vector<double> values(static_cast<size_t>(nnz));
  if (symla_read_matrix_market(path.c_str(), nullptr, nullptr, nullptr, nullptr, colptr.data(), rowidx.data(),
                               values.data()) != SYMLA_OK)
    return false;
#ifdef SYMLA_FACADE_EIGEN
  std::vector<Eigen::Triplet<double, int>> t;
  t.reserve(static_cast<size_t>(nnz));
  for (int c = 0; c < nc; ++c)
    for (int64_t k = colptr[c]; k < colptr[c + 1]; ++k) t.emplace_back(rowidx[k], c, values[k]);
  A.resize(nr, nc);
  A.setFromTriplets(t.begin(), t.end());
#else
  A = CscMatrix(nr, nc);
  A.colptr = std::move(colptr);
  A.rowidx = std::move(rowidx);
  A.values = std::move(values);
#endif
  return true;
}

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
    return last_info_.precision_used == SYMLA_PREC_FP32    ? Precision::FP32
           : last_info_.precision_used == SYMLA_PREC_BF16X3 ? Precision::BF16x3
                                                             : Precision::FP64;
  }

  void analyzePattern(const SparseMatrix& A, const SymbolicFactorOptions& options = SymbolicFactorOptions()) {
    if (A.rows() != A.cols()) throw std::invalid_argument("symla::computeOrdering: matrix must be square");
    release();
    detail::CscArrays a;
    detail::cscArrays(A, a);
    symla_symbolic* sym = nullptr;
    int ord = SYMLA_ORDER_AMD;
    if (ordering_ == OrderingType::Metis) ord = SYMLA_ORDER_METIS;
    if (ordering_ == OrderingType::Natural) ord = SYMLA_ORDER_NATURAL;
    detail::check(symla_analyze(a.n, a.colptr.data(), a.rowidx, ord, nullptr, options.max_relax_size,
                                options.max_relax_fill_fraction, &sym),
                  symla_last_error());
    sym_ = sym;
    detail::check(symla_plan_create(sym_, a.colptr.data(), a.rowidx, &plan_), symla_last_error());
    n_ = a.n;
    nnz_ = a.nnz;
    pattern_analyzed_ = true;
    factorized_ = false;
    symbolic_valid_ = false;
    numeric_valid_ = false;
  }

  void factorize(const SparseMatrix& A) {
    if (!pattern_analyzed_) throw std::logic_error("symla::SymLDLT::factorize: analyzePattern must be called first");
    detail::CscArrays a;
    detail::cscArrays(A, a);
    if (a.n != n_ || A.rows() != n_ || a.nnz != nnz_)
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
    o.precision = precision_ == Precision::FP32    ? SYMLA_PREC_FP32
                  : precision_ == Precision::BF16x3 ? SYMLA_PREC_BF16X3
                                                    : SYMLA_PREC_FP64;
    std::vector<int> signs;
    if (mode_ == Mode::StaticRegularized) {
      signs = buildExpectedSignOriginal(n_);
      if (use_kkt_blocks_ || use_explicit_signs_) o.expected_sign = signs.data();
    }
    factorized_ = false;
    numeric_valid_ = false;
    symla_factor_info info{};
    detail::check(symla_factorize(plan_, a.values, &o, &info), symla_plan_last_error(plan_));
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
    detail::check(symla_solve(plan_, B.data(), n_, static_cast<int>(B.cols()), X.data(), n_),
                  symla_plan_last_error(plan_));
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
    detail::CscArrays a;
    detail::cscArrays(A, a);
    if (a.nnz != nnz_)
      throw std::invalid_argument("symla::SymLDLT::solveWithRefinement: matrix pattern differs from the analyzed one");
    symla_refine_opts ro{options.max_iterations, options.tolerance};
    symla_refine_info ri{};
    detail::check(symla_solve_refined(plan_, a.values, B.data(), n_, static_cast<int>(B.cols()), out.x.data(), n_,
                                      &ro, &ri),
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

  // solver.hpp:261-262.  The symbolic factor is copied from the analysis on
  // first use; the numeric factor is copied from the device on first use
  // after each factorization (a test and inspection view: it moves every
  // front's L and D to the host).
  const SymbolicFactor& symbolicFactor() const {
    if (!pattern_analyzed_) throw std::logic_error("symla::SymLDLT::symbolicFactor: analyzePattern must be called first");
    if (!symbolic_valid_) loadSymbolic();
    return symbolic_;
  }
  const NumericFactor& numericFactor() const {
    if (!factorized_) throw std::logic_error("symla::SymLDLT::numericFactor: factorize() must be called first");
    if (!numeric_valid_) loadNumeric();
    return numeric_;
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
  mutable bool symbolic_valid_ = false;
  mutable bool numeric_valid_ = false;
  mutable SymbolicFactor symbolic_;
  mutable NumericFactor numeric_;

  void release() {
    if (plan_) symla_plan_destroy(plan_);
    if (sym_) symla_symbolic_free(sym_);
    plan_ = nullptr;
    sym_ = nullptr;
    symbolic_valid_ = false;
    numeric_valid_ = false;
  }

  void loadSymbolic() const {
    symla_symbolic_info_t info{};
    detail::check(symla_symbolic_info(sym_, &info), symla_last_error());
    std::vector<int> perm(info.n), parent(info.n), counts(info.n), first(info.nsn), ncols(info.nsn);
    std::vector<int64_t> rowptr(info.nsn + 1);
    std::vector<int> rows(static_cast<size_t>(info.pattern_total));
    detail::check(symla_symbolic_get(sym_, perm.data(), parent.data(), first.data(), ncols.data(), rowptr.data(),
                                     rows.data()),
                  symla_last_error());
    detail::check(symla_symbolic_colcounts(sym_, counts.data()), symla_last_error());
    SymbolicFactor sf;
    detail::toVector(perm, sf.perm);
    detail::toVector(parent, sf.etree.parent);
    detail::toVector(counts, sf.etree.colCount);
    sf.etree.n = info.n;
    sf.supernodes.resize(info.nsn);
    for (int s = 0; s < info.nsn; ++s) {
      sf.supernodes[s].firstCol = first[s];
      sf.supernodes[s].ncols = ncols[s];
      sf.supernodes[s].rowPattern.assign(rows.begin() + rowptr[s], rows.begin() + rowptr[s + 1]);
    }
    symbolic_ = std::move(sf);
    symbolic_valid_ = true;
  }

  void loadNumeric() const {
    const SymbolicFactor& sf = symbolicFactor();
    const int nsn = static_cast<int>(sf.supernodes.size());
    NumericFactor nf;
    nf.n = n_;
    nf.perm = sf.perm;
    nf.inertia = inertia_;
    nf.singular = singular_;
    nf.totalPerturbation = last_info_.total_perturbation;
    nf.numPerturbed = last_info_.num_perturbed;
    // singular columns in final order (the C-ABI reports original indices)
    std::vector<int> inv(n_);
    for (int k = 0; k < n_; ++k) inv[sf.perm[k]] = k;
    for (int c : singularColumns()) nf.singularCols.push_back(inv[c]);
    std::sort(nf.singularCols.begin(), nf.singularCols.end());
    // supernode tree from the elimination tree (multifrontal.hpp:312-332)
    std::vector<int> colToSN(n_, -1);
    for (int s = 0; s < nsn; ++s)
      for (int c = sf.supernodes[s].firstCol; c < sf.supernodes[s].firstCol + sf.supernodes[s].ncols; ++c)
        colToSN[c] = s;
    nf.parentSN.assign(nsn, -1);
    nf.childrenSN.assign(nsn, {});
    for (int s = 0; s < nsn; ++s) {
      const int p = sf.etree.parent[sf.supernodes[s].firstCol + sf.supernodes[s].ncols - 1];
      if (p >= 0) {
        nf.parentSN[s] = colToSN[p];
        nf.childrenSN[colToSN[p]].push_back(s);
      } else {
        nf.rootsSN.push_back(s);
      }
    }
    nf.fronts.resize(nsn);
    for (int s = 0; s < nsn; ++s) {
      int m = 0, k = 0, nb = 0;
      detail::check(symla_front_view(plan_, s, &m, &k, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr),
                    symla_plan_last_error(plan_));
      SupernodeFactor& f = nf.fronts[s];
      f.rowIndices.resize(m);
      f.nPivots = k;
      f.L = DenseMatrix(m, k);
      std::vector<double> dd(k), dof(k);
      std::vector<int> ps(k), pk(k);
      detail::check(symla_front_view(plan_, s, &m, &k, f.rowIndices.data(), f.L.data(), dd.data(), dof.data(),
                                     ps.data(), pk.data(), &nb),
                    symla_plan_last_error(plan_));
      f.D.diag.resize(k);
      f.D.offdiag.resize(k);
      for (int q = 0; q < k; ++q) {
        f.D.diag[q] = dd[q];
        f.D.offdiag[q] = dof[q];
      }
      f.pivotBlocks.resize(nb);
      for (int b = 0; b < nb; ++b) f.pivotBlocks[b] = PivotBlock{ps[b], pk[b] == 2 ? PivotKind::TwoByTwo : PivotKind::OneByOne};
      f.Ltop = DenseMatrix(k, k);
      for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) f.Ltop(i, j) = f.L(i, j);
      for (const PivotBlock& b : f.pivotBlocks)
        if (b.kind == PivotKind::TwoByTwo) f.Ltop(b.start + 1, b.start) = 0.0;
    }
    numeric_ = std::move(nf);
    numeric_valid_ = true;
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
