// symla_oracle.cpp -- CPU restatement of the reference numeric path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
// timed CPU baseline.  Nothing in the product links or calls this file.
//
// Each routine follows one reference function, cited by file:line (paths are
// relative to the reference repository root).  The restatement keeps the
// reference's data layout and arithmetic order where it matters for pivot
// decisions: full-square dense fronts with both triangles, lazy per-column
// panel corrections recomputed from a per-panel raw snapshot, full-square
// trailing GEMM followed by tiled re-symmetrization, first-failure delay.
// Eigen is not available here, so the dense products are written out; the
// GEMM micro-kernel uses fused multiply-adds like Eigen's vectorized GEMM
// does, and everything else is compiled with -ffp-contract=off like the
// reference build (CMakeLists.txt:99-121).

#include "symla_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#if defined(__FMA__) && defined(__AVX2__)
#include <immintrin.h>
#define ORACLE_HAVE_AVX2 1
#endif

namespace oracle {

// ---------------------------------------------------------------------------
// Small column-major dense matrix (stands in for Eigen::MatrixXd).
// ---------------------------------------------------------------------------
struct Mat {
  int r = 0, c = 0;
  std::vector<double> v;
  Mat() = default;
  Mat(int rows, int cols) : r(rows), c(cols), v(static_cast<size_t>(rows) * cols, 0.0) {}
  double& operator()(int i, int j) { return v[i + static_cast<size_t>(j) * r]; }
  double operator()(int i, int j) const { return v[i + static_cast<size_t>(j) * r]; }
  double* col(int j) { return v.data() + static_cast<size_t>(j) * r; }
  const double* col(int j) const { return v.data() + static_cast<size_t>(j) * r; }
};

// View onto caller storage with a leading dimension.
struct View {
  double* p;
  int n;    // rows == cols
  int ld;
  double& operator()(int i, int j) { return p[i + static_cast<size_t>(j) * ld]; }
  double operator()(int i, int j) const { return p[i + static_cast<size_t>(j) * ld]; }
  double* col(int j) { return p + static_cast<size_t>(j) * ld; }
};

struct Inertia {
  int n_pos = 0, n_neg = 0, n_zero = 0;
};

// BlockDiagonalD (dense_kernel.hpp:137-199): diagonal plus one d21 per 2x2.
struct BlockD {
  std::vector<double> diag, off;
  void resize(int n) {
    diag.assign(n, 0.0);
    off.assign(n, 0.0);
  }
};

enum Kind { OneByOne = 1, TwoByTwo = 2 };
struct PivotBlock {
  int start;
  int kind;
};

struct DenseResult {
  int n_factored = 0;
  std::vector<int> delayed_cols;
  std::vector<PivotBlock> pivots;
  std::vector<int> perm;
  Inertia inertia;
  double max_growth = 1.0;
  double total_perturbation = 0.0;
  int num_perturbed = 0;
};

struct DenseOpts {
  double pivot_threshold = 0.01;
  double zero_tolerance = 1e-300;
  double relative_pivot_floor = 1e-12;
  int panel_size = 64;
};

struct StaticOpts {
  double delta = 0.0;
  double relative_pivot_floor = 1e-8;
  double absolute_floor = 1e-300;
  int panel_size = 64;
};

// ---------------------------------------------------------------------------
// Dense products used by the trailing updates.
// ---------------------------------------------------------------------------

// LD = Lblk * dense(Dbuf.head(p)) where Lblk is rows x p taken from Lbuf at
// row offset `r0`.  Eigen evaluates `Lpanel * Dpanel` into a temporary first
// (dense_kernel.hpp:538, :817); the dense D is block diagonal so each output
// column sums the (at most two) nonzero D entries of that column in
// increasing row order.
static void formLD(const Mat& Lbuf, int r0, int rows, int p, const BlockD& Dbuf, Mat& LD) {
  LD = Mat(rows, p);
  for (int q = 0; q < p; ++q) {
    // column q of dense D: D(q-1,q) = off[q-1] (if a 2x2 starts at q-1),
    // D(q,q) = diag[q], D(q+1,q) = off[q] (if a 2x2 starts at q)
    const double dm = (q > 0) ? Dbuf.off[q - 1] : 0.0;
    const double dd = Dbuf.diag[q];
    const double dp = (q + 1 < p) ? Dbuf.off[q] : 0.0;
    double* out = LD.col(q);
    const double* lm = (q > 0) ? Lbuf.col(q - 1) + r0 : nullptr;
    const double* l0 = Lbuf.col(q) + r0;
    const double* lp = (q + 1 < p) ? Lbuf.col(q + 1) + r0 : nullptr;
    for (int i = 0; i < rows; ++i) {
      double s = 0.0;
      if (lm && dm != 0.0) s = s + lm[i] * dm;
      s = s + l0[i] * dd;
      if (lp && dp != 0.0) s = s + lp[i] * dp;
      out[i] = s;
    }
  }
}

// T(0:n,0:n) -= LD * Lblk^T over the FULL square (both triangles), as the
// reference's Eigen GEMM does (dense_kernel.hpp:817, :1050).  LD is n x p;
// Lblk(i,q) = Lbuf(r0+i, q).  Blocked over columns; the inner axpy uses FMA
// (mirrors Eigen's gebp micro-kernel).
static void gemmFullSub(View T, int t0, int n, const Mat& LD, const Mat& Lsrc, int r0, int p) {
  if (n <= 0 || p <= 0) return;
  constexpr int JB = 64, IB = 256;
  for (int jb = 0; jb < n; jb += JB) {
    const int jn = std::min(JB, n - jb);
    for (int ib = 0; ib < n; ib += IB) {
      const int in = std::min(IB, n - ib);
      for (int j = jb; j < jb + jn; ++j) {
        double* tc = T.col(t0 + j) + t0 + ib;
        for (int q = 0; q < p; ++q) {
          const double s = Lsrc(r0 + j, q);
          if (s == 0.0) continue;
          const double* a = LD.col(q) + ib;
          int i = 0;
#ifdef ORACLE_HAVE_AVX2
          const __m256d vs = _mm256_set1_pd(-s);
          for (; i + 4 <= in; i += 4) {
            __m256d t = _mm256_loadu_pd(tc + i);
            t = _mm256_fmadd_pd(_mm256_loadu_pd(a + i), vs, t);
            _mm256_storeu_pd(tc + i, t);
          }
#endif
          for (; i < in; ++i) tc[i] = std::fma(a[i], -s, tc[i]);
        }
      }
    }
  }
}

// resymmetrizeTrailingTiled (dense_kernel.hpp:269-297): per-entry average of
// both triangles, 64x64 tiles.
static void resymmetrize(View T, int t0, int extent) {
  if (extent == 0) return;
  constexpr int kTile = 64;
  auto at = [&](int i, int j) -> double& { return T(t0 + i, t0 + j); };
  for (int jb = 0; jb < extent; jb += kTile) {
    const int jn = std::min(kTile, extent - jb);
    for (int ib = jb; ib < extent; ib += kTile) {
      const int in = std::min(kTile, extent - ib);
      if (ib == jb) {
        for (int j = jb; j < jb + jn; ++j)
          for (int i = j + 1; i < jb + jn; ++i) {
            const double avg = 0.5 * (at(i, j) + at(j, i));
            at(i, j) = avg;
            at(j, i) = avg;
          }
      } else {
        for (int j = jb; j < jb + jn; ++j)
          for (int i = ib; i < ib + in; ++i) {
            const double avg = 0.5 * (at(i, j) + at(j, i));
            at(i, j) = avg;
            at(j, i) = avg;
          }
      }
    }
  }
}

static double maxAbsBlock(View T, int t0, int extent) {
  double m = 0.0;
  for (int j = 0; j < extent; ++j) {
    const double* c = T.col(t0 + j) + t0;
    for (int i = 0; i < extent; ++i) m = std::max(m, std::abs(c[i]));
  }
  return m;
}

// ---------------------------------------------------------------------------
// DenseLDLT::factor (dense_kernel.hpp:393-837): panel-blocked threshold
// Bunch-Kaufman with 1x1/2x2 pivots restricted to the eligible block,
// first-failure delay.
// ---------------------------------------------------------------------------
static DenseResult denseFactor(View A, BlockD& D_out, const DenseOpts& options, int n_eligible) {
  const int n = A.n;
  const int effEnd = (n_eligible < 0) ? n : n_eligible;
  if (effEnd < 0 || effEnd > n) throw std::invalid_argument("DenseLDLT::factor: n_eligible out of range");

  const double alpha = (1.0 + std::sqrt(17.0)) / 8.0;  // :400
  const int panelSizeOpt = options.panel_size > 0 ? options.panel_size : 64;

  DenseResult result;
  result.perm.resize(n);
  for (int i = 0; i < n; ++i) result.perm[i] = i;
  D_out.resize(n);

  // mirror lower -> upper (:410-411)
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) A(j, i) = A(i, j);

  double orig_max = 0.0;  // :413-415
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) orig_max = std::max(orig_max, std::abs(A(i, j)));
  if (orig_max < options.zero_tolerance) orig_max = 1.0;
  double running_max = orig_max;

  auto swap_rowcol = [&](int i, int j) {  // :458-463
    if (i == j) return;
    for (int c = 0; c < n; ++c) std::swap(A(i, c), A(j, c));
    for (int r = 0; r < n; ++r) std::swap(A(r, i), A(r, j));
    std::swap(result.perm[i], result.perm[j]);
  };

  int k = 0;
  while (k < effEnd) {
    const int panelStart = k;
    const int panelWidthTarget = std::min(panelSizeOpt, effEnd - panelStart);
    const int panelEnd = panelStart + panelWidthTarget;
    const int panelCapacity = panelWidthTarget + 1;  // :474
    const int m0 = n - panelStart;

    Mat Lbuf(m0, panelCapacity);  // :477
    BlockD Dbuf;
    Dbuf.resize(panelCapacity);
    Mat Araw(m0, m0);  // :489
    for (int j = 0; j < m0; ++j)
      for (int i = 0; i < m0; ++i) Araw(i, j) = A(panelStart + i, panelStart + j);
    std::vector<int> appliedCount(m0, 0);
    bool panelStoppedEarly = false;

    // setColumnRaw (:516-545)
    auto setColumnRaw = [&](int col, int p) {
      const int idx = col - panelStart;
      const int rowsFrom = k - panelStart;
      const int rowsCount = m0 - rowsFrom;
      if (rowsCount <= 0) {
        appliedCount[idx] = p;
        return;
      }
      if (p == 0) {
        for (int t = 0; t < rowsCount; ++t) A(k + t, col) = Araw(rowsFrom + t, idx);
      } else {
        Mat LD;
        formLD(Lbuf, rowsFrom, rowsCount, p, Dbuf, LD);
        // corr = LD * Lrow (GEMV, FMA accumulation as Eigen's GEMV kernel)
        std::vector<double> corr(rowsCount, 0.0);
        for (int q = 0; q < p; ++q) {
          const double s = Lbuf(idx, q);
          const double* a = LD.col(q);
          for (int t = 0; t < rowsCount; ++t) corr[t] = std::fma(a[t], s, corr[t]);
        }
        for (int t = 0; t < rowsCount; ++t) A(k + t, col) = Araw(rowsFrom + t, idx) - corr[t];
      }
      for (int t = 0; t < rowsCount; ++t) A(col, k + t) = A(k + t, col);
      appliedCount[idx] = p;
    };
    auto ensureColumnCurrent = [&](int col, int p) {  // :555-559
      const int idx = col - panelStart;
      if (appliedCount[idx] == p) return;
      setColumnRaw(col, p);
    };
    auto panelSwap = [&](int i, int j) {  // :565-573
      if (i == j) return;
      swap_rowcol(i, j);
      const int ii = i - panelStart, jj = j - panelStart;
      for (int c = 0; c < panelCapacity; ++c) std::swap(Lbuf(ii, c), Lbuf(jj, c));
      for (int c = 0; c < m0; ++c) std::swap(Araw(ii, c), Araw(jj, c));
      for (int r = 0; r < m0; ++r) std::swap(Araw(r, ii), Araw(r, jj));
      std::swap(appliedCount[ii], appliedCount[jj]);
    };

    int p = 0;
    while (k < panelEnd) {
      const int m = n - k;
      ensureColumnCurrent(k, p);  // :583

      double lambda = 0.0;  // :588-596
      int r = -1;
      for (int i = k + 1; i < effEnd; ++i) {
        const double v = std::abs(A(i, k));
        if (v > lambda) {
          lambda = v;
          r = i;
        }
      }

      bool accept_1x1_at_k = false;
      bool accept_1x1_swap_kr = false;
      if (lambda <= options.zero_tolerance) {  // :601-616
        const double isolatedFloor = std::max(options.zero_tolerance, options.relative_pivot_floor * orig_max);
        if (std::abs(A(k, k)) <= isolatedFloor) {
          panelStoppedEarly = true;
          break;
        }
        accept_1x1_at_k = true;
      } else if (std::abs(A(k, k)) >= alpha * lambda) {  // :617
        accept_1x1_at_k = true;
      } else {
        ensureColumnCurrent(r, p);  // :624
        double sigma = 0.0;         // :628-633
        for (int i = k; i < effEnd; ++i) {
          if (i == r) continue;
          const double v = std::abs(A(i, r));
          if (v > sigma) sigma = v;
        }
        if (std::abs(A(k, k)) * sigma >= alpha * lambda * lambda) {  // :641-645
          accept_1x1_at_k = true;
        } else if (std::abs(A(r, r)) >= alpha * sigma) {
          accept_1x1_swap_kr = true;
        }
      }

      if (accept_1x1_at_k || accept_1x1_swap_kr) {  // :665-714
        if (accept_1x1_swap_kr) panelSwap(k, r);
        running_max = std::max(running_max, std::abs(A(k, k)));
        const double d = A(k, k);
        D_out.diag[k] = d;
        Dbuf.diag[p] = d;
        result.pivots.push_back({k, OneByOne});
        if (d > 0)
          ++result.inertia.n_pos;
        else if (d < 0)
          ++result.inertia.n_neg;
        else
          ++result.inertia.n_zero;
        if (m > 1) {
          double cm = 0.0;
          for (int i = k + 1; i < n; ++i) cm = std::max(cm, std::abs(A(i, k)));
          running_max = std::max(running_max, cm);
          for (int i = k + 1; i < n; ++i) {
            const double l = A(i, k) / d;
            Lbuf(i - panelStart, p) = l;
            A(i, k) = l;
            A(k, i) = l;
          }
        }
        ++k;
        ++p;
        continue;
      }

      // 2x2 pivot at (k, r) (:721-792)
      if (r != k + 1) panelSwap(k + 1, r);
      const double d11 = A(k, k);
      const double d21 = A(k + 1, k);
      const double d22 = A(k + 1, k + 1);
      const double det = d11 * d22 - d21 * d21;
      if (std::abs(det) <=
          options.zero_tolerance * std::max(1.0, running_max) * std::max(1.0, running_max)) {  // :728
        panelStoppedEarly = true;
        break;
      }
      running_max = std::max(running_max, std::max({std::abs(d11), std::abs(d21), std::abs(d22)}));
      D_out.diag[k] = d11;
      D_out.off[k] = d21;
      D_out.diag[k + 1] = d22;
      Dbuf.diag[p] = d11;
      Dbuf.off[p] = d21;
      Dbuf.diag[p + 1] = d22;
      result.pivots.push_back({k, TwoByTwo});
      if (det < 0) {  // :753-768
        ++result.inertia.n_pos;
        ++result.inertia.n_neg;
      } else {
        const double tr = d11 + d22;
        if (tr > 0) {
          result.inertia.n_pos += 2;
        } else if (tr < 0) {
          result.inertia.n_neg += 2;
        } else {
          ++result.inertia.n_pos;
          ++result.inertia.n_neg;
        }
      }
      const int mm = n - (k + 2);
      if (mm > 0) {  // :771-788
        const double inv11 = d22 / det;
        const double inv21 = -d21 / det;
        const double inv22 = d11 / det;
        double cm = 0.0;
        for (int i = k + 2; i < n; ++i) cm = std::max(cm, std::max(std::abs(A(i, k)), std::abs(A(i, k + 1))));
        running_max = std::max(running_max, cm);
        for (int i = k + 2; i < n; ++i) {
          const double a0 = A(i, k), a1 = A(i, k + 1);
          const double l0 = a0 * inv11 + a1 * inv21;
          const double l1 = a0 * inv21 + a1 * inv22;
          Lbuf(i - panelStart, p) = l0;
          Lbuf(i - panelStart, p + 1) = l1;
          A(i, k) = l0;
          A(i, k + 1) = l1;
          A(k, i) = l0;
          A(k + 1, i) = l1;
        }
      }
      k += 2;
      p += 2;
    }

    // end of panel (:806-821)
    if (p > 0) {
      for (int col = k; col < n; ++col) {
        const int idx = col - panelStart;
        if (appliedCount[idx] != 0) setColumnRaw(col, 0);
      }
      const int rowsCount = n - k;
      if (rowsCount > 0) {
        Mat LD;
        formLD(Lbuf, k - panelStart, rowsCount, p, Dbuf, LD);
        gemmFullSub(A, k, rowsCount, LD, Lbuf, k - panelStart, p);
        resymmetrize(A, k, rowsCount);
        running_max = std::max(running_max, maxAbsBlock(A, k, rowsCount));
      }
    }
    if (panelStoppedEarly) break;
  }

  result.n_factored = k;  // :829-835
  for (int j = k; j < effEnd; ++j) result.delayed_cols.push_back(j);
  result.max_growth = running_max / orig_max;
  return result;
}

// ---------------------------------------------------------------------------
// DenseLDLT::factorStatic (dense_kernel.hpp:919-1059): fixed-order 1x1
// pivots with sign-matched perturbation.
// ---------------------------------------------------------------------------
static DenseResult denseFactorStatic(View A, BlockD& D_out, const StaticOpts& options,
                                     const std::vector<int>& expectedSign, int n_eligible) {
  const int n = A.n;
  const int effEnd = (n_eligible < 0) ? n : n_eligible;
  if (effEnd < 0 || effEnd > n) throw std::invalid_argument("DenseLDLT::factorStatic: n_eligible out of range");

  DenseResult result;
  result.perm.resize(n);
  for (int i = 0; i < n; ++i) result.perm[i] = i;
  D_out.resize(n);
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) A(j, i) = A(i, j);

  double orig_max = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) orig_max = std::max(orig_max, std::abs(A(i, j)));
  if (orig_max < options.absolute_floor) orig_max = 1.0;
  double running_max = orig_max;

  const double delta =
      options.delta > 0.0 ? options.delta : std::sqrt(std::numeric_limits<double>::epsilon()) * orig_max;  // :938
  const double floor = std::max(options.absolute_floor, options.relative_pivot_floor * orig_max);      // :939
  const int panelSizeOpt = options.panel_size > 0 ? options.panel_size : 64;

  int k = 0;
  while (k < effEnd) {
    const int panelStart = k;
    const int panelWidth = std::min(panelSizeOpt, effEnd - panelStart);
    const int panelEnd = panelStart + panelWidth;
    const int m0 = n - panelStart;
    Mat Araw(m0, panelWidth);  // :957
    for (int j = 0; j < panelWidth; ++j)
      for (int i = 0; i < m0; ++i) Araw(i, j) = A(panelStart + i, panelStart + j);
    Mat Lbuf(m0, panelWidth);
    std::vector<double> Dbuf(panelWidth, 0.0);

    for (; k < panelEnd; ++k) {
      const int idx = k - panelStart;
      const int rowsCount = m0 - idx;
      std::vector<double> col(rowsCount);
      if (idx == 0) {  // :974-980
        for (int t = 0; t < rowsCount; ++t) col[t] = Araw(t, 0);
      } else {
        std::vector<double> Dl(idx);
        for (int q = 0; q < idx; ++q) Dl[q] = Lbuf(idx, q) * Dbuf[q];
        std::vector<double> prod(rowsCount, 0.0);
        for (int q = 0; q < idx; ++q) {
          const double s = Dl[q];
          for (int t = 0; t < rowsCount; ++t) prod[t] = std::fma(Lbuf(idx + t, q), s, prod[t]);
        }
        for (int t = 0; t < rowsCount; ++t) col[t] = Araw(idx + t, idx) - prod[t];
      }
      const double a_kk = col[0];
      const int sign = (k < static_cast<int>(expectedSign.size())) ? expectedSign[k] : 0;
      double d = a_kk;
      double applied = 0.0;
      bool needPerturb;
      if (sign > 0)
        needPerturb = a_kk < floor;
      else if (sign < 0)
        needPerturb = a_kk > -floor;
      else
        needPerturb = std::abs(a_kk) < floor;
      if (needPerturb) {  // :996-1011
        const double s = (sign != 0) ? static_cast<double>(sign) : (a_kk >= 0.0 ? 1.0 : -1.0);
        d = a_kk + s * delta;
        applied = d - a_kk;
      }
      D_out.diag[k] = d;
      Dbuf[idx] = d;
      result.pivots.push_back({k, OneByOne});
      if (applied != 0.0) {
        result.total_perturbation += std::abs(applied);
        ++result.num_perturbed;
      }
      if (d > 0)
        ++result.inertia.n_pos;
      else if (d < 0)
        ++result.inertia.n_neg;
      else
        ++result.inertia.n_zero;
      running_max = std::max(running_max, std::abs(d));
      if (rowsCount > 1) {
        double cm = 0.0;
        for (int t = 1; t < rowsCount; ++t) cm = std::max(cm, std::abs(col[t]));
        running_max = std::max(running_max, cm);
        for (int t = 1; t < rowsCount; ++t) {
          const double l = col[t] / d;
          Lbuf(idx + t, idx) = l;
          A(k + t, k) = l;
          A(k, k + t) = l;
        }
      }
    }

    const int rowsBeyond = n - panelEnd;  // :1045-1053
    if (rowsBeyond > 0) {
      const int r0 = panelEnd - panelStart;
      Mat LD(rowsBeyond, panelWidth);
      for (int q = 0; q < panelWidth; ++q)
        for (int i = 0; i < rowsBeyond; ++i) LD(i, q) = Lbuf(r0 + i, q) * Dbuf[q];
      gemmFullSub(A, panelEnd, rowsBeyond, LD, Lbuf, r0, panelWidth);
      resymmetrize(A, panelEnd, rowsBeyond);
      running_max = std::max(running_max, maxAbsBlock(A, panelEnd, rowsBeyond));
    }
  }
  result.n_factored = effEnd;
  result.max_growth = running_max / orig_max;
  return result;
}

// ---------------------------------------------------------------------------
// Multifrontal driver (multifrontal.hpp:118-186, :199-242, :299-805).
// ---------------------------------------------------------------------------
struct SupernodeFactor {  // multifrontal.hpp:118-135
  std::vector<int> rowIndices;
  int nPivots = 0;
  Mat L;  // m x nPivots
  BlockD D;
  std::vector<PivotBlock> pivotBlocks;
  Mat Ltop;  // nPivots x nPivots, 2x2 d21 slots zeroed
};

struct NumericFactor {  // multifrontal.hpp:148-186
  std::vector<SupernodeFactor> fronts;
  std::vector<int> perm;
  Inertia inertia;
  bool singular = false;
  std::vector<int> singularCols;
  int n = 0;
  double totalPerturbation = 0.0;
  int numPerturbed = 0;
  std::vector<int> parentSN;
  std::vector<std::vector<int>> childrenSN;
  std::vector<int> rootsSN;
};

struct Symbolic {
  int n = 0;
  std::vector<int> perm, etreeParent;
  struct SN {
    int firstCol, ncols;
    std::vector<int> rowPattern;
  };
  std::vector<SN> supernodes;
};

struct Csc {
  int n;
  const long long* colptr;
  const int* rowidx;
  const double* vals;
};

struct MfOpts {
  bool parallel = false;
  int num_threads = 0;
  long long task_cutoff = -1;
  bool static_pivoting = false;
  DenseOpts dense;
  StaticOpts stat;
  std::vector<int> expectedSign;  // final order
  int snLimit = 0;                // > 0: supernodes [0, snLimit) only (timing sample)
};

// parallel::runTaskDag (task_graph.hpp:48-89)
template <typename ProcessFn>
static void runTaskDag(const std::vector<int>& roots, const std::vector<std::vector<int>>& children,
                       const std::vector<long long>& subtreeWeight, long long taskCutoff, ProcessFn&& process) {
#ifdef _OPENMP
  std::function<void(int)> visit = [&](int node) {
    bool spawnedAny = false;
    for (int c : children[node]) {
      if (subtreeWeight[c] >= taskCutoff) {
        spawnedAny = true;
#pragma omp task default(shared) firstprivate(c)
        visit(c);
      } else {
        visit(c);
      }
    }
    if (spawnedAny) {
#pragma omp taskwait
    }
    process(node);
  };
  for (int r : roots) {
#pragma omp task default(shared) firstprivate(r)
    visit(r);
  }
#pragma omp taskwait
#else
  std::function<void(int)> visit = [&](int node) {
    for (int c : children[node]) visit(c);
    process(node);
  };
  for (int r : roots) visit(r);
#endif
}

// parallel::runTaskDagTopDown (task_graph.hpp:108-153)
template <typename ProcessFn>
static void runTaskDagTopDown(const std::vector<int>& roots, const std::vector<std::vector<int>>& children,
                              const std::vector<long long>& subtreeWeight, long long taskCutoff,
                              ProcessFn&& process) {
#ifdef _OPENMP
  std::function<void(int)> visit = [&](int node) {
    process(node);
    for (int c : children[node]) {
      if (subtreeWeight[c] >= taskCutoff) {
#pragma omp task default(shared) firstprivate(c)
        visit(c);
      } else {
        visit(c);
      }
    }
  };
#pragma omp taskgroup
  {
    for (int r : roots) {
#pragma omp task default(shared) firstprivate(r)
      visit(r);
    }
  }
#else
  std::function<void(int)> visit = [&](int node) {
    process(node);
    for (int c : children[node]) visit(c);
  };
  for (int r : roots) visit(r);
#endif
}

static int maxThreadsAvailable() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static NumericFactor mfFactorize(const Symbolic& sf, const Csc& A, const MfOpts& mfOptions) {
  const int n = sf.n;
  NumericFactor nf;
  nf.perm = sf.perm;
  nf.n = n;
  if (n == 0) return nf;
  const auto& supernodes = sf.supernodes;
  const int numSNAll = static_cast<int>(supernodes.size());
  // a timing sample stops after the first snLimit supernodes in postorder;
  // every loop below runs over [0, numSN)
  const int numSN = (mfOptions.snLimit > 0 && mfOptions.snLimit < numSNAll) ? mfOptions.snLimit : numSNAll;

  // supernode tree (:312-332)
  std::vector<int> colToSupernode(n, -1);
  for (int si = 0; si < numSNAll; ++si)
    for (int c = supernodes[si].firstCol; c < supernodes[si].firstCol + supernodes[si].ncols; ++c)
      colToSupernode[c] = si;
  std::vector<int> parentSN(numSN, -1);
  std::vector<std::vector<int>> childrenSN(numSN);
  std::vector<int> rootsSN;
  for (int si = 0; si < numSN; ++si) {
    const int lastCol = supernodes[si].firstCol + supernodes[si].ncols - 1;
    const int p = sf.etreeParent[lastCol];
    if (p != -1) {
      const int psn = colToSupernode[p];
      if (psn >= numSN) continue;  // outside a timing sample: the front is a sample root
      parentSN[si] = psn;
      childrenSN[psn].push_back(si);
    }
  }
  for (int si = 0; si < numSN; ++si)
    if (parentSN[si] == -1) rootsSN.push_back(si);
  nf.parentSN = parentSN;
  nf.childrenSN = childrenSN;
  nf.rootsSN = rootsSN;

  // PermutedLower::build (:199-242)
  std::vector<int> newIndexOf(n);
  for (int k = 0; k < n; ++k) newIndexOf[sf.perm[k]] = k;
  std::vector<std::vector<int>> aRows(n);
  std::vector<std::vector<double>> aVals(n);
  for (int c = 0; c < n; ++c) {
    for (long long t = A.colptr[c]; t < A.colptr[c + 1]; ++t) {
      const int r = A.rowidx[t];
      if (r < c) continue;
      const int nr = newIndexOf[r], nc = newIndexOf[c];
      const int lo = std::min(nr, nc), hi = std::max(nr, nc);
      aRows[lo].push_back(hi);
      aVals[lo].push_back(A.vals[t]);
    }
  }
  for (int c = 0; c < n; ++c) {
    std::vector<int> order(aRows[c].size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = static_cast<int>(k);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return aRows[c][a] < aRows[c][b]; });
    std::vector<int> nr(order.size());
    std::vector<double> nv(order.size());
    for (size_t k = 0; k < order.size(); ++k) {
      nr[k] = aRows[c][order[k]];
      nv[k] = aVals[c][order[k]];
    }
    aRows[c].swap(nr);
    aVals[c].swap(nv);
  }

  struct GeneratedElement {  // :354-357
    std::vector<int> indices;
    Mat matrix;
  };
  std::vector<GeneratedElement> genElem(numSN);
  nf.fronts.resize(numSN);
  std::vector<Inertia> perFrontInertia(numSN);
  std::vector<double> perFrontPerturbation(numSN, 0.0);
  std::vector<int> perFrontNumPerturbed(numSN, 0);
  std::mutex singularMutex;

  const int maxThreads =
      mfOptions.parallel ? std::max(1, mfOptions.num_threads > 0 ? mfOptions.num_threads : maxThreadsAvailable()) : 1;
  std::vector<std::vector<int>> globalToLocalTls(maxThreads, std::vector<int>(n, -1));

  std::vector<long long> subtreeWeight(numSN, 0);  // :418-429
  for (int si = 0; si < numSN; ++si) {
    const long long m = static_cast<long long>(supernodes[si].rowPattern.size());
    const long long snCols = supernodes[si].ncols;
    subtreeWeight[si] += std::max<long long>(1, snCols * m * m);
    if (parentSN[si] != -1) subtreeWeight[parentSN[si]] += subtreeWeight[si];
  }
  long long totalCost = 0;
  for (int r : rootsSN) totalCost += subtreeWeight[r];

  auto processSupernode = [&](int si) {  // :436-644
#ifdef _OPENMP
    const int tid = mfOptions.parallel ? omp_get_thread_num() : 0;
#else
    const int tid = 0;
#endif
    std::vector<int>& globalToLocal = globalToLocalTls[tid];
    const auto& sn = supernodes[si];
    const int ncols = sn.ncols;

    std::vector<int> extraFromChildren;  // :453-467
    for (int ci : childrenSN[si]) {
      for (int idx : genElem[ci].indices)
        if (!std::binary_search(sn.rowPattern.begin(), sn.rowPattern.end(), idx)) extraFromChildren.push_back(idx);
    }
    std::sort(extraFromChildren.begin(), extraFromChildren.end());
    extraFromChildren.erase(std::unique(extraFromChildren.begin(), extraFromChildren.end()), extraFromChildren.end());
    const int nExtra = static_cast<int>(extraFromChildren.size());
    const int nEligible = ncols + nExtra;
    const int m = static_cast<int>(sn.rowPattern.size()) + nExtra;

    std::vector<int> frontIdx(m);  // :472-479
    for (int j = 0; j < ncols; ++j) frontIdx[j] = sn.firstCol + j;
    for (int j = 0; j < nExtra; ++j) frontIdx[ncols + j] = extraFromChildren[j];
    for (int j = ncols; j < static_cast<int>(sn.rowPattern.size()); ++j)
      frontIdx[nEligible + (j - ncols)] = sn.rowPattern[j];
    for (int t = 0; t < m; ++t) globalToLocal[frontIdx[t]] = t;

    Mat F(m, m);  // :498
    for (int j = 0; j < ncols; ++j) {  // extend (:501-509)
      const int gj = sn.firstCol + j;
      const auto& rr = aRows[gj];
      const auto& vv = aVals[gj];
      for (size_t t = 0; t < rr.size(); ++t) F(globalToLocal[rr[t]], j) += vv[t];
    }
    for (int ci : childrenSN[si]) {  // add (:518-535)
      auto& ge = genElem[ci];
      const int p = static_cast<int>(ge.indices.size());
      for (int a = 0; a < p; ++a) {
        const int la = globalToLocal[ge.indices[a]];
        for (int b = 0; b <= a; ++b) {
          const int lb = globalToLocal[ge.indices[b]];
          const double v = ge.matrix(a, b);
          if (la >= lb)
            F(la, lb) += v;
          else
            F(lb, la) += v;
        }
      }
      ge.matrix = Mat();
      ge.indices.clear();
      ge.indices.shrink_to_fit();
    }
    for (int t = 0; t < m; ++t) globalToLocal[frontIdx[t]] = -1;

    constexpr int kPanelBlockingMinFront = 256;  // :564-566
    BlockD D;
    DenseResult res;
    View Fv{F.v.data(), m, m};
    if (mfOptions.static_pivoting) {  // :569-593
      std::vector<int> localExpectedSign(m, 0);
      if (!mfOptions.expectedSign.empty()) {
        for (int t = 0; t < m; ++t) {
          const int g = frontIdx[t];
          if (g >= 0 && g < static_cast<int>(mfOptions.expectedSign.size()))
            localExpectedSign[t] = mfOptions.expectedSign[g];
        }
      }
      StaticOpts so = mfOptions.stat;
      if (nEligible < kPanelBlockingMinFront) so.panel_size = 1;
      res = denseFactorStatic(Fv, D, so, localExpectedSign, nEligible);
    } else {
      DenseOpts fo = mfOptions.dense;
      if (nEligible < kPanelBlockingMinFront) fo.panel_size = 1;
      res = denseFactor(Fv, D, fo, nEligible);
    }
    perFrontInertia[si] = res.inertia;
    perFrontPerturbation[si] = res.total_perturbation;
    perFrontNumPerturbed[si] = res.num_perturbed;

    const int nFactored = res.n_factored;
    auto globalOf = [&](int i) { return frontIdx[res.perm[i]]; };  // :612

    SupernodeFactor sfac;  // :614-625
    sfac.nPivots = nFactored;
    sfac.rowIndices.resize(m);
    for (int i = 0; i < m; ++i) sfac.rowIndices[i] = globalOf(i);
    sfac.L = Mat(m, nFactored);
    for (int j = 0; j < nFactored; ++j)
      for (int i = 0; i < m; ++i) sfac.L(i, j) = F(i, j);
    sfac.D.diag.assign(D.diag.begin(), D.diag.begin() + nFactored);
    sfac.D.off.assign(D.off.begin(), D.off.begin() + nFactored);
    sfac.pivotBlocks = res.pivots;
    sfac.Ltop = Mat(nFactored, nFactored);
    for (int j = 0; j < nFactored; ++j)
      for (int i = 0; i < nFactored; ++i) sfac.Ltop(i, j) = sfac.L(i, j);
    for (const auto& pb : sfac.pivotBlocks)
      if (pb.kind == TwoByTwo) sfac.Ltop(pb.start + 1, pb.start) = 0.0;
    nf.fronts[si] = std::move(sfac);

    const int leftover = m - nFactored;  // :629-643
    if (leftover > 0) {
      const int psn = parentSN[si];
      if (psn == -1) {
        std::lock_guard<std::mutex> lock(singularMutex);
        nf.singular = true;
        for (int i = nFactored; i < m; ++i) nf.singularCols.push_back(globalOf(i));
      } else {
        GeneratedElement ge;
        ge.indices.resize(leftover);
        for (int i = 0; i < leftover; ++i) ge.indices[i] = globalOf(nFactored + i);
        ge.matrix = Mat(leftover, leftover);
        for (int j = 0; j < leftover; ++j)
          for (int i = 0; i < leftover; ++i) ge.matrix(i, j) = F(nFactored + i, nFactored + j);
        genElem[si] = std::move(ge);
      }
    }
  };

#ifdef _OPENMP
  if (mfOptions.parallel) {  // :646-787
    int effectiveThreads = maxThreads;
    if (mfOptions.num_threads <= 0) {
      constexpr long long kCostPerThread = 8'000'000;
      effectiveThreads = static_cast<int>(
          std::min<long long>(maxThreads, std::max<long long>(1, totalCost / kCostPerThread)));
    }
    long long cutoff = mfOptions.task_cutoff;
    if (cutoff < 0) {
      constexpr long long kTargetTasksPerThread = 4;
      const long long targetTasks = std::max<long long>(1, kTargetTasksPerThread * effectiveThreads);
      cutoff = totalCost > 0 ? std::max<long long>(1, totalCost / targetTasks) : 1;
      constexpr long long kMinTaskCost = 1 << 17;
      cutoff = std::max(cutoff, kMinTaskCost);
    }
    constexpr long long kMinTotalCostForParallel = (1 << 17) * 4;
    if (totalCost < kMinTotalCostForParallel) {
      for (int si = 0; si < numSN; ++si) processSupernode(si);
    } else {
      const int savedMaxActiveLevels = omp_get_max_active_levels();
      omp_set_max_active_levels(1);
      const int requestedThreads = mfOptions.num_threads > 0 ? mfOptions.num_threads : effectiveThreads;
#pragma omp parallel num_threads(requestedThreads)
      {
#pragma omp single
        { runTaskDag(rootsSN, childrenSN, subtreeWeight, cutoff, processSupernode); }
      }
      omp_set_max_active_levels(savedMaxActiveLevels);
    }
  } else {
    for (int si = 0; si < numSN; ++si) processSupernode(si);
  }
#else
  for (int si = 0; si < numSN; ++si) processSupernode(si);
#endif

  for (int si = 0; si < numSN; ++si) {  // :795-801
    nf.inertia.n_pos += perFrontInertia[si].n_pos;
    nf.inertia.n_neg += perFrontInertia[si].n_neg;
    nf.inertia.n_zero += perFrontInertia[si].n_zero;
    nf.totalPerturbation += perFrontPerturbation[si];
    nf.numPerturbed += perFrontNumPerturbed[si];
  }
  std::sort(nf.singularCols.begin(), nf.singularCols.end());
  return nf;
}

// ---------------------------------------------------------------------------
// Solve (solve.hpp:159-508).  Dense blocks are n x nrhs column-major.
// ---------------------------------------------------------------------------

// Unit-lower in-place solve with Ltop (k x k), Y k x nrhs (ld k).
static void unitLowerSolve(const Mat& Lt, double* Y, int k, int nrhs) {
  for (int c = 0; c < nrhs; ++c) {
    double* y = Y + static_cast<size_t>(c) * k;
    for (int j = 0; j < k; ++j) {
      const double yj = y[j];
      if (yj == 0.0) continue;
      const double* l = Lt.col(j);
      for (int i = j + 1; i < k; ++i) y[i] -= l[i] * yj;
    }
  }
}
// Unit-lower transposed in-place solve (back substitution with Ltop^T).
static void unitLowerTransSolve(const Mat& Lt, double* Y, int k, int nrhs) {
  for (int c = 0; c < nrhs; ++c) {
    double* y = Y + static_cast<size_t>(c) * k;
    for (int j = k - 1; j >= 0; --j) {
      const double* l = Lt.col(j);
      double s = y[j];
      for (int i = j + 1; i < k; ++i) s -= l[i] * y[i];
      y[j] = s;
    }
  }
}

struct Rows {  // row-major-ish helper: matrix n x nrhs column-major
  double* p;
  int n, nrhs;
  double& at(int i, int c) { return p[i + static_cast<size_t>(c) * n]; }
};

static void diagonalSolve(const NumericFactor& nf, std::vector<double>& Y, std::vector<double>& Z, int nrhs) {
  const int n = nf.n;  // :434-455
  for (const auto& f : nf.fronts) {
    for (const auto& pb : f.pivotBlocks) {
      const int s = pb.start;
      if (pb.kind == OneByOne) {
        const int idx = f.rowIndices[s];
        const double d = f.D.diag[s];
        for (int c = 0; c < nrhs; ++c) Z[idx + static_cast<size_t>(c) * n] = Y[idx + static_cast<size_t>(c) * n] / d;
      } else {
        const int i0 = f.rowIndices[s], i1 = f.rowIndices[s + 1];
        const double d11 = f.D.diag[s], d21 = f.D.off[s], d22 = f.D.diag[s + 1];
        const double det = d11 * d22 - d21 * d21;
        for (int c = 0; c < nrhs; ++c) {
          const double y0 = Y[i0 + static_cast<size_t>(c) * n], y1 = Y[i1 + static_cast<size_t>(c) * n];
          Z[i0 + static_cast<size_t>(c) * n] = (d22 * y0 - d21 * y1) / det;
          Z[i1 + static_cast<size_t>(c) * n] = (-d21 * y0 + d11 * y1) / det;
        }
      }
    }
  }
}

static void forwardFront(const SupernodeFactor& f, std::vector<double>& Yw, int n, int nrhs) {
  const int m = static_cast<int>(f.rowIndices.size());
  const int k = f.nPivots;
  if (k == 0) return;
  std::vector<double> yOwn(static_cast<size_t>(k) * nrhs);
  for (int c = 0; c < nrhs; ++c)
    for (int p = 0; p < k; ++p) yOwn[p + static_cast<size_t>(c) * k] = Yw[f.rowIndices[p] + static_cast<size_t>(c) * n];
  unitLowerSolve(f.Ltop, yOwn.data(), k, nrhs);  // :469
  for (int c = 0; c < nrhs; ++c)
    for (int p = 0; p < k; ++p) Yw[f.rowIndices[p] + static_cast<size_t>(c) * n] = yOwn[p + static_cast<size_t>(c) * k];
  const int extra = m - k;
  if (extra > 0) {  // :473-477
    for (int c = 0; c < nrhs; ++c) {
      std::vector<double> upd(extra, 0.0);
      for (int q = 0; q < k; ++q) {
        const double y = yOwn[q + static_cast<size_t>(c) * k];
        const double* l = f.L.col(q) + k;
        for (int i = 0; i < extra; ++i) upd[i] = std::fma(l[i], y, upd[i]);
      }
      for (int i = 0; i < extra; ++i) Yw[f.rowIndices[k + i] + static_cast<size_t>(c) * n] -= upd[i];
    }
  }
}

static void backwardFront(const SupernodeFactor& f, const std::vector<double>& Z, std::vector<double>& Xw, int n,
                          int nrhs) {
  const int m = static_cast<int>(f.rowIndices.size());
  const int k = f.nPivots;
  if (k == 0) return;
  std::vector<double> xOwn(static_cast<size_t>(k) * nrhs);
  for (int c = 0; c < nrhs; ++c)
    for (int p = 0; p < k; ++p) xOwn[p + static_cast<size_t>(c) * k] = Z[f.rowIndices[p] + static_cast<size_t>(c) * n];
  const int extra = m - k;
  if (extra > 0) {  // :496-501
    for (int c = 0; c < nrhs; ++c) {
      for (int q = 0; q < k; ++q) {
        const double* l = f.L.col(q) + k;
        double s = 0.0;
        for (int i = 0; i < extra; ++i) s = std::fma(l[i], Xw[f.rowIndices[k + i] + static_cast<size_t>(c) * n], s);
        xOwn[q + static_cast<size_t>(c) * k] -= s;
      }
    }
  }
  unitLowerTransSolve(f.Ltop, xOwn.data(), k, nrhs);  // :503
  for (int c = 0; c < nrhs; ++c)
    for (int p = 0; p < k; ++p) Xw[f.rowIndices[p] + static_cast<size_t>(c) * n] = xOwn[p + static_cast<size_t>(c) * k];
}

static std::vector<double> mfSolve(const NumericFactor& nf, const double* B, int nrhs, bool parallel, int numThreads) {
  const int n = nf.n;
  std::vector<double> result(static_cast<size_t>(n) * nrhs);
  if (n == 0) return result;
  std::vector<double> Bp(static_cast<size_t>(n) * nrhs);  // :169-170
  for (int c = 0; c < nrhs; ++c)
    for (int t = 0; t < n; ++t) Bp[t + static_cast<size_t>(c) * n] = B[nf.perm[t] + static_cast<size_t>(c) * n];
  const int numSN = static_cast<int>(nf.fronts.size());

  // cost model and policy (:192-285)
  std::vector<long long> subtreeWeight(numSN, 0);
  for (int si = 0; si < numSN; ++si) {
    const auto& f = nf.fronts[si];
    const long long m = static_cast<long long>(f.rowIndices.size());
    const long long k = std::max<long long>(1, f.nPivots);
    subtreeWeight[si] += std::max<long long>(1, m * k) * std::max(1, nrhs);
    if (nf.parentSN[si] != -1) subtreeWeight[nf.parentSN[si]] += subtreeWeight[si];
  }
  long long totalCost = 0;
  for (int r : nf.rootsSN) totalCost += subtreeWeight[r];
  bool runParallel = false;
  long long cutoff = 0;
  int requestedThreads = 1;
#ifdef _OPENMP
  const int maxThreads = parallel ? std::max(1, numThreads > 0 ? numThreads : omp_get_max_threads()) : 1;
  if (parallel && numSN > 0) {
    constexpr long long kCostPerThread = 2'000'000;
    int effectiveThreads = maxThreads;
    if (numThreads <= 0)
      effectiveThreads =
          static_cast<int>(std::min<long long>(maxThreads, std::max<long long>(1, totalCost / kCostPerThread)));
    constexpr long long kTargetTasksPerThread = 4;
    const long long targetTasks = std::max<long long>(1, kTargetTasksPerThread * effectiveThreads);
    cutoff = totalCost > 0 ? std::max<long long>(1, totalCost / targetTasks) : 1;
    cutoff = std::max<long long>(cutoff, 1'000'000);
    long long maxSpawnable = 0;
    for (int si = 0; si < numSN; ++si)
      if (nf.parentSN[si] != -1) maxSpawnable = std::max(maxSpawnable, subtreeWeight[si]);
    if (nf.rootsSN.size() >= 2)
      for (int r : nf.rootsSN) maxSpawnable = std::max(maxSpawnable, subtreeWeight[r]);
    constexpr long long kMinTotalCostForParallel = 1'000'000 * 4;
    if (totalCost >= kMinTotalCostForParallel && maxSpawnable >= cutoff) {
      runParallel = true;
      requestedThreads = numThreads > 0 ? numThreads : effectiveThreads;
    }
  }
#else
  (void)parallel;
  (void)numThreads;
#endif

  std::vector<double> Y(static_cast<size_t>(n) * nrhs), Z(static_cast<size_t>(n) * nrhs),
      X(static_cast<size_t>(n) * nrhs);
  if (runParallel) {
#ifdef _OPENMP
    struct RhsElement {
      std::vector<int> indices;
      std::vector<double> matrix;  // p x nrhs
    };
    std::vector<RhsElement> rhsElem(numSN);
    std::vector<std::vector<int>> g2lTls(std::max(1, requestedThreads), std::vector<int>(n, -1));
    auto forwardProcess = [&](int si) {  // :303-362
      std::vector<int>& g2l = g2lTls[omp_get_thread_num()];
      const auto& f = nf.fronts[si];
      const int m = static_cast<int>(f.rowIndices.size());
      const int k = f.nPivots;
      if (m == 0) return;
      std::vector<double> G(static_cast<size_t>(m) * nrhs, 0.0);
      for (int t = 0; t < m; ++t) g2l[f.rowIndices[t]] = t;
      for (int ci : nf.childrenSN[si]) {
        RhsElement& ce = rhsElem[ci];
        const int p = static_cast<int>(ce.indices.size());
        for (int a = 0; a < p; ++a) {
          const int li = g2l[ce.indices[a]];
          for (int c = 0; c < nrhs; ++c) G[li + static_cast<size_t>(c) * m] += ce.matrix[a + static_cast<size_t>(c) * p];
        }
        ce.matrix.clear();
        ce.matrix.shrink_to_fit();
        ce.indices.clear();
        ce.indices.shrink_to_fit();
      }
      for (int t = 0; t < m; ++t) g2l[f.rowIndices[t]] = -1;
      std::vector<double> yOwn(static_cast<size_t>(k) * nrhs);
      if (k > 0) {
        for (int c = 0; c < nrhs; ++c)
          for (int p = 0; p < k; ++p)
            yOwn[p + static_cast<size_t>(c) * k] =
                Bp[f.rowIndices[p] + static_cast<size_t>(c) * n] - G[p + static_cast<size_t>(c) * m];
        unitLowerSolve(f.Ltop, yOwn.data(), k, nrhs);
        for (int c = 0; c < nrhs; ++c)
          for (int p = 0; p < k; ++p) Y[f.rowIndices[p] + static_cast<size_t>(c) * n] = yOwn[p + static_cast<size_t>(c) * k];
      }
      const int extra = m - k;
      if (extra > 0) {
        std::vector<double> ne(static_cast<size_t>(extra) * nrhs);
        for (int c = 0; c < nrhs; ++c) {
          for (int i = 0; i < extra; ++i) ne[i + static_cast<size_t>(c) * extra] = G[k + i + static_cast<size_t>(c) * m];
          if (k > 0) {
            std::vector<double> upd(extra, 0.0);
            for (int q = 0; q < k; ++q) {
              const double y = yOwn[q + static_cast<size_t>(c) * k];
              const double* l = f.L.col(q) + k;
              for (int i = 0; i < extra; ++i) upd[i] = std::fma(l[i], y, upd[i]);
            }
            for (int i = 0; i < extra; ++i) ne[i + static_cast<size_t>(c) * extra] += upd[i];
          }
        }
        rhsElem[si].indices.assign(f.rowIndices.begin() + k, f.rowIndices.end());
        rhsElem[si].matrix = std::move(ne);
      }
    };
    auto backwardProcess = [&](int si) { backwardFront(nf.fronts[si], Z, X, n, nrhs); };
    const int savedMaxActiveLevels = omp_get_max_active_levels();
    omp_set_max_active_levels(1);
#pragma omp parallel num_threads(requestedThreads)
    {
#pragma omp single
      { runTaskDag(nf.rootsSN, nf.childrenSN, subtreeWeight, cutoff, forwardProcess); }
    }
    diagonalSolve(nf, Y, Z, nrhs);
#pragma omp parallel num_threads(requestedThreads)
    {
#pragma omp single
      { runTaskDagTopDown(nf.rootsSN, nf.childrenSN, subtreeWeight, cutoff, backwardProcess); }
    }
    omp_set_max_active_levels(savedMaxActiveLevels);
#endif
  } else {
    std::vector<double> Yw = Bp;  // forwardSolveSerial (:457-481)
    for (const auto& f : nf.fronts) forwardFront(f, Yw, n, nrhs);
    Y = std::move(Yw);
    diagonalSolve(nf, Y, Z, nrhs);
    for (auto it = nf.fronts.rbegin(); it != nf.fronts.rend(); ++it) backwardFront(*it, Z, X, n, nrhs);  // :483-508
  }
  for (int c = 0; c < nrhs; ++c)  // :428-430
    for (int t = 0; t < n; ++t) result[nf.perm[t] + static_cast<size_t>(c) * n] = X[t + static_cast<size_t>(c) * n];
  return result;
}

// ---------------------------------------------------------------------------
// Iterative refinement (refine.hpp:69-144).
// ---------------------------------------------------------------------------
static double sparseMaxAbsRowSum(const Csc& A) {  // refine.hpp:69-81
  std::vector<double> rs(A.n, 0.0);
  for (int c = 0; c < A.n; ++c)
    for (long long t = A.colptr[c]; t < A.colptr[c + 1]; ++t) rs[A.rowidx[t]] += std::abs(A.vals[t]);
  double m = 0.0;
  for (double v : rs) m = std::max(m, v);
  return m;
}

// R = B - selfadjointView<Lower>(A) * X
static void residual(const Csc& A, const double* B, const double* X, int nrhs, std::vector<double>& R) {
  const int n = A.n;
  R.assign(B, B + static_cast<size_t>(n) * nrhs);
  for (int c = 0; c < nrhs; ++c) {
    const double* x = X + static_cast<size_t>(c) * n;
    double* r = R.data() + static_cast<size_t>(c) * n;
    for (int j = 0; j < n; ++j) {
      for (long long t = A.colptr[j]; t < A.colptr[j + 1]; ++t) {
        const int i = A.rowidx[t];
        if (i < j) continue;
        const double v = A.vals[t];
        r[i] -= v * x[j];
        if (i != j) r[j] -= v * x[i];
      }
    }
  }
}

static double infNorm(const double* p, size_t len) {
  double m = 0.0;
  for (size_t i = 0; i < len; ++i) m = std::max(m, std::abs(p[i]));
  return m;
}

static oracle_refine_result iterativeRefine(const NumericFactor& nf, const Csc& A, const double* B, int nrhs,
                                            std::vector<double>& X, int maxIt, double tol, bool parallel, int threads) {
  oracle_refine_result result{0, 0.0, 0.0};
  const size_t len = static_cast<size_t>(A.n) * nrhs;
  const double normA = sparseMaxAbsRowSum(A);
  const double normB = infNorm(B, len);
  auto rel = [&](const std::vector<double>& R, const std::vector<double>& Xc) {
    const double normR = infNorm(R.data(), len);
    const double normX = infNorm(Xc.data(), len);
    const double denom = normA * normX + normB;
    return denom > 0.0 ? normR / denom : normR;
  };
  std::vector<double> R;
  residual(A, B, X.data(), nrhs, R);
  double res = rel(R, X);
  result.initial_relative_residual = res;
  result.final_relative_residual = res;
  for (int it = 0; it < maxIt; ++it) {  // :121-141
    if (res <= tol) break;
    std::vector<double> dx = mfSolve(nf, R.data(), nrhs, parallel, threads);
    std::vector<double> Xnew(len);
    for (size_t i = 0; i < len; ++i) Xnew[i] = X[i] + dx[i];
    std::vector<double> Rnew;
    residual(A, B, Xnew.data(), nrhs, Rnew);
    const double resNew = rel(Rnew, Xnew);
    if (!(resNew < res)) break;
    X = std::move(Xnew);
    R = std::move(Rnew);
    res = resNew;
    result.iterations = it + 1;
    result.final_relative_residual = res;
  }
  return result;
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
using namespace oracle;

struct oracle_factor {
  NumericFactor nf;
};

static DenseOpts toDense(const oracle_dense_opts* o) {
  DenseOpts d;
  if (o) {
    d.pivot_threshold = o->pivot_threshold;
    d.zero_tolerance = o->zero_tolerance;
    d.relative_pivot_floor = o->relative_pivot_floor;
    d.panel_size = o->panel_size;
  }
  return d;
}
static StaticOpts toStatic(const oracle_static_opts* o) {
  StaticOpts s;
  if (o) {
    s.delta = o->delta;
    s.relative_pivot_floor = o->relative_pivot_floor;
    s.absolute_floor = o->absolute_floor;
    s.panel_size = o->panel_size;
  }
  return s;
}

static void fillDense(const DenseResult& r, const BlockD& D, int n, double* d_diag, double* d_off, int* perm,
                      int* ps, int* pk, int* delayed, oracle_dense_result* res) {
  for (int i = 0; i < n; ++i) {
    if (d_diag) d_diag[i] = D.diag[i];
    if (d_off) d_off[i] = D.off[i];
    if (perm) perm[i] = r.perm[i];
  }
  for (size_t i = 0; i < r.pivots.size(); ++i) {
    if (ps) ps[i] = r.pivots[i].start;
    if (pk) pk[i] = r.pivots[i].kind;
  }
  if (delayed)
    for (size_t i = 0; i < r.delayed_cols.size(); ++i) delayed[i] = r.delayed_cols[i];
  res->n_factored = r.n_factored;
  res->n_delayed = static_cast<int>(r.delayed_cols.size());
  res->n_pivots = static_cast<int>(r.pivots.size());
  res->n_pos = r.inertia.n_pos;
  res->n_neg = r.inertia.n_neg;
  res->n_zero = r.inertia.n_zero;
  res->max_growth = r.max_growth;
  res->total_perturbation = r.total_perturbation;
  res->num_perturbed = r.num_perturbed;
}

extern "C" int oracle_dense_factor(int n, double* A, int lda, int n_eligible, const oracle_dense_opts* opts,
                                   double* d_diag, double* d_off, int* perm, int* pivot_start, int* pivot_kind,
                                   int* delayed, oracle_dense_result* res) {
  try {
    BlockD D;
    DenseResult r = denseFactor(View{A, n, lda}, D, toDense(opts), n_eligible);
    fillDense(r, D, n, d_diag, d_off, perm, pivot_start, pivot_kind, delayed, res);
    return 0;
  } catch (...) {
    return -1;
  }
}

extern "C" int oracle_dense_factor_static(int n, double* A, int lda, int n_eligible, const oracle_static_opts* opts,
                                          const int* expected_sign, int n_sign, double* d_diag, double* d_off,
                                          int* perm, int* pivot_start, int* pivot_kind, oracle_dense_result* res) {
  try {
    BlockD D;
    std::vector<int> sg;
    if (expected_sign) sg.assign(expected_sign, expected_sign + n_sign);
    DenseResult r = denseFactorStatic(View{A, n, lda}, D, toStatic(opts), sg, n_eligible);
    fillDense(r, D, n, d_diag, d_off, perm, pivot_start, pivot_kind, nullptr, res);
    return 0;
  } catch (...) {
    return -1;
  }
}

static Symbolic makeSymbolic(int n, const int* perm, const int* parent, int nsn, const int* first, const int* ncols,
                             const long long* rowptr, const int* rows) {
  Symbolic s;
  s.n = n;
  s.perm.assign(perm, perm + n);
  s.etreeParent.assign(parent, parent + n);
  s.supernodes.resize(nsn);
  for (int i = 0; i < nsn; ++i) {
    s.supernodes[i].firstCol = first[i];
    s.supernodes[i].ncols = ncols[i];
    s.supernodes[i].rowPattern.assign(rows + rowptr[i], rows + rowptr[i + 1]);
  }
  return s;
}

static MfOpts toMf(const oracle_mf_opts* o) {
  MfOpts m;
  if (o) {
    m.parallel = o->parallel != 0;
    m.num_threads = o->num_threads;
    m.task_cutoff = o->task_cutoff;
    m.static_pivoting = o->static_pivoting != 0;
    m.dense = toDense(&o->dense);
    m.stat = toStatic(&o->stat);
    m.snLimit = o->sn_limit;
  }
  return m;
}

extern "C" oracle_factor* oracle_mf_factorize(int n, const int* perm, const int* etree_parent, int nsn,
                                              const int* sn_first, const int* sn_ncols, const long long* sn_rowptr,
                                              const int* sn_rows, const long long* a_colptr, const int* a_rowidx,
                                              const double* a_vals, const int* expected_sign_final,
                                              const oracle_mf_opts* opts) {
  try {
    Symbolic s = makeSymbolic(n, perm, etree_parent, nsn, sn_first, sn_ncols, sn_rowptr, sn_rows);
    MfOpts mo = toMf(opts);
    if (expected_sign_final) mo.expectedSign.assign(expected_sign_final, expected_sign_final + n);
    auto* f = new oracle_factor;
    f->nf = mfFactorize(s, Csc{n, a_colptr, a_rowidx, a_vals}, mo);
    return f;
  } catch (...) {
    return nullptr;
  }
}

extern "C" void oracle_mf_free(oracle_factor* f) { delete f; }

extern "C" void oracle_mf_get_info(const oracle_factor* f, oracle_mf_info* info) {
  const NumericFactor& nf = f->nf;
  info->n = nf.n;
  info->n_pos = nf.inertia.n_pos;
  info->n_neg = nf.inertia.n_neg;
  info->n_zero = nf.inertia.n_zero;
  info->singular = nf.singular ? 1 : 0;
  info->n_singular = static_cast<int>(nf.singularCols.size());
  info->total_perturbation = nf.totalPerturbation;
  info->num_perturbed = nf.numPerturbed;
  info->nsn = static_cast<int>(nf.fronts.size());
  long long tp = 0;
  int n2 = 0, mf = 0;
  for (const auto& fr : nf.fronts) {
    tp += fr.nPivots;
    mf = std::max(mf, static_cast<int>(fr.rowIndices.size()));
    for (const auto& pb : fr.pivotBlocks) n2 += pb.kind == TwoByTwo;
  }
  info->total_pivots = tp;
  info->n_two_by_two = n2;
  info->max_front = mf;
}

extern "C" int oracle_mf_singular_cols(const oracle_factor* f, int* out, int cap) {
  const auto& sc = f->nf.singularCols;
  const int cnt = std::min(cap, static_cast<int>(sc.size()));
  for (int i = 0; i < cnt; ++i) out[i] = sc[i];
  return cnt;
}

extern "C" int oracle_mf_front_m(const oracle_factor* f, int si) {
  return static_cast<int>(f->nf.fronts[si].rowIndices.size());
}
extern "C" int oracle_mf_front_npiv(const oracle_factor* f, int si) { return f->nf.fronts[si].nPivots; }

extern "C" int oracle_mf_front_get(const oracle_factor* f, int si, int* rows, double* L, double* d_diag,
                                   double* d_off, int* pivot_start, int* pivot_kind) {
  const auto& fr = f->nf.fronts[si];
  const int m = static_cast<int>(fr.rowIndices.size());
  const int k = fr.nPivots;
  for (int i = 0; i < m; ++i)
    if (rows) rows[i] = fr.rowIndices[i];
  if (L) std::memcpy(L, fr.L.v.data(), sizeof(double) * static_cast<size_t>(m) * k);
  for (int i = 0; i < k; ++i) {
    if (d_diag) d_diag[i] = fr.D.diag[i];
    if (d_off) d_off[i] = fr.D.off[i];
  }
  for (size_t i = 0; i < fr.pivotBlocks.size(); ++i) {
    if (pivot_start) pivot_start[i] = fr.pivotBlocks[i].start;
    if (pivot_kind) pivot_kind[i] = fr.pivotBlocks[i].kind;
  }
  return static_cast<int>(fr.pivotBlocks.size());
}

extern "C" int oracle_mf_solve(const oracle_factor* f, const double* B, int nrhs, double* X, int parallel,
                               int num_threads) {
  try {
    std::vector<double> x = mfSolve(f->nf, B, nrhs, parallel != 0, num_threads);
    std::memcpy(X, x.data(), sizeof(double) * x.size());
    return 0;
  } catch (...) {
    return -1;
  }
}

extern "C" int oracle_refine(const oracle_factor* f, int n, const long long* a_colptr, const int* a_rowidx,
                             const double* a_vals, const double* B, int nrhs, double* X, int max_iterations,
                             double tolerance, int parallel, int num_threads, oracle_refine_result* out) {
  try {
    std::vector<double> x(X, X + static_cast<size_t>(n) * nrhs);
    *out = iterativeRefine(f->nf, Csc{n, a_colptr, a_rowidx, a_vals}, B, nrhs, x, max_iterations, tolerance,
                           parallel != 0, num_threads);
    std::memcpy(X, x.data(), sizeof(double) * x.size());
    return 0;
  } catch (...) {
    return -1;
  }
}

// factorizeStaticRegularized (solver.hpp:328-417)
extern "C" oracle_factor* oracle_static_regularized(int n, const int* perm, const int* etree_parent, int nsn,
                                                    const int* sn_first, const int* sn_ncols,
                                                    const long long* sn_rowptr, const int* sn_rows,
                                                    const long long* a_colptr, const int* a_rowidx,
                                                    const double* a_vals, const int* signs_orig, double initial_delta,
                                                    int max_retries, double growth, double rel_floor,
                                                    double abs_floor, int parallel, int num_threads,
                                                    double* last_delta, int* retries_used, int* threw, char* msg,
                                                    int cap) {
  *threw = 0;
  try {
    Symbolic s = makeSymbolic(n, perm, etree_parent, nsn, sn_first, sn_ncols, sn_rowptr, sn_rows);
    const bool haveExpectation = signs_orig != nullptr;
    std::vector<int> signsOrig(n, 0);
    if (signs_orig) signsOrig.assign(signs_orig, signs_orig + n);
    std::vector<int> signsFinal(n, 0);
    for (int t = 0; t < n; ++t) signsFinal[t] = signsOrig[s.perm[t]];
    int expectedPos = 0, expectedNeg = 0;
    for (int v : signsOrig) {
      if (v > 0)
        ++expectedPos;
      else if (v < 0)
        ++expectedNeg;
    }
    double delta = initial_delta;
    if (delta <= 0.0) {  // :346-359
      double maxAbs = 0.0;
      for (long long t = 0; t < a_colptr[n]; ++t) maxAbs = std::max(maxAbs, std::abs(a_vals[t]));
      if (maxAbs <= 0.0) maxAbs = 1.0;
      delta = std::sqrt(std::numeric_limits<double>::epsilon()) * maxAbs;
    }
    const int maxAttempts = std::max(1, max_retries + 1);
    for (int attempt = 0; attempt < maxAttempts; ++attempt) {
      MfOpts mo;
      mo.parallel = parallel != 0;
      mo.num_threads = num_threads;
      mo.static_pivoting = true;
      mo.stat.delta = delta;
      mo.stat.relative_pivot_floor = rel_floor;
      mo.stat.absolute_floor = abs_floor;
      mo.expectedSign = signsFinal;
      NumericFactor r = mfFactorize(s, Csc{n, a_colptr, a_rowidx, a_vals}, mo);
      *last_delta = delta;
      *retries_used = attempt;
      const bool fullCoverage = (expectedPos + expectedNeg == n);
      const bool ok = !haveExpectation ||
                      (r.inertia.n_zero == 0 && !r.singular &&
                       (!fullCoverage || (r.inertia.n_pos == expectedPos && r.inertia.n_neg == expectedNeg)));
      if (ok) {
        auto* f = new oracle_factor;
        f->nf = std::move(r);
        return f;
      }
      if (attempt + 1 >= maxAttempts) {
        std::ostringstream oss;
        oss << "inertia-controlled regularization retry budget (" << max_retries
            << " retries) exhausted; expected (n_pos=" << expectedPos << ", n_neg=" << expectedNeg
            << ", n_zero=0), got (n_pos=" << r.inertia.n_pos << ", n_neg=" << r.inertia.n_neg
            << ", n_zero=" << r.inertia.n_zero << ") at delta=" << delta;
        *threw = 1;
        if (msg && cap > 0) std::snprintf(msg, cap, "%s", oss.str().c_str());
        return nullptr;
      }
      delta *= growth;
    }
  } catch (const std::exception& e) {
    *threw = 1;
    if (msg && cap > 0) std::snprintf(msg, cap, "%s", e.what());
  }
  return nullptr;
}
