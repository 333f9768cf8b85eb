/*
 * symla_oracle.h -- C-ABI of the CPU oracle for the numeric phase.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a CPU restatement of the reference
 * numeric path (symla: include/symla/dense_kernel.hpp, multifrontal.hpp,
 * solve.hpp, refine.hpp, solver.hpp, parallel/task_graph.hpp), used as the
 * checker for the HIP product path and as the timed CPU baseline.  Only the
 * test suite, the smoke check and the benchmark's CPU-baseline leg may load
 * it.  The product library never links, loads or calls it.
 *
 * Conventions follow the reference: column-major dense storage, `perm[t]` =
 * original index of the t-th pivot in final order, sparse matrices in CSC
 * with only the lower triangle (row >= col) read.
 */
#ifndef SYMLA_ORACLE_H
#define SYMLA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Options of DenseLDLT::factor (dense_kernel.hpp:57-116). */
typedef struct {
  double pivot_threshold;      /* carried, unused by the decision logic (dense_kernel.hpp:57-69) */
  double zero_tolerance;       /* 1e-300 */
  double relative_pivot_floor; /* 1e-12 */
  int panel_size;              /* 64 */
} oracle_dense_opts;

/* Options of DenseLDLT::factorStatic (dense_kernel.hpp:227-253). */
typedef struct {
  double delta;                /* <= 0: sqrt(eps) * orig_max */
  double relative_pivot_floor; /* 1e-8 */
  double absolute_floor;       /* 1e-300 */
  int panel_size;              /* 64 */
} oracle_static_opts;

/* Result of one dense factorization (dense_kernel.hpp:209-223).  Arrays are
 * caller-allocated with capacity n. */
typedef struct {
  int n_factored;
  int n_delayed;
  int n_pivots;
  int n_pos, n_neg, n_zero;
  double max_growth;
  double total_perturbation;
  int num_perturbed;
} oracle_dense_result;

/* DenseLDLT::factor on an n x n column-major matrix A (lower triangle read,
 * factored in place).  Outputs: d_diag[n], d_off[n] (BlockDiagonalD),
 * perm[n], pivot_start[n], pivot_kind[n] (1 or 2), delayed[n]. */
int oracle_dense_factor(int n, double* A, int lda, int n_eligible, const oracle_dense_opts* opts,
                        double* d_diag, double* d_off, int* perm, int* pivot_start, int* pivot_kind,
                        int* delayed, oracle_dense_result* res);

/* DenseLDLT::factorStatic; expected_sign may be NULL (all 0). */
int oracle_dense_factor_static(int n, double* A, int lda, int n_eligible, const oracle_static_opts* opts,
                               const int* expected_sign, int n_sign, double* d_diag, double* d_off,
                               int* perm, int* pivot_start, int* pivot_kind, oracle_dense_result* res);

/* ---- Multifrontal factorization (multifrontal.hpp:299-805) ---- */

typedef struct oracle_factor oracle_factor;

typedef struct {
  int parallel;        /* MultifrontalOptions::parallel */
  int num_threads;     /* 0 = auto (multifrontal.hpp:676-687) */
  long long task_cutoff; /* < 0 = auto */
  int static_pivoting; /* use factorStatic per front */
  oracle_dense_opts dense;
  oracle_static_opts stat;
  /* > 0: factor only supernodes [0, sn_limit) -- in postorder a union of
   * complete subtrees; the contribution blocks of their topmost fronts are
   * left unconsumed.  A timing sample (benchmarks), not a factorization. */
  int sn_limit;
} oracle_mf_opts;

/* Symbolic input (the caller's analysis): perm[n], etree_parent[n] (final
 * order), supernodes (first column, ncols, row pattern in CSR-like arrays).
 * A in CSC (colptr[n+1], rowidx, vals); only row >= col entries are read.
 * expected_sign (final order, size n) may be NULL. */
oracle_factor* oracle_mf_factorize(int n, const int* perm, const int* etree_parent, int nsn,
                                   const int* sn_first, const int* sn_ncols, const long long* sn_rowptr,
                                   const int* sn_rows, const long long* a_colptr, const int* a_rowidx,
                                   const double* a_vals, const int* expected_sign_final,
                                   const oracle_mf_opts* opts);
void oracle_mf_free(oracle_factor* f);

/* Summary of a factorization. */
typedef struct {
  int n;
  int n_pos, n_neg, n_zero;
  int singular;
  int n_singular;
  double total_perturbation;
  int num_perturbed;
  int nsn;
  long long total_pivots;
  int n_two_by_two;
  int max_front;
} oracle_mf_info;
void oracle_mf_get_info(const oracle_factor* f, oracle_mf_info* info);
/* singular columns in final order, sorted; returns count written. */
int oracle_mf_singular_cols(const oracle_factor* f, int* out, int cap);
/* per-front accessors (SupernodeFactor, multifrontal.hpp:118-135). */
int oracle_mf_front_m(const oracle_factor* f, int si);
int oracle_mf_front_npiv(const oracle_factor* f, int si);
/* rows[m], L (m x npiv col-major, ld=m), d_diag/d_off[npiv], pivot_start/kind[npiv]; returns #pivot blocks */
int oracle_mf_front_get(const oracle_factor* f, int si, int* rows, double* L, double* d_diag, double* d_off,
                        int* pivot_start, int* pivot_kind);

/* ---- Solve (solve.hpp:159-508) ---- B, X: n x nrhs column-major. */
int oracle_mf_solve(const oracle_factor* f, const double* B, int nrhs, double* X, int parallel, int num_threads);

/* ---- Iterative refinement (refine.hpp:97-144) ---- X in/out. */
typedef struct {
  int iterations;
  double initial_relative_residual;
  double final_relative_residual;
} oracle_refine_result;
int oracle_refine(const oracle_factor* f, int n, const long long* a_colptr, const int* a_rowidx,
                  const double* a_vals, const double* B, int nrhs, double* X, int max_iterations,
                  double tolerance, int parallel, int num_threads, oracle_refine_result* out);

/* ---- Mode::StaticRegularized retry loop (solver.hpp:328-417) ----
 * signs_orig: expected sign per ORIGINAL index (size n) or NULL (no
 * expectation).  Returns the accepted factorization, or NULL and sets
 * *threw = 1 with the message in msg (cap bytes) when the budget is spent. */
oracle_factor* oracle_static_regularized(int n, const int* perm, const int* etree_parent, int nsn,
                                         const int* sn_first, const int* sn_ncols, const long long* sn_rowptr,
                                         const int* sn_rows, const long long* a_colptr, const int* a_rowidx,
                                         const double* a_vals, const int* signs_orig, double initial_delta,
                                         int max_retries, double growth, double rel_floor, double abs_floor,
                                         int parallel, int num_threads, double* last_delta, int* retries_used,
                                         int* threw, char* msg, int cap);

#ifdef __cplusplus
}
#endif
#endif
