/*
 * symla_hip.h -- C-ABI of the MI355X numeric phase of symla's multifrontal
 * LDL^T solver (libsymla_hip.so).
 *
 * Plain pointers and sizes only.  Status codes: 0 = ok; otherwise one of the
 * SYMLA_ERR_* values, with a message from symla_last_error() (thread-local)
 * or symla_plan_last_error(plan).  The C++ facade in include/symla/solver.hpp
 * turns these codes into the exceptions the reference throws.
 *
 * Every entry point below cites the reference interface (file:line in the
 * symla repository) whose role it takes.
 */
#ifndef SYMLA_HIP_H
#define SYMLA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SYMLA_OK = 0,
  SYMLA_ERR_INVALID = 1, /* std::invalid_argument in the reference */
  SYMLA_ERR_LOGIC = 2,   /* std::logic_error */
  SYMLA_ERR_RUNTIME = 3, /* std::runtime_error */
  SYMLA_ERR_DEVICE = 4   /* HIP runtime failure */
};

/* OrderingType (include/symla/ordering.hpp:31-35) plus a caller permutation. */
enum { SYMLA_ORDER_AMD = 0, SYMLA_ORDER_METIS = 1, SYMLA_ORDER_NATURAL = 2, SYMLA_ORDER_USER = 3 };

/* Mode (include/symla/solver.hpp:20-23). */
enum { SYMLA_MODE_THRESHOLD = 0, SYMLA_MODE_STATIC = 1 };

/* Front storage precision (symla_factor_opts.precision).  SYMLA_PREC_BF16X3:
 * FP32 fronts whose deferred contribution-block updates (the largest fronts'
 * Schur complements) run as BF16x3 split products accumulated in FP32. */
enum { SYMLA_PREC_FP64 = 0, SYMLA_PREC_FP32 = 1, SYMLA_PREC_BF16X3 = 2 };

typedef struct symla_symbolic symla_symbolic;
typedef struct symla_plan symla_plan;

const char* symla_last_error(void);

/* ---- Symbolic analysis (CPU) ----
 * Replaces SymbolicFactor::analyze (include/symla/symbolic.hpp:195-399) as
 * called by SymLDLT::analyzePattern (include/symla/solver.hpp:133-137).
 * A: n x n CSC (colptr[n+1], rowidx); either triangle or both may be stored.
 * user_perm (perm[k] = original index of pivot k) is read when ordering is
 * SYMLA_ORDER_USER.  max_relax_size / max_relax_fill: SymbolicFactorOptions
 * (symbolic.hpp:161-184; defaults 64 and 0.25). */
int symla_analyze(int n, const int64_t* colptr, const int* rowidx, int ordering, const int* user_perm,
                  int max_relax_size, double max_relax_fill, symla_symbolic** out);

typedef struct {
  int n;
  int nsn;               /* supernodes */
  int64_t nnz_l;         /* sum of column counts */
  int64_t pattern_total; /* sum of supernode row-pattern lengths */
  int max_front;         /* largest row pattern */
} symla_symbolic_info_t;
int symla_symbolic_info(const symla_symbolic* s, symla_symbolic_info_t* info);
/* Copies SymbolicFactor's fields: perm[n], etree_parent[n] (etree.parent),
 * sn_first[nsn], sn_ncols[nsn], sn_rowptr[nsn+1], sn_rows[pattern_total]
 * (Supernode, symbolic.hpp:138-159).  Any pointer may be NULL. */
int symla_symbolic_get(const symla_symbolic* s, int* perm, int* etree_parent, int* sn_first, int* sn_ncols,
                       int64_t* sn_rowptr, int* sn_rows);
/* EliminationTree::colCount (include/symla/elimination_tree.hpp:50-53):
 * nnz(L(:, j)) including the diagonal, final order, n entries. */
int symla_symbolic_colcounts(const symla_symbolic* s, int* col_count);
void symla_symbolic_free(symla_symbolic* s);

/* Matrix Market coordinate reader: the role of Eigen::loadMarket in the
 * reference's real-matrix tests and benches (test/correctness/
 * real_matrix_test.cpp:84, bench/tier1_timing_main.cpp:24).  Real, integer
 * or pattern fields; general, symmetric or skew-symmetric storage, entries
 * kept as stored (a symmetric file's one triangle), duplicates summed, rows
 * sorted per column.  *symmetry = 0 general, 1 symmetric, 2 skew.  Call
 * once with NULL colptr/rowidx/values for the sizes, then with arrays of
 * ncols+1 / nnz / nnz entries. */
int symla_read_matrix_market(const char* path, int* nrows, int* ncols, int64_t* nnz, int* symmetry, int64_t* colptr,
                             int* rowidx, double* values);

/* ---- Numeric plan ----
 * Takes the setup part of MultifrontalFactorizer::factorize
 * (include/symla/multifrontal.hpp:299-429): supernode tree, value
 * permutation (PermutedLower, :199-242) as a precomputed scatter map, and the
 * device pool that holds every front, its contribution block and its L/D
 * factors.  A's pattern must be the one given to symla_analyze. */
int symla_plan_create(const symla_symbolic* sym, const int64_t* a_colptr, const int* a_rowidx, symla_plan** out);
void symla_plan_destroy(symla_plan* p);
const char* symla_plan_last_error(const symla_plan* p);

typedef struct {
  int mode;                     /* SYMLA_MODE_* */
  double pivot_threshold;       /* carried like the reference; no effect on decisions (dense_kernel.hpp:57-69) */
  double zero_tolerance;        /* DenseLDLTOptions::zero_tolerance (1e-300) */
  double relative_pivot_floor;  /* DenseLDLTOptions::relative_pivot_floor (1e-12) */
  /* Mode::StaticRegularized (solver.hpp:54-111, :328-417) */
  const int* expected_sign;     /* original index space, size n, values -1/0/+1; NULL = no expectation */
  double initial_delta;         /* <= 0: sqrt(eps) * max|A| */
  int max_retries;              /* 10 */
  double growth;                /* 10 */
  double static_rel_floor;      /* 1e-8 */
  double static_abs_floor;      /* 1e-300 */
  /* Front storage precision (no counterpart in the reference, which factors
   * in FP64): SYMLA_PREC_FP64, or SYMLA_PREC_FP32 = fronts, multipliers and
   * Schur updates in FP32 (FP32 MFMA), with pivot tests, 2x2 inversion, D
   * and inertia in FP64; pair with symla_solve_refined for FP64 accuracy.
   * SYMLA_PREC_BF16X3: as FP32, with the deferred contribution-block updates
   * as BF16x3 split products (hi = bf16(x), lo = bf16(x - hi); a b ~ hi hi +
   * hi lo + lo hi, FP32 accumulation). */
  int precision;
} symla_factor_opts;

typedef struct {
  int n_pos, n_neg, n_zero;     /* Inertia (inertia.hpp:14-18) */
  int singular;                 /* NumericFactor::singular */
  int n_singular;
  double total_perturbation;    /* NumericFactor::totalPerturbation */
  int num_perturbed;
  int retries_used;             /* regularizationRetriesUsed() */
  double last_delta;            /* lastRegularizationDelta() */
  int delayed_fronts;           /* fronts that handed delayed columns to their parent */
  int repaired;                 /* 1 when the delayed-pivot repair pass ran */
  /* SYMLA_PREC_* of the returned factorization.  Static-regularized mode
   * with FP32 fronts falls back to the reference's FP64 retry sequence when
   * the FP32 attempt misses the expected inertia (no reference counterpart). */
  int precision_used;
} symla_factor_info;

/* SymLDLT::factorize / refactorize (solver.hpp:140-172) through
 * MultifrontalFactorizer::factorize (multifrontal.hpp:299-805); in static
 * mode the inertia-controlled retry loop of factorizeStaticRegularized
 * (solver.hpp:328-417).  a_vals: values in the CSC order of plan_create's
 * pattern; host memory (symla_factorize) or device memory
 * (symla_factorize_device).  Returns SYMLA_ERR_RUNTIME when the static retry
 * budget is exhausted (the plan is then not factorized). */
int symla_factorize(symla_plan* p, const double* a_vals, const symla_factor_opts* opts, symla_factor_info* info);
int symla_factorize_device(symla_plan* p, const double* d_a_vals, const symla_factor_opts* opts,
                           symla_factor_info* info);

/* SymLDLT::singularColumns (solver.hpp:251-257): original indices. */
int symla_singular_cols(const symla_plan* p, int* out, int cap);

/* Inertia and singular-column count per group of original indices (no
 * counterpart in the reference, which factors one system per call).  For a
 * block-diagonal matrix assembled from independent systems, group g owns the
 * original indices [group_start[g], group_start[g+1]) (group_start[0] = 0,
 * group_start[ngroups] = n); every supernode lies inside one system, so the
 * per-front counts add up per system.  inertia3: 3 ints per group {pos, neg,
 * zero}; singular (may be NULL): singular columns per group. */
int symla_group_inertia(symla_plan* p, int ngroups, const int* group_start, int* inertia3, int* singular);

/* ---- Batches of independent systems (no counterpart in the reference,
 * which factors one system per SymLDLT, include/symla/solver.hpp:29-418).
 * System s is an n[s] x n[s] CSC pattern (colptr[s], rowidx[s]).  The
 * batch is factored and solved as the block-diagonal matrix of its systems:
 * every launch of the numeric phase covers the fronts of all systems at one
 * tree level.
 * symla_batch_analyze runs symla_analyze on each system (on host threads)
 * and joins the results as the analysis of the block-diagonal matrix;
 * symla_batch_plan_create builds its plan from the per-system patterns.
 * Then symla_factorize / symla_solve take the systems' values and
 * right-hand sides concatenated (system s's values after system s-1's, its
 * rows after system s-1's rows), symla_group_inertia returns per-system
 * inertia (group_start = running sums of n), and symla_batch_solve_refined
 * runs iterativeRefine (refine.hpp:97-144) for each system on its own: its
 * own norms and stopping tests, per_system[s] its iterations and backward
 * error (sys_start: nsys + 1 running sums of n). */
int symla_batch_analyze(int nsys, const int* n, const int64_t* const* colptr, const int* const* rowidx, int ordering,
                        int max_relax_size, double max_relax_fill, symla_symbolic** out);
int symla_batch_plan_create(const symla_symbolic* sym, int nsys, const int* n, const int64_t* const* colptr,
                            const int* const* rowidx, symla_plan** out);
/* SymLDLT::solve (solver.hpp:198-218) / MultifrontalSolver::solve
 * (solve.hpp:159-431).  B, X: n x nrhs column-major.  Host pointers
 * (symla_solve) or device pointers (symla_solve_device).  SYMLA_ERR_LOGIC
 * before a factorization or on a singular one. */
int symla_solve(symla_plan* p, const double* B, int64_t ldb, int nrhs, double* X, int64_t ldx);
int symla_solve_device(symla_plan* p, const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);

/* No counterpart in the reference: symla_solve_device for callers that
 * cannot allocate device memory (code under stream capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same kernels and results as
 * symla_solve_device; the per-call workspace, (3 n + w) x nrhs doubles
 * (w: the solve's work rows, the sum of the front orders plus the partial
 * rows of the split backward products), comes from a static 128 MiB device array of the library
 * instead of plan buffers, and the call makes no hipMalloc / hipFree.  A
 * solve needing more returns SYMLA_ERR_INVALID (nothing runs).  Call
 * symla_factorize first, on the current device (the plan's).  The array of
 * a device is set up by the first symla_plan_create on it and held until
 * the process ends.  All calls on one GPU share that GPU's array: calls that
 * overlap on one GPU race. */
int symla_solve_device_noalloc(symla_plan* p, const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);

/* SymLDLT::solveWithRefinement (solver.hpp:234-240) with iterativeRefine
 * (refine.hpp:97-144).  a_vals must be the values last factorized. */
typedef struct {
  int max_iterations; /* 10 */
  double tolerance;   /* 1e-14 */
} symla_refine_opts;
typedef struct {
  int iterations;
  double initial_relative_residual; /* refine.hpp's residual (its own |A| row-sum norm) */
  double final_relative_residual;
  double backward_error;            /* max_c ||r_c||inf / (||A||inf ||x_c||inf + ||b_c||inf), symmetric ||A||inf */
} symla_refine_info;
int symla_solve_refined(symla_plan* p, const double* a_vals, const double* B, int64_t ldb, int nrhs, double* X,
                        int64_t ldx, const symla_refine_opts* opts, symla_refine_info* info);
int symla_solve_refined_device(symla_plan* p, const double* d_a_vals, const double* dB, int64_t ldb, int nrhs,
                               double* dX, int64_t ldx, const symla_refine_opts* opts, symla_refine_info* info);

/* Per-system refinement of a batch (see symla_batch_analyze). */
int symla_batch_solve_refined(symla_plan* p, int nsys, const int* sys_start, const double* a_vals, const double* B,
                              int64_t ldb, int nrhs, double* X, int64_t ldx, const symla_refine_opts* opts,
                              symla_refine_info* per_system);
int symla_batch_solve_refined_device(symla_plan* p, int nsys, const int* sys_start, const double* d_a_vals,
                                     const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx,
                                     const symla_refine_opts* opts, symla_refine_info* per_system);

/* Recorded refactorize + solve (no counterpart in the reference; it serves
 * callers that refactorize one pattern many times, e.g. inside a time or
 * Newton loop).  symla_graph_capture records, as one hipGraph on the plan's
 * stream, the threshold-pivoting factorization of the values at d_a_vals and
 * the solve of dB into dX (all device pointers, kept by the plan).
 * symla_graph_replay re-runs both on the current contents of those buffers
 * and fills info; a replay whose factorization delays pivots is repaired on
 * the regular path (info->repaired = 1) and solved again, and on a singular
 * factorization dX is not meaningful (info->singular = 1).  The graph is
 * re-recorded by itself when a buffer it refers to has moved (another solve
 * size, a repair).  Static-regularized mode is not recorded (its retry loop
 * decides on the host): SYMLA_ERR_INVALID. */
int symla_graph_capture(symla_plan* p, const double* d_a_vals, const double* dB, int64_t ldb, int nrhs, double* dX,
                        int64_t ldx, const symla_factor_opts* opts);
int symla_graph_replay(symla_plan* p, symla_factor_info* info);
/* The same with FP64 iterative refinement of the recorded solve
 * (SymLDLT::solveWithRefinement, solver.hpp:234-240; iterativeRefine,
 * refine.hpp:97-144): symla_graph_capture_refined records the
 * factorization, the norms of A and B, the first solve and its residual as
 * one graph, and one refinement step (correction solve, update, residual) as
 * two more (ping-pong buffers).  symla_graph_replay_refined launches the
 * first, then one step graph per iteration while refine.hpp's tests (stop at
 * ropts->tolerance, when the residual no longer decreases, or after
 * ropts->max_iterations) ask for another, and fills rinfo like
 * symla_solve_refined.  A replay that delays pivots is repaired and refined
 * on the regular path. */
int symla_graph_capture_refined(symla_plan* p, const double* d_a_vals, const double* dB, int64_t ldb, int nrhs,
                                double* dX, int64_t ldx, const symla_factor_opts* opts,
                                const symla_refine_opts* ropts);
int symla_graph_replay_refined(symla_plan* p, symla_factor_info* info, symla_refine_info* rinfo);
int symla_graph_release(symla_plan* p);

/* NumericFactor::fronts (multifrontal.hpp:118-186), copied to the host on
 * request: one SupernodeFactor.  rows[m] (final-order rowIndices), L
 * (m x npiv column-major, ld m), d_diag/d_off[npiv] (BlockDiagonalD),
 * pivot_start/pivot_kind[npiv] (PivotBlock; kind 1 or 2), *nblocks = pivot
 * blocks.  Pass NULL arrays to query m and npiv first. */
int symla_front_count(const symla_plan* p);
int symla_front_view(symla_plan* p, int si, int* m, int* npiv, int* rows, double* L, double* d_diag,
                     double* d_off, int* pivot_start, int* pivot_kind, int* nblocks);

/* DenseLDLT::factor / factorStatic (include/symla/dense_kernel.hpp:393-837,
 * :919-1059) on one dense n x n column-major matrix, run by the same device
 * front kernels (test entry).  A is factored in place (lower triangle
 * meaningful).  Outputs have capacity n; inertia3 = {pos, neg, zero}. */
typedef struct {
  double zero_tolerance;
  double relative_pivot_floor;
  int static_mode;
  double static_delta;
  double static_rel_floor;
  double static_abs_floor;
  const int* expected_sign; /* static mode, size n or NULL */
  int exact_panel;          /* 1: no speculative panel step (pivot-by-pivot kernel only) */
  int deferred_cb;          /* 1: rows and columns [n_eligible, n) take one rank-n_eligible update after the
                               last panel instead of one per panel (the plan default for large blocks) */
} symla_dense_opts;
int symla_dense_factor(int n, double* A, int lda, int n_eligible, const symla_dense_opts* opts, double* d_diag,
                       double* d_off, int* perm, int* pivot_start, int* pivot_kind, int* n_pivot_blocks,
                       int* n_factored, int* inertia3, double* total_perturbation, int* num_perturbed);

/* Scheduling choices of a plan (no counterpart in the reference, whose
 * front kernel is one CPU routine): fronts of order <= small_front_max
 * (at most 192, default 64) are factored whole in LDS, larger ones by the
 * panel path; assembly runs in column tiles of assemble_tile (a power of two
 * in [16, 1024]).  With speculative_panel set, each panel of a front with
 * more than speculative_min_rows eligible rows is first factored as 1x1
 * pivots in order and the threshold tests checked for all its columns at
 * once; the exact pivot-by-pivot kernel continues from the first column that
 * fails (the same pivot decisions, up to rounding of the tested values).
 * symla_plan_autotune times factorize + solve of the values at d_a_vals
 * (device) with nrhs fixed pseudo-random right-hand sides, best of reps, in
 * the given precision, by coordinate descent from the plan's current
 * tuning: each choice in turn over its candidates (small_front_max 0..192,
 * assemble_tile 32/64/128, the speculative step off or from 0/1,024/2,048/
 * 4,096 rows, the deferred contribution-block bound off or 1,024/2,048/
 * 4,096) with the others fixed, the fastest kept, until a sweep changes
 * nothing.  It keeps the fastest (returned in *chosen) and reports every
 * trial (up to cap records, *ntrials = number timed).  The plan then holds a
 * factorization of those values.  symla_plan_autotune_cached first looks up
 * cache_path (a text file, one line per result) for the same key -- GPU
 * architecture, precision, nrhs and histogram of front orders -- and applies
 * a stored tuning without timing (*from_cache = 1, *ntrials = 0); otherwise it
 * tunes and appends the result. */
typedef struct {
  int small_front_max;
  int assemble_tile;
  int speculative_panel;    /* 1 (default): speculative blocked step on large fronts; 0: exact steps only */
  int speculative_min_rows; /* fronts with more eligible rows than this take it (default 2048; 0 when
                               a front has more than 4096 eligible rows) */
  int deferred_cb_min_rows; /* large fronts with at least 256 eligible columns and a contribution block of at
                               least this many rows update that block once, after their last panel, as
                               library SYRKs (default 2048; -1: every block panel by panel) */
} symla_tuning;
typedef struct {
  symla_tuning tuning;
  double ms;
} symla_tuning_trial;
int symla_plan_get_tuning(const symla_plan* p, symla_tuning* out);
int symla_plan_set_tuning(symla_plan* p, const symla_tuning* t);
int symla_plan_autotune(symla_plan* p, const double* d_a_vals, int nrhs, int precision, int reps,
                        symla_tuning* chosen, symla_tuning_trial* trials, int cap, int* ntrials);
int symla_plan_autotune_cached(symla_plan* p, const double* d_a_vals, int nrhs, int precision, int reps,
                               const char* cache_path, symla_tuning* chosen, symla_tuning_trial* trials, int cap,
                               int* ntrials, int* from_cache);

/* Device-side timing of the last factorize / solve (milliseconds, HIP
 * events on the plan's stream). */
int symla_plan_timing(const symla_plan* p, double* factor_ms, double* solve_ms);

/* Per-kernel-class profile (no counterpart in the reference, which only
 * has the compile-time SYMLA_PROFILE accumulators of multifrontal.hpp:56-78).
 * When enabled, every launch of the listed kernel classes is bracketed by
 * HIP events on the plan's stream; symla_plan_get_profile returns the sums
 * since the previous call and resets them.  schur_flops is the algorithmic
 * flop count p*r*(r+1) of the panels' MFMA trailing updates (lower triangle,
 * 2 flops per multiply-add), cb_flops that of the deferred contribution-block
 * updates. */
typedef struct {
  double syrk_ms;
  long long syrk_launches;
  double schur_flops;
  double panel_ms;
  long long panel_launches;
  double panel_rows_ms;
  long long panel_rows_launches;
  double small_ms;
  long long small_launches;
  double assemble_ms;
  long long assemble_launches;
  double fwd_ms;
  long long fwd_launches;
  double bwd_ms;
  long long bwd_launches;
  double cb_ms;        /* deferred contribution-block updates (operand kernel + library SYRK per chunk) */
  long long cb_launches;
  double cb_flops;     /* their algorithmic flops, 2 k r (r + 1) / 2 per chunk of k pivots and r rows */
} symla_profile;
int symla_plan_set_profiling(symla_plan* p, int enable);
int symla_plan_get_profile(symla_plan* p, symla_profile* out);

#ifdef __cplusplus
}
#endif
#endif
