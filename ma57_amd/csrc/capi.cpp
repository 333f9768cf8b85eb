// capi.cpp -- extern "C" entry points of libsymla_hip.so (include/symla_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/symla_hip.h"
#include "kernels.hpp"
#include "mmio.hpp"
#include "plan.hpp"
#include "symbolic.hpp"

using namespace symla_hip;

struct symla_symbolic {
  SymbolicResult r;
};

struct symla_plan {
  std::unique_ptr<Plan> plan;
  std::string err;
};

namespace {

thread_local std::string g_lastError;

int fail(int code, const std::string& msg) {
  g_lastError = msg;
  return code;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_lastError.clear();
    // every HIP call and launch above was checked; leave no per-thread
    // error state behind for the caller's own HIP work
    (void)hipGetLastError();
    return SYMLA_OK;
  } catch (const std::invalid_argument& e) {
    return fail(SYMLA_ERR_INVALID, e.what());
  } catch (const std::logic_error& e) {
    return fail(SYMLA_ERR_LOGIC, e.what());
  } catch (const DeviceError& e) {
    return fail(SYMLA_ERR_DEVICE, e.what());
  } catch (const std::runtime_error& e) {
    return fail(SYMLA_ERR_RUNTIME, e.what());
  } catch (const std::exception& e) {
    return fail(SYMLA_ERR_RUNTIME, e.what());
  }
}

int planGuarded(symla_plan* p, const std::function<void()>& f) {
  const int rc = guarded(f);
  if (p) p->err = (rc == SYMLA_OK) ? std::string() : g_lastError;
  return rc;
}

FactorOptions toFactorOptions(const symla_factor_opts* o, int n) {
  FactorOptions f;
  if (!o) return f;
  f.mode = o->mode;
  f.zeroTol = o->zero_tolerance;
  f.relFloor = o->relative_pivot_floor;
  if (o->expected_sign) {
    f.expectedSignOrig.assign(o->expected_sign, o->expected_sign + n);
    f.haveExpectation = true;
  }
  f.initialDelta = o->initial_delta;
  f.maxRetries = o->max_retries;
  f.growth = o->growth;
  f.staticRelFloor = o->static_rel_floor;
  f.staticAbsFloor = o->static_abs_floor;
  // SYMLA_PREC_BF16X3 keeps FP32 fronts; only the deferred contribution-block updates change
  if (o->precision == SYMLA_PREC_BF16X3) {
    f.precision = SYMLA_PREC_FP32;
    f.bf16x3 = true;
  } else {
    f.precision = o->precision;
  }
  return f;
}

void fillInfo(const FactorInfo& fi, symla_factor_info* info) {
  if (!info) return;
  info->n_pos = fi.pos;
  info->n_neg = fi.neg;
  info->n_zero = fi.zero;
  info->singular = fi.singular ? 1 : 0;
  info->n_singular = fi.nSingular;
  info->total_perturbation = fi.totalPerturbation;
  info->num_perturbed = fi.numPerturbed;
  info->retries_used = fi.retriesUsed;
  info->last_delta = fi.lastDelta;
  info->delayed_fronts = fi.delayedFronts;
  info->repaired = fi.repaired ? 1 : 0;
  info->precision_used = fi.precisionUsed;
}

}  // namespace

extern "C" {

const char* symla_last_error(void) { return g_lastError.c_str(); }

int symla_analyze(int n, const int64_t* colptr, const int* rowidx, int ordering, const int* user_perm,
                  int max_relax_size, double max_relax_fill, symla_symbolic** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("symla_analyze: out is null");
    *out = nullptr;
    if (n < 0) throw std::invalid_argument("symla_analyze: n must be non-negative");
    if (n > 0 && (!colptr || (colptr[n] > 0 && !rowidx))) throw std::invalid_argument("symla_analyze: null pattern");
    if (ordering < 0 || ordering > 3) throw std::invalid_argument("symla::computeOrdering: unknown OrderingType");
    SymbolicOptions so;
    so.max_relax_size = max_relax_size;
    so.max_relax_fill_fraction = max_relax_fill;
    auto s = std::make_unique<symla_symbolic>();
    s->r = analyze(n, colptr, rowidx, static_cast<Ordering>(ordering), so, user_perm);
    *out = s.release();
  });
}

int symla_batch_analyze(int nsys, const int* n, const int64_t* const* colptr, const int* const* rowidx, int ordering,
                        int max_relax_size, double max_relax_fill, symla_symbolic** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("symla_batch_analyze: out is null");
    *out = nullptr;
    if (nsys <= 0 || !n || !colptr || !rowidx) throw std::invalid_argument("symla_batch_analyze: empty batch");
    if (ordering < 0 || ordering > 2) throw std::invalid_argument("symla::computeOrdering: unknown OrderingType");
    for (int s = 0; s < nsys; ++s) {
      if (n[s] < 0) throw std::invalid_argument("symla_batch_analyze: n must be non-negative");
      if (n[s] > 0 && (!colptr[s] || (colptr[s][n[s]] > 0 && !rowidx[s])))
        throw std::invalid_argument("symla_batch_analyze: null pattern");
    }
    SymbolicOptions so;
    so.max_relax_size = max_relax_size;
    so.max_relax_fill_fraction = max_relax_fill;
    // the systems are independent: analyze them on host threads
    std::vector<SymbolicResult> parts(nsys);
    std::vector<std::string> errs(nsys);
    const int nthreads = std::max(1, std::min<int>({nsys, 16, static_cast<int>(std::thread::hardware_concurrency())}));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
      pool.emplace_back([&, t] {
        for (int s = t; s < nsys; s += nthreads) {
          try {
            parts[s] = analyze(n[s], colptr[s], rowidx[s], static_cast<Ordering>(ordering), so, nullptr);
          } catch (const std::exception& e) {
            errs[s] = e.what();
          }
        }
      });
    for (auto& th : pool) th.join();
    for (int s = 0; s < nsys; ++s)
      if (!errs[s].empty()) throw std::invalid_argument("symla_batch_analyze: system " + std::to_string(s) + ": " + errs[s]);
    auto sym = std::make_unique<symla_symbolic>();
    sym->r = concatenateAnalyses(parts);
    *out = sym.release();
  });
}

int symla_symbolic_info(const symla_symbolic* s, symla_symbolic_info_t* info) {
  return guarded([&] {
    if (!s || !info) throw std::invalid_argument("symla_symbolic_info: null argument");
    info->n = s->r.n;
    info->nsn = static_cast<int>(s->r.supernodes.size());
    info->nnz_l = s->r.nnzL;
    int64_t tot = 0;
    int mx = 0;
    for (const auto& sn : s->r.supernodes) {
      tot += static_cast<int64_t>(sn.rowPattern.size());
      mx = std::max(mx, static_cast<int>(sn.rowPattern.size()));
    }
    info->pattern_total = tot;
    info->max_front = mx;
  });
}

int symla_symbolic_get(const symla_symbolic* s, int* perm, int* etree_parent, int* sn_first, int* sn_ncols,
                       int64_t* sn_rowptr, int* sn_rows) {
  return guarded([&] {
    if (!s) throw std::invalid_argument("symla_symbolic_get: null symbolic");
    const auto& r = s->r;
    if (perm) std::copy(r.perm.begin(), r.perm.end(), perm);
    if (etree_parent) std::copy(r.etreeParent.begin(), r.etreeParent.end(), etree_parent);
    int64_t off = 0;
    for (size_t i = 0; i < r.supernodes.size(); ++i) {
      const auto& sn = r.supernodes[i];
      if (sn_first) sn_first[i] = sn.firstCol;
      if (sn_ncols) sn_ncols[i] = sn.ncols;
      if (sn_rowptr) sn_rowptr[i] = off;
      if (sn_rows) std::copy(sn.rowPattern.begin(), sn.rowPattern.end(), sn_rows + off);
      off += static_cast<int64_t>(sn.rowPattern.size());
    }
    if (sn_rowptr) sn_rowptr[r.supernodes.size()] = off;
  });
}

int symla_symbolic_colcounts(const symla_symbolic* s, int* col_count) {
  return guarded([&] {
    if (!s || (!col_count && s->r.n > 0)) throw std::invalid_argument("symla_symbolic_colcounts: null argument");
    std::copy(s->r.colCount.begin(), s->r.colCount.end(), col_count);
  });
}

void symla_symbolic_free(symla_symbolic* s) { delete s; }

int symla_read_matrix_market(const char* path, int* nrows, int* ncols, int64_t* nnz, int* symmetry, int64_t* colptr,
                             int* rowidx, double* values) {
  return guarded([&] {
    if (!path) throw std::invalid_argument("symla_read_matrix_market: null path");
    const MarketMatrix M = readMatrixMarket(path);
    if (nrows) *nrows = M.rows;
    if (ncols) *ncols = M.cols;
    if (nnz) *nnz = static_cast<int64_t>(M.rowidx.size());
    if (symmetry) *symmetry = M.symmetry;
    if (colptr) std::copy(M.colptr.begin(), M.colptr.end(), colptr);
    if (rowidx) std::copy(M.rowidx.begin(), M.rowidx.end(), rowidx);
    if (values) std::copy(M.values.begin(), M.values.end(), values);
  });
}

int symla_plan_create(const symla_symbolic* sym, const int64_t* a_colptr, const int* a_rowidx, symla_plan** out) {
  return guarded([&] {
    if (!sym || !out) throw std::invalid_argument("symla_plan_create: null argument");
    *out = nullptr;
    if (sym->r.n > 0 && !a_colptr) throw std::invalid_argument("symla_plan_create: null pattern");
    auto p = std::make_unique<symla_plan>();
    p->plan = std::make_unique<Plan>(sym->r, a_colptr, a_rowidx);
    // set-up of the allocation-free solve: the static scratch of this device
    // is loaded and its address looked up now, so that entry point's calls
    // take no device memory at all
    int dev = 0;
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    (void)noAllocScratch(dev);
    *out = p.release();
  });
}

void symla_plan_destroy(symla_plan* p) { delete p; }

const char* symla_plan_last_error(const symla_plan* p) { return p ? p->err.c_str() : g_lastError.c_str(); }

static int factorizeImpl(symla_plan* p, const double* host, const double* dev, const symla_factor_opts* opts,
                         symla_factor_info* info) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_factorize: null plan");
    FactorOptions fo = toFactorOptions(opts, p->plan->n());
    FactorInfo fi;
    p->plan->factorize(host, dev, fo, fi);
    fillInfo(fi, info);
  });
}

int symla_factorize(symla_plan* p, const double* a_vals, const symla_factor_opts* opts, symla_factor_info* info) {
  return factorizeImpl(p, a_vals, nullptr, opts, info);
}

int symla_factorize_device(symla_plan* p, const double* d_a_vals, const symla_factor_opts* opts,
                           symla_factor_info* info) {
  return factorizeImpl(p, nullptr, d_a_vals, opts, info);
}

int symla_batch_plan_create(const symla_symbolic* sym, int nsys, const int* n, const int64_t* const* colptr,
                            const int* const* rowidx, symla_plan** out) {
  return guarded([&] {
    if (!sym || !out || nsys <= 0 || !n || !colptr || !rowidx)
      throw std::invalid_argument("symla_batch_plan_create: null argument");
    *out = nullptr;
    // the block-diagonal pattern: system s's columns follow system s-1's, its rows shifted by the same offset
    int total = 0;
    int64_t nnz = 0;
    for (int s = 0; s < nsys; ++s) {
      total += n[s];
      nnz += n[s] > 0 ? colptr[s][n[s]] : 0;
    }
    if (total != sym->r.n) throw std::invalid_argument("symla_batch_plan_create: sizes differ from the analysis");
    std::vector<int64_t> cp;
    cp.reserve(static_cast<size_t>(total) + 1);
    cp.push_back(0);
    std::vector<int> ri;
    ri.reserve(static_cast<size_t>(nnz));
    int off = 0;
    for (int s = 0; s < nsys; ++s) {
      for (int c = 0; c < n[s]; ++c) {
        for (int64_t t = colptr[s][c]; t < colptr[s][c + 1]; ++t) {
          const int r = rowidx[s][t];
          if (r < 0 || r >= n[s]) throw std::invalid_argument("symla: row index out of range");
          ri.push_back(r + off);
        }
        cp.push_back(static_cast<int64_t>(ri.size()));
      }
      off += n[s];
    }
    auto p = std::make_unique<symla_plan>();
    p->plan = std::make_unique<Plan>(sym->r, cp.data(), ri.data());
    int dev = 0;
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    (void)noAllocScratch(dev);
    *out = p.release();
  });
}

static int batchRefinedImpl(symla_plan* p, int nsys, const int* sys_start, const double* aHost, const double* aDev,
                            const double* B, int64_t ldb, int nrhs, double* X, int64_t ldx, bool dev,
                            const symla_refine_opts* opts, symla_refine_info* per_system) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_batch_solve_refined: null plan");
    if (ldb < p->plan->n() || ldx < p->plan->n())
      throw std::invalid_argument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size");
    const int maxIt = opts ? opts->max_iterations : 10;
    const double tol = opts ? opts->tolerance : 1e-14;
    const std::vector<RefineResult> r =
        p->plan->solveRefinedGroups(aHost, aDev, B, ldb, nrhs, X, ldx, dev, maxIt, tol, nsys, sys_start);
    if (per_system)
      for (int s = 0; s < nsys; ++s) {
        per_system[s].iterations = r[s].iterations;
        per_system[s].initial_relative_residual = r[s].initialRel;
        per_system[s].final_relative_residual = r[s].finalRel;
        per_system[s].backward_error = r[s].backwardError;
      }
  });
}

int symla_batch_solve_refined(symla_plan* p, int nsys, const int* sys_start, const double* a_vals, const double* B,
                              int64_t ldb, int nrhs, double* X, int64_t ldx, const symla_refine_opts* opts,
                              symla_refine_info* per_system) {
  return batchRefinedImpl(p, nsys, sys_start, a_vals, nullptr, B, ldb, nrhs, X, ldx, false, opts, per_system);
}

int symla_batch_solve_refined_device(symla_plan* p, int nsys, const int* sys_start, const double* d_a_vals,
                                     const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx,
                                     const symla_refine_opts* opts, symla_refine_info* per_system) {
  return batchRefinedImpl(p, nsys, sys_start, nullptr, d_a_vals, dB, ldb, nrhs, dX, ldx, true, opts, per_system);
}

int symla_singular_cols(const symla_plan* p, int* out, int cap) {
  if (!p) return fail(SYMLA_ERR_INVALID, "symla_singular_cols: null plan"), -1;
  const auto& sc = p->plan->singularColsOrig();
  const int cnt = std::min<int>(cap, static_cast<int>(sc.size()));
  for (int i = 0; i < cnt && out; ++i) out[i] = sc[i];
  return static_cast<int>(sc.size());
}

int symla_group_inertia(symla_plan* p, int ngroups, const int* group_start, int* inertia3, int* singular) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_group_inertia: null plan");
    p->plan->groupInertia(ngroups, group_start, inertia3, singular);
  });
}

int symla_graph_capture(symla_plan* p, const double* d_a_vals, const double* dB, int64_t ldb, int nrhs, double* dX,
                        int64_t ldx, const symla_factor_opts* opts) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_graph_capture: null plan");
    p->plan->graphCapture(d_a_vals, dB, ldb, nrhs, dX, ldx, toFactorOptions(opts, p->plan->n()));
  });
}

int symla_graph_replay(symla_plan* p, symla_factor_info* info) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_graph_replay: null plan");
    FactorInfo fi;
    p->plan->graphReplay(fi);
    fillInfo(fi, info);
  });
}

int symla_graph_capture_refined(symla_plan* p, const double* d_a_vals, const double* dB, int64_t ldb, int nrhs,
                                double* dX, int64_t ldx, const symla_factor_opts* opts,
                                const symla_refine_opts* ropts) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_graph_capture_refined: null plan");
    p->plan->graphCaptureRefined(d_a_vals, dB, ldb, nrhs, dX, ldx, toFactorOptions(opts, p->plan->n()),
                                 ropts ? ropts->max_iterations : 10, ropts ? ropts->tolerance : 1e-14);
  });
}

int symla_graph_replay_refined(symla_plan* p, symla_factor_info* info, symla_refine_info* rinfo) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_graph_replay_refined: null plan");
    FactorInfo fi;
    RefineResult r;
    p->plan->graphReplay(fi, &r);
    fillInfo(fi, info);
    if (rinfo) {
      rinfo->iterations = r.iterations;
      rinfo->initial_relative_residual = r.initialRel;
      rinfo->final_relative_residual = r.finalRel;
      rinfo->backward_error = r.backwardError;
    }
  });
}

int symla_graph_release(symla_plan* p) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_graph_release: null plan");
    p->plan->graphRelease();
  });
}

int symla_plan_get_tuning(const symla_plan* p, symla_tuning* out) {
  if (!p || !out) return fail(SYMLA_ERR_INVALID, "symla_plan_get_tuning: null argument");
  out->small_front_max = p->plan->tuning().smallMax;
  out->assemble_tile = p->plan->tuning().asmTile;
  out->speculative_panel = p->plan->tuning().specPanel;
  out->speculative_min_rows = p->plan->tuning().specMinRows;
  out->deferred_cb_min_rows = p->plan->tuning().cbMinRows;
  return SYMLA_OK;
}

int symla_plan_set_tuning(symla_plan* p, const symla_tuning* t) {
  return planGuarded(p, [&] {
    if (!p || !t) throw std::invalid_argument("symla_plan_set_tuning: null argument");
    if (t->speculative_panel != 0 && t->speculative_panel != 1)
      throw std::invalid_argument("symla_plan_set_tuning: speculative_panel must be 0 or 1");
    p->plan->setTuning(
        Plan::Tuning{t->small_front_max, t->assemble_tile, t->speculative_panel, t->speculative_min_rows,
                     t->deferred_cb_min_rows});
  });
}

static int autotuneImpl(symla_plan* p, const double* d_a_vals, int nrhs, int precision, int reps,
                        const char* cache_path, symla_tuning* chosen, symla_tuning_trial* trials, int cap,
                        int* ntrials, int* from_cache) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_plan_autotune: null plan");
    bool cached = false;
    const auto tr = p->plan->autotune(d_a_vals, nrhs, precision, reps, cache_path, &cached);
    if (from_cache) *from_cache = cached ? 1 : 0;
    if (chosen) {
      chosen->small_front_max = p->plan->tuning().smallMax;
      chosen->assemble_tile = p->plan->tuning().asmTile;
      chosen->speculative_panel = p->plan->tuning().specPanel;
      chosen->speculative_min_rows = p->plan->tuning().specMinRows;
      chosen->deferred_cb_min_rows = p->plan->tuning().cbMinRows;
    }
    if (ntrials) *ntrials = static_cast<int>(tr.size());
    for (int i = 0; trials && i < cap && i < static_cast<int>(tr.size()); ++i) {
      trials[i].tuning.small_front_max = tr[i].t.smallMax;
      trials[i].tuning.assemble_tile = tr[i].t.asmTile;
      trials[i].tuning.speculative_panel = tr[i].t.specPanel;
      trials[i].tuning.speculative_min_rows = tr[i].t.specMinRows;
      trials[i].tuning.deferred_cb_min_rows = tr[i].t.cbMinRows;
      trials[i].ms = tr[i].ms;
    }
  });
}

int symla_plan_autotune(symla_plan* p, const double* d_a_vals, int nrhs, int precision, int reps,
                        symla_tuning* chosen, symla_tuning_trial* trials, int cap, int* ntrials) {
  return autotuneImpl(p, d_a_vals, nrhs, precision, reps, nullptr, chosen, trials, cap, ntrials, nullptr);
}

int symla_plan_autotune_cached(symla_plan* p, const double* d_a_vals, int nrhs, int precision, int reps,
                               const char* cache_path, symla_tuning* chosen, symla_tuning_trial* trials, int cap,
                               int* ntrials, int* from_cache) {
  if (!cache_path) return fail(SYMLA_ERR_INVALID, "symla_plan_autotune_cached: null cache path");
  return autotuneImpl(p, d_a_vals, nrhs, precision, reps, cache_path, chosen, trials, cap, ntrials, from_cache);
}

int symla_solve(symla_plan* p, const double* B, int64_t ldb, int nrhs, double* X, int64_t ldx) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_solve: null plan");
    if (ldb < p->plan->n() || ldx < p->plan->n())
      throw std::invalid_argument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size");
    p->plan->solve(B, ldb, nrhs, X, ldx, false);
  });
}

int symla_solve_device(symla_plan* p, const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_solve_device: null plan");
    if (ldb < p->plan->n() || ldx < p->plan->n())
      throw std::invalid_argument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size");
    p->plan->solve(dB, ldb, nrhs, dX, ldx, true);
  });
}

int symla_solve_device_noalloc(symla_plan* p, const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_solve_device_noalloc: null plan");
    if (ldb < p->plan->n() || ldx < p->plan->n())
      throw std::invalid_argument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size");
    int dev = 0;
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    p->plan->solveNoAlloc(dB, ldb, nrhs, dX, ldx, noAllocScratch(dev), kNoAllocScratchBytes / sizeof(double));
  });
}

static int refinedImpl(symla_plan* p, const double* aHost, const double* aDev, const double* B, int64_t ldb,
                       int nrhs, double* X, int64_t ldx, bool dev, const symla_refine_opts* opts,
                       symla_refine_info* info) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_solve_refined: null plan");
    if (ldb < p->plan->n() || ldx < p->plan->n())
      throw std::invalid_argument("symla::SymLDLT::solve: B.rows() must match the factorized matrix size");
    const int maxIt = opts ? opts->max_iterations : 10;
    const double tol = opts ? opts->tolerance : 1e-14;
    RefineResult r = p->plan->solveRefined(aHost, aDev, B, ldb, nrhs, X, ldx, dev, maxIt, tol);
    if (info) {
      info->iterations = r.iterations;
      info->initial_relative_residual = r.initialRel;
      info->final_relative_residual = r.finalRel;
      info->backward_error = r.backwardError;
    }
  });
}

int symla_solve_refined(symla_plan* p, const double* a_vals, const double* B, int64_t ldb, int nrhs, double* X,
                        int64_t ldx, const symla_refine_opts* opts, symla_refine_info* info) {
  return refinedImpl(p, a_vals, nullptr, B, ldb, nrhs, X, ldx, false, opts, info);
}

int symla_solve_refined_device(symla_plan* p, const double* d_a_vals, const double* dB, int64_t ldb, int nrhs,
                               double* dX, int64_t ldx, const symla_refine_opts* opts, symla_refine_info* info) {
  return refinedImpl(p, nullptr, d_a_vals, dB, ldb, nrhs, dX, ldx, true, opts, info);
}

int symla_front_count(const symla_plan* p) { return p ? p->plan->numSN() : -1; }

int symla_front_view(symla_plan* p, int si, int* m, int* npiv, int* rows, double* L, double* d_diag, double* d_off,
                     int* pivot_start, int* pivot_kind, int* nblocks) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_front_view: null plan");
    p->plan->frontView(si, m, npiv, rows, L, d_diag, d_off, pivot_start, pivot_kind, nblocks);
  });
}

int symla_plan_set_profiling(symla_plan* p, int enable) {
  return planGuarded(p, [&] {
    if (!p) throw std::invalid_argument("symla_plan_set_profiling: null plan");
    p->plan->setProfiling(enable != 0);
  });
}

int symla_plan_get_profile(symla_plan* p, symla_profile* out) {
  return planGuarded(p, [&] {
    if (!p || !out) throw std::invalid_argument("symla_plan_get_profile: null argument");
    const auto r = p->plan->takeProfile();
    out->syrk_ms = r.ms[Plan::kProfSyrk];
    out->syrk_launches = r.launches[Plan::kProfSyrk];
    out->schur_flops = r.schurFlops;
    out->panel_ms = r.ms[Plan::kProfPanel];
    out->panel_launches = r.launches[Plan::kProfPanel];
    out->panel_rows_ms = r.ms[Plan::kProfPanelRows];
    out->panel_rows_launches = r.launches[Plan::kProfPanelRows];
    out->small_ms = r.ms[Plan::kProfSmall];
    out->small_launches = r.launches[Plan::kProfSmall];
    out->assemble_ms = r.ms[Plan::kProfAssemble];
    out->assemble_launches = r.launches[Plan::kProfAssemble];
    out->fwd_ms = r.ms[Plan::kProfFwd];
    out->fwd_launches = r.launches[Plan::kProfFwd];
    out->bwd_ms = r.ms[Plan::kProfBwd];
    out->bwd_launches = r.launches[Plan::kProfBwd];
    out->cb_ms = r.ms[Plan::kProfCb];
    out->cb_launches = r.launches[Plan::kProfCb];
    out->cb_flops = r.cbFlops;
  });
}

int symla_plan_timing(const symla_plan* p, double* factor_ms, double* solve_ms) {
  if (!p) return fail(SYMLA_ERR_INVALID, "symla_plan_timing: null plan");
  if (factor_ms) *factor_ms = p->plan->lastFactorMs();
  if (solve_ms) *solve_ms = p->plan->lastSolveMs();
  return SYMLA_OK;
}

// One dense front through the device kernels (DenseLDLT::factor /
// factorStatic test entry).
int symla_dense_factor(int n, double* A, int lda, int n_eligible, const symla_dense_opts* opts, double* d_diag,
                       double* d_off, int* perm, int* pivot_start, int* pivot_kind, int* n_pivot_blocks,
                       int* n_factored, int* inertia3, double* total_perturbation, int* num_perturbed) {
  return guarded([&] {
    if (n < 0 || lda < n) throw std::invalid_argument("DenseLDLT::factor: A must be square");
    const int nE = (n_eligible < 0) ? n : n_eligible;
    if (nE < 0 || nE > n) throw std::invalid_argument("DenseLDLT::factor: n_eligible out of range");
    if (n == 0) {
      if (n_factored) *n_factored = 0;
      if (n_pivot_blocks) *n_pivot_blocks = 0;
      if (inertia3) inertia3[0] = inertia3[1] = inertia3[2] = 0;
      return;
    }
    hipStream_t st;
    hipCheck(hipStreamCreate(&st), "stream");
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    auto upI = [&](DBuf<int>& b, const std::vector<int>& v) { b.upload(v, st); };
    DBuf<double> dA;
    dA.alloc(static_cast<size_t>(n) * n);
    hipCheck(hipMemcpy2DAsync(dA.p, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                              hipMemcpyHostToDevice, st),
             "A upload");
    DBuf<int> dm, dnE, dnc, dfirst, dfi, drel, dcp, dcl, daRow, dperm, drow, dpk, dnF, dinert, dkcur, dpc, dstop,
        dnpert, dlist, dsign, dcb, dcut;
    DBuf<int64_t> dio, dacp, daSrc;
    DBuf<double> ddd, ddo, dpert, dws;
    DBuf<void*> dF;
    DBuf<unsigned long long> dom, drm;
    upI(dm, {n});
    upI(dnE, {nE});
    upI(dnc, {nE});
    upI(dfirst, {0});
    const bool deferCb = opts && opts->deferred_cb && nE >= 1 && nE < n && n > kSmallFrontMax;
    const std::vector<int> cutv{deferCb ? nE : n};
    upI(dcut, cutv);
    std::vector<int> ident(n);
    for (int i = 0; i < n; ++i) ident[i] = i;
    upI(dfi, ident);
    upI(drel, std::vector<int>(n, -1));
    upI(dcp, {0, 0});
    dcl.alloc(1);
    daRow.alloc(1);
    daSrc.alloc(1);
    dio.upload(std::vector<int64_t>{0}, st);
    dacp.upload(std::vector<int64_t>(n + 1, 0), st);
    dperm.alloc(n);
    drow.alloc(n);
    dpk.alloc(n);
    upI(dpk, std::vector<int>(n, 0));
    ddd.alloc(n);
    ddo.alloc(n);
    dnF.alloc(1);
    dinert.alloc(3);
    dkcur.alloc(1);
    dpc.alloc(1);
    dstop.alloc(1);
    dcb.alloc(1);
    dnpert.alloc(1);
    dpert.alloc(1);
    dom.alloc(1);
    drm.alloc(1);
    dws.alloc(static_cast<size_t>(panelWorkspace(n, nE, !(opts && opts->exact_panel))));
    dF.upload(std::vector<void*>{dA.p}, st);
    upI(dlist, {0});
    const int* signPtr = nullptr;
    if (opts && opts->static_mode && opts->expected_sign) {
      upI(dsign, std::vector<int>(opts->expected_sign, opts->expected_sign + n));
      signPtr = dsign.p;
    }
    DevFronts f;
    f.m = dm.p;
    f.nE = dnE.p;
    f.ncols = dnc.p;
    f.F = dF.p;
    f.fp32 = 0;
    f.idxOff = dio.p;
    f.frontIdx = dfi.p;
    f.relmap = drel.p;
    f.childPtr = dcp.p;
    f.childList = dcl.p;
    f.firstCol = dfirst.p;
    f.acolPtr = dacp.p;
    f.aRow = daRow.p;
    f.aSrc = daSrc.p;
    f.schurCut = dcut.p;
    f.perm = dperm.p;
    f.rowIdx = drow.p;
    f.dDiag = ddd.p;
    f.dOff = ddo.p;
    f.pkind = dpk.p;
    f.nF = dnF.p;
    f.inertia = dinert.p;
    f.origMaxBits = dom.p;
    f.runMaxBits = drm.p;
    f.kcur = dkcur.p;
    f.pcount = dpc.p;
    f.stopped = dstop.p;
    f.cbDone = dcb.p;
    f.pert = dpert.p;
    f.npert = dnpert.p;
    f.schurFlops = nullptr;
    FactorParams prm;
    prm.zeroTol = opts ? opts->zero_tolerance : 1e-300;
    prm.relFloor = opts ? opts->relative_pivot_floor : 1e-12;
    prm.staticMode = (opts && opts->static_mode) ? 1 : 0;
    prm.staticDelta = opts ? opts->static_delta : 0.0;
    prm.staticRelFloor = opts ? opts->static_rel_floor : 1e-8;
    prm.staticAbsFloor = opts ? opts->static_abs_floor : 1e-300;
    prm.expectedSign = signPtr;
    prm.specPanel = (opts && opts->exact_panel) ? 0 : 1;
    prm.specMinRows = 0;  // every panel of the dense entry takes the speculative step
    launchResetFronts(1, f, st);
    launchResetPerm(dlist.p, 1, f, st);
    if (n <= kSmallFrontMax) {
      launchFactorSmall(dlist.p, 1, n, f, prm, st);
    } else {
      std::vector<int> mv{n}, nev{nE};
      LargeSchedHost lh;
      LargeSched ls = buildLargeSched({0}, mv, nev, cutv, lh, (opts && opts->exact_panel) ? -1 : 0);
      DBuf<int> dl, dr, dt, dtp, dmw, dsp, derr;
      DBuf<int64_t> dw;
      DBuf<MwSync> dsync;
      dl.upload(lh.lists, st);
      dr.upload(lh.rowStart, st);
      dt.upload(lh.tileStart, st);
      dtp.upload(lh.tileTp, st);
      dmw.upload(lh.mwStart, st);
      dsp.upload(lh.specStart, st);
      dw.upload(std::vector<int64_t>{0}, st);
      dsync.alloc(kMwSlots);
      derr.alloc(1);
      hipCheck(hipMemsetAsync(dsync.p, 0, sizeof(MwSync) * kMwSlots, st), "memset");
      hipCheck(hipMemsetAsync(derr.p, 0, sizeof(int), st), "memset");
      LargeDev dv;
      dv.lists = dl.p;
      dv.rowStart = dr.p;
      dv.tileStart = dt.p;
      dv.tileTp = dtp.p;
      dv.mwStart = dmw.p;
      dv.specStart = dsp.p;
      dv.sync = dsync.p;
      dv.err = derr.p;
      // operand workspace of the deferred contribution-block update
      DBuf<double> dW;
      const std::vector<void*> Fv{dA.p};
      CbUpdate cb;
      if (!ls.cb.empty()) {
        dW.alloc(static_cast<size_t>(n - nE) * kCbChunk);
        cb.w = dW.p;
        cb.wElems = dW.n;
        cb.F = &Fv;
      }
      runLargeSched(ls, dv, f, prm, dws.p, dw.p, st, &cb, nullptr);
      int herr = 0;
      hipCheck(hipMemcpyAsync(&herr, derr.p, sizeof(int), hipMemcpyDeviceToHost, st), "error word");
      hipCheck(hipStreamSynchronize(st), "dense large front");
      if (herr != 0) throw DeviceError("symla: a cooperative panel step exceeded its bounded wait");
    }
    hipCheck(hipMemcpy2DAsync(A, sizeof(double) * lda, dA.p, sizeof(double) * n, sizeof(double) * n, n,
                              hipMemcpyDeviceToHost, st),
             "A download");
    std::vector<int> hperm(n), hpk(n);
    std::vector<double> hdd(n), hdo(n);
    int hnF = 0, hin[3] = {0, 0, 0}, hnp = 0;
    double hpert = 0.0;
    hipCheck(hipMemcpyAsync(hperm.data(), dperm.p, sizeof(int) * n, hipMemcpyDeviceToHost, st), "perm");
    hipCheck(hipMemcpyAsync(hpk.data(), dpk.p, sizeof(int) * n, hipMemcpyDeviceToHost, st), "pkind");
    hipCheck(hipMemcpyAsync(hdd.data(), ddd.p, sizeof(double) * n, hipMemcpyDeviceToHost, st), "d");
    hipCheck(hipMemcpyAsync(hdo.data(), ddo.p, sizeof(double) * n, hipMemcpyDeviceToHost, st), "d");
    hipCheck(hipMemcpyAsync(&hnF, dnF.p, sizeof(int), hipMemcpyDeviceToHost, st), "nF");
    hipCheck(hipMemcpyAsync(hin, dinert.p, 3 * sizeof(int), hipMemcpyDeviceToHost, st), "inertia");
    hipCheck(hipMemcpyAsync(&hpert, dpert.p, sizeof(double), hipMemcpyDeviceToHost, st), "pert");
    hipCheck(hipMemcpyAsync(&hnp, dnpert.p, sizeof(int), hipMemcpyDeviceToHost, st), "npert");
    hipCheck(hipStreamSynchronize(st), "dense factor");
    int nb = 0;
    for (int i = 0; i < n; ++i) {
      if (perm) perm[i] = hperm[i];
      if (d_diag) d_diag[i] = (i < hnF) ? hdd[i] : 0.0;
      if (d_off) d_off[i] = (i < hnF && hpk[i] == 2) ? hdo[i] : 0.0;
      if (i < hnF && (hpk[i] == 1 || hpk[i] == 2)) {
        if (pivot_start) pivot_start[nb] = i;
        if (pivot_kind) pivot_kind[nb] = hpk[i];
        ++nb;
      }
    }
    if (n_pivot_blocks) *n_pivot_blocks = nb;
    if (n_factored) *n_factored = hnF;
    if (inertia3) {
      inertia3[0] = hin[0];
      inertia3[1] = hin[1];
      inertia3[2] = hin[2];
    }
    if (total_perturbation) *total_perturbation = hpert;
    if (num_perturbed) *num_perturbed = hnp;
  });
}

}  // extern "C"
