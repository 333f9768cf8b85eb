// plan.hpp -- host-side owner of the device pool and the level schedule of
// the multifrontal numeric phase.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "device.hpp"
#include "symbolic.hpp"

namespace symla_hip {

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hipCheck(hipError_t e, const char* what);

// Moves on every allocation or release of plan device memory (DBuf, the
// per-front blocks of a repair pass, grown index arrays).  A recorded graph
// refers to device addresses, so it is stale once this has moved since its
// recording.  Process-wide: another plan's allocations also count, which
// only costs an extra re-record.
inline std::atomic<unsigned long long> g_devAllocEpoch{0};

// Device buffer with RAII free.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      ++g_devAllocEpoch;
    }
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipCheck(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)), "hipMalloc");
    ++g_devAllocEpoch;
    n = count;
  }
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  // Returns once the copy is done: callers may free or reuse the host vector
  // right away (set-up paths only; never under stream capture).
  void upload(const std::vector<T>& v, hipStream_t st) {
    ensure(v.size());
    if (!v.empty()) {
      hipCheck(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st), "upload");
      hipCheck(hipStreamSynchronize(st), "upload");
    }
  }
};

struct FactorOptions {
  int mode = 0;
  int precision = 0;  // 0: FP64 fronts; 1: FP32 fronts (FP64 pivots, D and refinement)
  bool bf16x3 = false;  // FP32 fronts whose deferred contribution-block updates use BF16x3 split products
  double zeroTol = 1e-300;
  double relFloor = 1e-12;
  std::vector<int> expectedSignOrig;  // empty = none
  bool haveExpectation = false;
  double initialDelta = 0.0;
  int maxRetries = 10;
  double growth = 10.0;
  double staticRelFloor = 1e-8;
  double staticAbsFloor = 1e-300;
};

struct FactorInfo {
  int pos = 0, neg = 0, zero = 0;
  bool singular = false;
  int nSingular = 0;
  double totalPerturbation = 0.0;
  int numPerturbed = 0;
  int retriesUsed = 0;
  double lastDelta = 0.0;
  int delayedFronts = 0;
  bool repaired = false;
  int precisionUsed = 0;  // front precision of the returned factorization
};

struct RefineResult {
  int iterations = 0;
  double initialRel = 0.0, finalRel = 0.0, backwardError = 0.0;
};

// Batched schedule of the large-front path for one set of independent fronts
// (one tree level): per panel iteration, the fronts that still have eligible
// columns, their row blocks and their trailing-update tiles.  Fronts with
// more than kMwThreshold eligible rows left take the cooperative panel
// (several workgroups per front); the others one workgroup each.
constexpr int kMwThreshold = 4096;
// Fronts with more eligible rows than this take the speculative panel step by default.
constexpr int kSpecMinRowsDefault = 2048;
// plans with a front of more eligible rows than this default to the
// speculative step on every panel front
constexpr int kSpecAllFrontsNE = 4096;
constexpr int kMwMaxBlocks = 240;  // cooperative workgroups per launch (all resident at once)
struct LargeIter {
  int64_t listOff = 0;  // all fronts of the iteration (rows, trailing update)
  int count = 0;
  int64_t rowOff = 0;
  int rowBlocks = 0;
  int64_t tileOff = 0;
  int tiles = 0;
  int64_t regOff = 0;  // single-workgroup panels
  int regCount = 0;
  int regMaxNE = 0;    // widest eligible-row count among them
  int64_t mwOff = 0;   // cooperative panels: fronts, and workgroup starts at mwStartOff
  int mwCount = 0;
  int64_t mwStartOff = 0;
  int mwBlocks = 0;
  int64_t specOff = 0;  // speculative step: workgroup starts of the iteration's fronts
  int specBlocks = 0;
};
// Deferred contribution-block updates: a large front with at least
// kCbMinCols eligible columns and a contribution block of at least the
// tuned row count (Tuning::cbMinRows) stops its panels' trailing updates at
// column nE; after its last panel the block [nE, m)^2 takes one rank-nE
// update, kCbChunk pivot columns per MFMA SYRK launch (cb_syrk.hip).
constexpr int kCbMinCols = 256;
constexpr int kCbMinRowsDefault = 2048;
constexpr int kCbChunk = 2048;
// fronts with at most this many 32-row own blocks solve their own rows on
// one workgroup; longer chains take one workgroup per block
constexpr int kOwnChainBlocks = 10;
// the backward ancestors' product: chunks of at least this many rows, split
// until a level has about kAncLevelBlocks workgroups, at most kAncMaxSplit
constexpr int kAncChunkRows = 512;
constexpr int kAncLevelBlocks = 2048;
constexpr int kAncMaxSplit = 16;
struct CbTask {
  int s, m, nE;
};
struct LargeSched {
  int64_t allOff = 0;
  int allCount = 0;
  int maxM = 0;
  std::vector<LargeIter> iters;
  std::vector<CbTask> cb;  // fronts whose contribution block is updated after their last panel
};
struct LargeSchedHost {
  std::vector<int> lists, rowStart, tileStart, tileTp, mwStart, specStart;
};
// Device copies of a LargeSchedHost plus the cooperative-panel records.
struct LargeDev {
  const int* lists = nullptr;
  const int* rowStart = nullptr;
  const int* tileStart = nullptr;
  const int* tileTp = nullptr;
  const int* mwStart = nullptr;
  const int* specStart = nullptr;
  MwSync* sync = nullptr;
  int* err = nullptr;
  bool frontMaxDone = false;  // the assembly already folded max |F| into origMaxBits
};
// What the deferred contribution-block updates run with: a workspace of at
// least kCbChunk times the largest contribution-block order (elements of the
// front type) and the fronts' device pointers by front id.
struct CbUpdate {
  void* w = nullptr;
  size_t wElems = 0;
  const std::vector<void*>* F = nullptr;
  bool fp32 = false;
  bool bf16x3 = false;  // FP32 fronts: BF16x3 split products instead of FP32 MFMA
};
// m, nE and cut are indexed by front id (cut[s] < m[s]: the contribution
// block is deferred); allowMw = false keeps every panel on one workgroup.
// specMinRows < 0: no speculative step; else fronts with nE > specMinRows get its workgroups.
LargeSched buildLargeSched(const std::vector<int>& fronts, const std::vector<int>& m, const std::vector<int>& nE,
                           const std::vector<int>& cut, LargeSchedHost& h, int specMinRows, bool allowMw = true);
class Plan;
// Launch sequence of a LargeSched; `prof` (may be null) brackets launches
// with events; `cb` may be null only when no front defers its contribution block.
void runLargeSched(const LargeSched& ls, const LargeDev& dv, const DevFronts& fd, const FactorParams& prm, double* ws,
                   const int64_t* wsOff, hipStream_t st, const CbUpdate* cb, Plan* prof);
// noalloc.hip: the static solve scratch of `device` (its device address,
// looked up once per device); kNoAllocScratchBytes is its size.
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
double* noAllocScratch(int device);
// cb_syrk.hip: C -= W B^T over the lower triangle of the N x N block C
// (column-major; W, B: N x K) on the matrix cores, in the front precision
void launchCbSyrk(bool fp32, int N, int K, const void* W, int ldw, const void* B, int ldb, void* C, int ldc,
                  hipStream_t st);
// the same on FP32 fronts as BF16x3 split products accumulated in FP32
void launchCbSyrkBf16x3(int N, int K, const float* W, int ldw, const float* B, int ldb, float* C, int ldc,
                        hipStream_t st);

class Plan {
 public:
  Plan(const SymbolicResult& sym, const int64_t* colptr, const int* rowidx);
  ~Plan();

  // values: device pointer (already resident) after `setValuesDevice`.
  void factorize(const double* aValsHostOrNull, const double* aValsDev, const FactorOptions& opts, FactorInfo& info);
  void solve(const double* B, int64_t ldb, int nrhs, double* X, int64_t ldx, bool deviceIo);
  // Refinement of a batch of independent systems (a block-diagonal matrix
  // whose system g owns the original indices [groupStart[g],
  // groupStart[g+1])): refine.hpp's loop applied to each system on its own --
  // own norms, own stopping tests, own iteration count -- while the
  // correction solves run on the whole batch.  Returns one result per system.
  std::vector<RefineResult> solveRefinedGroups(const double* aHost, const double* aDev, const double* B, int64_t ldb,
                                               int nrhs, double* X, int64_t ldx, bool deviceIo, int maxIt,
                                               double tol, int ngroups, const int* groupStart);
  // The device solve with its per-call workspace in `scratch` (capacity in
  // doubles) instead of plan buffers: allocates nothing.  Throws
  // std::invalid_argument when (3 n + solve work rows) x nrhs doubles
  // exceed the capacity.
  void solveNoAlloc(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx, double* scratch,
                    size_t capacity);
  RefineResult solveRefined(const double* aHost, const double* aDev, const double* B, int64_t ldb, int nrhs,
                            double* X, int64_t ldx, bool deviceIo, int maxIt, double tol);

  int n() const { return n_; }
  int numSN() const { return numSN_; }
  bool factorized() const { return factorized_; }
  bool singular() const { return singular_; }
  const std::vector<int>& singularColsOrig() const { return singularColsOrig_; }
  // Inertia and singular-column count per group of original indices
  // [groupStart[g], groupStart[g+1]) (independent systems of a block-diagonal
  // matrix: every front lies inside one of them).
  void groupInertia(int ngroups, const int* groupStart, int* inertia3, int* singular);

  // hipGraph of one factorization (threshold pivoting, static schedule) and
  // one solve on fixed device buffers.  Replay re-runs both on the current
  // buffer contents; a replay whose factorization delayed pivots repairs it
  // on the regular path and solves again.  The graph is re-recorded when a
  // buffer it refers to has moved.
  void graphCapture(const double* dVals, const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx,
                    const FactorOptions& opts);
  // The same with FP64 iterative refinement (refine.hpp:97-144): the graph
  // holds the factorization, the norms, the first solve and its residual;
  // two more graphs hold one refinement step each (ping-pong buffers), and
  // replay launches them while the host's refine.hpp tests ask for more.
  void graphCaptureRefined(const double* dVals, const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx,
                           const FactorOptions& opts, int maxIt, double tol);
  void graphReplay(FactorInfo& info, RefineResult* rres = nullptr);
  void graphRelease();
  bool hasGraph() const { return graphExec_ != nullptr; }

  // Scheduling choices: fronts of order <= smallMax take the LDS kernel,
  // larger ones the panel path; assembly runs in column tiles of asmTile.
  struct Tuning {
    // fronts up to this order are factored whole in LDS; larger ones (also the
    // few mid-size fronts of the upper levels, where one workgroup per front
    // is latency-bound) by the panel path
    int smallMax = kSmallFrontDefault;
    int asmTile = 64;
    int specPanel = 1;
    int specMinRows = kSpecMinRowsDefault;
    int cbMinRows = kCbMinRowsDefault;  // < 0: every front updates its contribution block panel by panel
  };
  struct TuneTrial {
    Tuning t;
    double ms = 0.0;
  };
  const Tuning& tuning() const { return tune_; }
  void setTuning(const Tuning& t);
  // Times factorize + solve (values at dVals, nrhs pseudo-random right-hand
  // sides) by coordinate descent over the scheduling knobs (small-front
  // bound, assembly tile, speculative panel row bound or off, deferred
  // contribution-block row bound or off), keeps the fastest and returns every
  // trial.  With cachePath, a tuning stored there for the same key (GPU
  // architecture, precision, nrhs, histogram of front orders) is applied
  // without timing (*fromCache = true, no trials), and a new result is
  // appended to the file.
  std::vector<TuneTrial> autotune(const double* dVals, int nrhs, int precision, int reps,
                                  const char* cachePath = nullptr, bool* fromCache = nullptr);
  std::string tuneKey(int nrhs, int precision) const;
  void frontView(int si, int* m, int* npiv, int* rows, double* L, double* dd, double* dof, int* ps, int* pk,
                 int* nblocks);
  double lastFactorMs() const { return factorMs_; }
  double lastSolveMs() const { return solveMs_; }
  std::string lastError;

  // per-kernel-class event profile
  enum ProfCat {
    kProfSyrk = 0,
    kProfPanel,
    kProfPanelRows,
    kProfSmall,
    kProfAssemble,
    kProfFwd,
    kProfBwd,
    kProfCb,
    kNumCats
  };
  struct Profile {
    double ms[kNumCats] = {};
    long long launches[kNumCats] = {};
    double schurFlops = 0.0;  // the panels' trailing updates (k_schur)
    double cbFlops = 0.0;     // the deferred contribution-block updates (planned pivots)
  };
  void addCbFlops(double f) { profAcc_.cbFlops += f; }
  void setProfiling(bool on);
  Profile takeProfile();
  void profBegin(int cat);
  void profEnd(int cat);

 private:
  // symbolic
  int n_ = 0, numSN_ = 0;
  std::vector<int> perm_;
  std::vector<Supernode> sn_;
  std::vector<int> parentSN_, roots_;
  std::vector<std::vector<int>> children_;
  std::vector<int> level_;
  std::vector<std::vector<int>> levelFronts_;
  std::vector<int> colToSN_;
  // A pattern
  int64_t nnzA_ = 0;
  int64_t nnzSym_ = 0;  // entries of the symmetric CSR (both triangles)
  std::vector<int64_t> aColptr_;
  std::vector<int> aRowidx_;
  // host layout mirrors
  std::vector<int> hm_, hnE_, hncols_, hfirst_, hchildPtr_, hchildList_;
  std::vector<int64_t> hidxOff_, hacolPtr_, haSrc_;
  std::vector<int> hfrontIdx_, hrelmap_, haRow_;
  int64_t idxUsed_ = 0;
  std::vector<int> hnF_, hrowIdx_;  // host copies after factorization (dynamic mode / views)
  // device layout
  hipStream_t stream_ = nullptr;
  DBuf<double> pool_;      // front storage (FP64, or FP32 packed in the same bytes)
  std::vector<void*> hF_;
  std::vector<void*> dynAllocs_;
  DBuf<void*> dF_;
  DBuf<int> dm_, dnE_, dncols_, dfirst_, dchildPtr_, dchildList_, dfrontIdx_, drelmap_, daRow_;
  DBuf<int64_t> didxOff_, dacolPtr_, daSrc_;
  DBuf<int> dperm_, drowIdx_, dpkind_, dnF_, dinertia_, dkcur_, dpcount_, dstopped_, dnpert_, dcbDone_;
  DBuf<double> ddDiag_, ddOff_, dpert_;
  DBuf<unsigned long long> dorigMax_, drunMax_;
  DBuf<long long> dstatI_;
  DBuf<double> dstatD_;
  DBuf<double> dvals_;
  DBuf<double> dws_;
  DBuf<int> dsignFinal_;
  // static schedule
  struct LevelSched {
    int64_t smallOff = 0;
    int smallCount = 0;
    int smallMaxM = 0;
    int64_t taskOff = 0;
    int taskCount = 0;
    std::vector<int> large;
    LargeSched lsched;
  };
  std::vector<LevelSched> sched_;
  DBuf<int> dsmall_;
  DBuf<int4> dtasks_;
  DBuf<int> dallFronts_;
  DBuf<int> dlevelFronts_;
  std::vector<int64_t> levelFrontOff_;
  int maxNE_ = 0;
  // solve
  DBuf<int> dpermVec_;
  DBuf<double> dBp_, dZ_, dX_, dDl_, dBin_, dXout_;
  DBuf<int64_t> dgOff_;  // running sum of the front orders: front s's work blocks start at dgOff_[s] * nrhs
  int64_t sumM_ = 0;
  // work rows of a solve: the fronts' blocks (sumM_) then the split
  // ancestors' products' partial rows (dancPart_, one entry per front)
  int64_t workRows_ = 0;
  DBuf<AncPart> dancPart_;
  bool solveLayoutReady_ = false;
  // forward gather map: per front position (idxOff-based), the children's
  // delta entries (child, child position) that sum into it, in child order
  DBuf<int64_t> dmapPtr_;
  DBuf<int64_t> dmapRow_;  // the entries' child work rows (gOff[child] + child position)
  // inverses of the sweeps' diagonal blocks: fronts above kTinyFront rows,
  // their flat workgroup starts (one per kSolveBlock eligible rows) and
  // per-front offsets into dinv_
  DBuf<int> dinvList_, dinvStart_;
  DBuf<int64_t> dinvOff_;
  DBuf<double> dinv_;
  int invCount_ = 0, invBlocks_ = 0;
  void launchInverse();
  // Per level, the solve's launches: the gather over every front (positions
  // (front, row) at dpos_ + posOff),
  // fronts whose own rows take one workgroup (own), fronts whose own rows
  // take one workgroup per block (chain), fronts with rows past their own
  // (delta), every front above kTinyFront rows (anc: the backward ancestors'
  // part), and fronts of at most kTinyFront rows (tiny).  Lists are offsets
  // into dsolveLists_, workgroup starts offsets into dsolveStarts_.
  struct SolveLaunch {
    int64_t list = 0, start = 0;
    int count = 0, blocks = 0;
  };
  struct SolveLevel {
    int64_t posOff = 0, positions = 0;
    SolveLaunch own, chain, delta, anc, tiny;
  };
  std::vector<SolveLevel> solveLv_;
  DBuf<int> dsolveLists_, dsolveStarts_;
  DBuf<int2> dpos_;
  DBuf<PosInfo> dposInfo_;  // the gather's records of the positions in dpos_, formed per solve
  int64_t posTotal_ = 0;
  // dependency flags of the sweeps (one per own-row block), the solve
  // generation they are compared with, and the bounded-wait error word
  DBuf<int64_t> dflagOff_;
  DBuf<unsigned> dflagsF_, dflagsB_, dgen_;
  DBuf<int> dsolveErr_;
  void checkSolveError();
  // refinement CSR
  DBuf<int64_t> dsymPtr_, dsymSrc_, dstPtr_, dstSrc_;
  DBuf<int> dsymCol_;
  DBuf<double> dsymVal_;  // symmetric CSR values of the matrix being refined against
  DBuf<double> dR_, dDx_, dXn_, dRn_, dBr_, dXt_;
  DBuf<unsigned long long> dnorms_;
  // per-system refinement of a batch: row -> system map (for the group
  // boundaries in hgroupStart_), per-system norm slots, accept flags
  std::vector<int> hgroupStart_;
  DBuf<int> drowGroup_, dgAccept_;
  DBuf<unsigned long long> dgNorms_;
  // state
  bool factorized_ = false;
  bool singular_ = false;
  std::vector<int> singularColsOrig_;
  bool layoutDirty_ = false;  // dynamic mode changed the layout
  int prec_ = 0;              // precision of the current front storage
  bool cbBf16x3_ = false;     // FP32 fronts: deferred contribution-block updates as BF16x3 products
  // SYMLA_PREC_* of the current factorization (2: FP32 fronts with BF16x3 contribution-block updates)
  int precisionCode() const { return prec_ == 0 ? 0 : (cbBf16x3_ ? 2 : 1); }
  size_t elemSize() const { return prec_ ? sizeof(float) : sizeof(double); }
  double factorMs_ = 0.0, solveMs_ = 0.0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  // recorded refactorize + solve
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graphExec_ = nullptr;
  struct GraphArgs {
    const double* vals = nullptr;
    const double* B = nullptr;
    int64_t ldb = 0;
    int nrhs = 0;
    double* X = nullptr;
    int64_t ldx = 0;
    FactorOptions opts;
    bool refine = false;
    int maxIt = 10;
    double tol = 1e-14;
  } graphArgs_;
  hipGraphExec_t stepExec_[2] = {nullptr, nullptr};  // recorded refinement steps
  // g_devAllocEpoch when the graph was recorded; graphStale_ forces a
  // re-record (new schedule)
  unsigned long long graphEpoch_ = 0;
  bool graphStale_ = true;
  Tuning tune_;
  bool graphCurrent() const { return !graphStale_ && graphEpoch_ == g_devAllocEpoch.load(); }
  void graphRecord();
  bool profiling_ = false;
  struct ProfEvents {
    std::vector<hipEvent_t> ev;
    size_t used = 0;
  };
  ProfEvents prof_[kNumCats];
  Profile profAcc_;
  DBuf<unsigned long long> dschur_;
  void profHarvest();  // after a stream sync: fold recorded event pairs into profAcc_

  void buildStaticLayout();
  void uploadLayout();
  void buildSchedule();
  void buildRefineCsr();
  DevFronts devFronts();
  FactorParams params(const FactorOptions& o, bool staticMode, double delta, const int* signDev) const;
  void runStatic(const double* aValsDev, const FactorParams& prm);
  void runDynamic(const double* aValsDev, const FactorParams& prm);
  DBuf<int> dlargeLists_, dlargeRowStart_, dlargeTileStart_, dlargeTileTp_, dlargeMwStart_, dlargeSpecStart_;
  DBuf<MwSync> dmwSync_;
  DBuf<int> dfactErr_;  // cooperative-panel wait errors
  LargeDev largeDev(const int* lists, const int* rowStart, const int* tileStart, const int* tileTp,
                    const int* mwStart, const int* specStart);
  // deferred contribution-block updates: per-front trailing-update cut
  // (host copy from the current sizes and tuning), operand workspace
  std::vector<int> hcut_;
  DBuf<int> dschurCut_;
  DBuf<double> dcbW_;
  void computeCut();
  // workspace for the contribution blocks of `fronts` (outside any capture)
  CbUpdate cbUpdate(const std::vector<CbTask>& tasks);
  DBuf<int64_t> dwsOff_;
  std::vector<int64_t> hwsOff_;
  void collectStats(FactorInfo& info, bool& anyDelay);
  void finishSingular();
  // nrhs-independent solve set-up (gather map, launch shapes, flags) for
  // the current layout; built after each factorization
  void ensureSolveLayout();
  void ensureSolveWs(int nrhs);
  // per-call solve workspace: permuted B, Z, X (n x nrhs each, row-major) and
  // the work rows Dl (workRows_ x nrhs, row-major)
  struct SolveWs {
    double *Bp = nullptr, *Z = nullptr, *X = nullptr, *Dl = nullptr;
  };
  // ext: caller-provided workspace (nothing is allocated); null: the plan's own
  void solveCore(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx, const SolveWs* ext = nullptr);
  // refinement pieces (shared by the eager loop and the recorded graphs)
  struct RefineNorms {
    double normA = 0.0, normAsym = 0.0, normB = 0.0, rel = 0.0;
  };
  void ensureRefineWs(int nrhs);
  void refineSetup(const double* vals, int nrhs);
  void refineStep(const double* Xc, const double* Rc, double* Xn, double* Rn, int nrhs);
  RefineNorms readNorms();
  double backwardError(const double* R, const double* X, int nrhs, double normAsym);
  // iterations of refine.hpp:121-141 from the state refineSetup left (X in
  // dXout_, R in dR_); steps: two recorded refinement steps (X/R from
  // dXout_/dR_ into dXn_/dRn_ and back), or null to launch them eagerly.
  // Returns the buffer holding the accepted solution.
  const double* refineLoop(const RefineNorms& nm, int nrhs, int maxIt, double tol, hipGraphExec_t* steps,
                           RefineResult& res);
  void ensureIdxCapacity(int64_t need);
};

}  // namespace symla_hip
