// plan.cpp -- device pool, level schedule, factorization driver, solve and
// refinement of the MI355X numeric phase.
//
// Factorization runs level by level over the supernode tree (children before
// parents; multifrontal.hpp:13-18 makes supernode index order a postorder).
// The static schedule assumes no delayed pivots and runs every level without
// a host round trip.  One readback at the end tells whether some front
// delayed columns to its parent; only then a host-driven pass re-runs the
// factorization level by level, sizing each front with the delayed columns it
// actually receives (multifrontal.hpp:453-479).
#include "plan.hpp"

#include <chrono>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <sstream>
#include <stdexcept>

#include "kernels.hpp"

namespace symla_hip {

void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::ostringstream o;
    o << "HIP error in " << what << ": " << hipGetErrorString(e);
    throw DeviceError(o.str());
  }
}

namespace {
double bitsToDoubleHost(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof(d));
  return d;
}
}  // namespace

Plan::Plan(const SymbolicResult& sym, const int64_t* colptr, const int* rowidx) {
  n_ = sym.n;
  perm_ = sym.perm;
  sn_ = sym.supernodes;
  numSN_ = static_cast<int>(sn_.size());
  hipCheck(hipStreamCreate(&stream_), "hipStreamCreate");
  hipCheck(hipEventCreate(&ev0_), "hipEventCreate");
  hipCheck(hipEventCreate(&ev1_), "hipEventCreate");

  // supernode tree (multifrontal.hpp:312-332)
  colToSN_.assign(n_, -1);
  for (int s = 0; s < numSN_; ++s)
    for (int c = sn_[s].firstCol; c < sn_[s].firstCol + sn_[s].ncols; ++c) colToSN_[c] = s;
  parentSN_.assign(numSN_, -1);
  children_.assign(numSN_, {});
  for (int s = 0; s < numSN_; ++s) {
    const int last = sn_[s].firstCol + sn_[s].ncols - 1;
    const int p = sym.etreeParent[last];
    if (p != -1) {
      parentSN_[s] = colToSN_[p];
      children_[colToSN_[p]].push_back(s);
    }
  }
  for (int s = 0; s < numSN_; ++s)
    if (parentSN_[s] == -1) roots_.push_back(s);
  level_.assign(numSN_, 0);
  int maxLevel = 0;
  for (int s = 0; s < numSN_; ++s) {
    int l = 0;
    for (int c : children_[s]) l = std::max(l, level_[c] + 1);
    level_[s] = l;
    maxLevel = std::max(maxLevel, l);
  }
  levelFronts_.assign(numSN_ ? maxLevel + 1 : 0, {});
  for (int s = 0; s < numSN_; ++s) levelFronts_[level_[s]].push_back(s);

  // A pattern
  aColptr_.assign(colptr, colptr + n_ + 1);
  nnzA_ = aColptr_[n_];
  aRowidx_.assign(rowidx, rowidx + nnzA_);
  for (int64_t t = 0; t < nnzA_; ++t)
    if (aRowidx_[t] < 0 || aRowidx_[t] >= n_) throw std::invalid_argument("symla: row index out of range");

  buildStaticLayout();
  // Default speculative threshold from the largest front: with very large
  // fronts the mid-size ones carry long contribution blocks too, whose rows
  // the step maps in the same pass (measured faster on every panel front).
  {
    int maxNE = 0;
    for (int s = 0; s < numSN_; ++s) maxNE = std::max(maxNE, hnE_[s]);
    if (maxNE > kSpecAllFrontsNE) tune_.specMinRows = 0;
  }
  // per-front state
  dnF_.alloc(numSN_);
  dinertia_.alloc(3 * static_cast<size_t>(numSN_));
  dkcur_.alloc(numSN_);
  dpcount_.alloc(numSN_);
  dstopped_.alloc(numSN_);
  dcbDone_.alloc(numSN_);
  dnpert_.alloc(numSN_);
  dpert_.alloc(numSN_);
  dorigMax_.alloc(numSN_);
  drunMax_.alloc(numSN_);
  dstatI_.alloc(8);
  dstatD_.alloc(4);
  dvals_.alloc(nnzA_);
  uploadLayout();
  buildSchedule();
  buildRefineCsr();
  {
    std::vector<int> pv(perm_);
    dpermVec_.upload(pv, stream_);
  }
  dnorms_.alloc(16);
  dschur_.alloc(1);
  hipCheck(hipMemset(dschur_.p, 0, sizeof(unsigned long long)), "memset");
  // records of the cooperative panels (tags only grow; zero once)
  dmwSync_.alloc(kMwSlots);
  hipCheck(hipMemset(dmwSync_.p, 0, sizeof(MwSync) * kMwSlots), "memset");
  dfactErr_.alloc(1);
  hipCheck(hipMemset(dfactErr_.p, 0, sizeof(int)), "memset");
  hipCheck(hipStreamSynchronize(stream_), "plan setup");
}

void Plan::setProfiling(bool on) {
  profiling_ = on;
  profAcc_ = Profile{};
  for (auto& pc : prof_) pc.used = 0;
}

void Plan::profBegin(int cat) {
  if (!profiling_) return;
  auto& pc = prof_[cat];
  while (pc.ev.size() < pc.used + 2) {
    hipEvent_t e;
    hipCheck(hipEventCreate(&e), "hipEventCreate");
    pc.ev.push_back(e);
  }
  hipCheck(hipEventRecord(pc.ev[pc.used], stream_), "hipEventRecord");
}

void Plan::profEnd(int cat) {
  if (!profiling_) return;
  auto& pc = prof_[cat];
  hipCheck(hipEventRecord(pc.ev[pc.used + 1], stream_), "hipEventRecord");
  pc.used += 2;
  profAcc_.launches[cat] += 1;
}

void Plan::profHarvest() {
  if (!profiling_) return;
  hipCheck(hipStreamSynchronize(stream_), "profile sync");
  for (int c = 0; c < kNumCats; ++c) {
    auto& pc = prof_[c];
    for (size_t i = 0; i + 1 < pc.used; i += 2) {
      float ms = 0.f;
      hipCheck(hipEventElapsedTime(&ms, pc.ev[i], pc.ev[i + 1]), "hipEventElapsedTime");
      profAcc_.ms[c] += ms;
    }
    pc.used = 0;
  }
}

Plan::Profile Plan::takeProfile() {
  profHarvest();
  Profile out = profAcc_;
  profAcc_ = Profile{};
  return out;
}

Plan::~Plan() {
  graphRelease();
  for (auto& pc : prof_)
    for (hipEvent_t e : pc.ev) (void)hipEventDestroy(e);
  for (void* p : dynAllocs_) (void)hipFree(p);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Static layout: front s is [own columns | symbolic update rows] =
// rowPattern (symbolic.hpp:138-147), F_s is m x m column-major in the pool.
void Plan::buildStaticLayout() {
  hm_.resize(numSN_);
  hnE_.resize(numSN_);
  hncols_.resize(numSN_);
  hfirst_.resize(numSN_);
  hidxOff_.assign(numSN_ + 1, 0);
  for (int s = 0; s < numSN_; ++s) {
    hm_[s] = static_cast<int>(sn_[s].rowPattern.size());
    hnE_[s] = sn_[s].ncols;
    hncols_[s] = sn_[s].ncols;
    hfirst_[s] = sn_[s].firstCol;
    hidxOff_[s + 1] = hidxOff_[s] + hm_[s];
  }
  idxUsed_ = hidxOff_[numSN_];
  hfrontIdx_.assign(idxUsed_, 0);
  hrelmap_.assign(idxUsed_, -1);
  for (int s = 0; s < numSN_; ++s)
    std::copy(sn_[s].rowPattern.begin(), sn_[s].rowPattern.end(), hfrontIdx_.begin() + hidxOff_[s]);
  for (int c = 0; c < numSN_; ++c) {
    const int p = parentSN_[c];
    if (p < 0) continue;
    const auto& pp = sn_[p].rowPattern;
    for (int i = sn_[c].ncols; i < hm_[c]; ++i) {
      const int g = sn_[c].rowPattern[i];
      auto it = std::lower_bound(pp.begin(), pp.end(), g);
      if (it == pp.end() || *it != g) throw std::runtime_error("symla: child row missing from parent pattern");
      hrelmap_[hidxOff_[c] + i] = static_cast<int>(it - pp.begin());
    }
  }
  hchildPtr_.assign(numSN_ + 1, 0);
  hchildList_.clear();
  for (int s = 0; s < numSN_; ++s) {
    for (int c : children_[s]) hchildList_.push_back(c);
    hchildPtr_[s + 1] = static_cast<int>(hchildList_.size());
  }
  // A scatter map (PermutedLower::build, multifrontal.hpp:199-242): lower
  // entries of the original matrix, placed by their permuted column.
  std::vector<int> inv(n_);
  for (int k = 0; k < n_; ++k) inv[perm_[k]] = k;
  std::vector<int64_t> cnt(n_ + 1, 0);
  for (int c = 0; c < n_; ++c)
    for (int64_t t = aColptr_[c]; t < aColptr_[c + 1]; ++t) {
      const int r = aRowidx_[t];
      if (r < c) continue;
      ++cnt[std::min(inv[r], inv[c]) + 1];
    }
  hacolPtr_.assign(n_ + 1, 0);
  for (int k = 0; k < n_; ++k) hacolPtr_[k + 1] = hacolPtr_[k] + cnt[k + 1];
  std::vector<int64_t> pos(hacolPtr_.begin(), hacolPtr_.end() - 1);
  std::vector<int> hiRow(hacolPtr_[n_]);
  haSrc_.assign(hacolPtr_[n_], 0);
  for (int c = 0; c < n_; ++c)
    for (int64_t t = aColptr_[c]; t < aColptr_[c + 1]; ++t) {
      const int r = aRowidx_[t];
      if (r < c) continue;
      const int lo = std::min(inv[r], inv[c]), hi = std::max(inv[r], inv[c]);
      hiRow[pos[lo]] = hi;
      haSrc_[pos[lo]++] = t;
    }
  haRow_.assign(hacolPtr_[n_], 0);
  for (int k = 0; k < n_; ++k) {
    const int64_t b = hacolPtr_[k], e = hacolPtr_[k + 1];
    std::vector<int64_t> ord(e - b);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return hiRow[b + x] < hiRow[b + y]; });
    std::vector<int> hr(e - b);
    std::vector<int64_t> sr(e - b);
    for (size_t q = 0; q < ord.size(); ++q) {
      hr[q] = hiRow[b + ord[q]];
      sr[q] = haSrc_[b + ord[q]];
    }
    const int s = colToSN_[k];
    const auto& pat = sn_[s].rowPattern;
    for (size_t q = 0; q < ord.size(); ++q) {
      auto it = std::lower_bound(pat.begin(), pat.end(), hr[q]);
      if (it == pat.end() || *it != hr[q]) throw std::runtime_error("symla: A entry outside the front pattern");
      haRow_[b + q] = static_cast<int>(it - pat.begin());
      haSrc_[b + q] = sr[q];
    }
  }
  // pool
  size_t total = 0;
  std::vector<size_t> off(numSN_);
  for (int s = 0; s < numSN_; ++s) {
    off[s] = total;
    total += static_cast<size_t>(hm_[s]) * hm_[s];
  }
  // offsets in elements of the current precision; the pool is sized in doubles
  const size_t es = elemSize();
  const size_t poolDoubles = (total * es + sizeof(double) - 1) / sizeof(double);
  if (pool_.n < poolDoubles) pool_.alloc(poolDoubles);
  hF_.resize(numSN_);
  for (int s = 0; s < numSN_; ++s) hF_[s] = reinterpret_cast<char*>(pool_.p) + off[s] * es;
  maxNE_ = 0;
  for (int s = 0; s < numSN_; ++s) maxNE_ = std::max(maxNE_, hnE_[s]);
  dws_.ensure(2 * static_cast<size_t>(std::max(maxNE_, 1)));
  if (!dynAllocs_.empty()) ++g_devAllocEpoch;
  for (void* p : dynAllocs_) (void)hipFree(p);
  dynAllocs_.clear();
}

void Plan::ensureIdxCapacity(int64_t need) {
  const size_t cap = dperm_.n;
  if (static_cast<size_t>(need) <= cap) return;
  const size_t ncap = std::max<size_t>(need, 2 * cap);
  auto grow = [&](auto& buf) {
    using T = std::remove_pointer_t<decltype(buf.p)>;
    T* np = nullptr;
    hipCheck(hipMalloc(reinterpret_cast<void**>(&np), ncap * sizeof(T)), "hipMalloc grow");
    if (buf.p && idxUsed_ > 0)
      hipCheck(hipMemcpyAsync(np, buf.p, idxUsed_ * sizeof(T), hipMemcpyDeviceToDevice, stream_), "grow copy");
    hipCheck(hipStreamSynchronize(stream_), "grow sync");
    if (buf.p) (void)hipFree(buf.p);
    ++g_devAllocEpoch;
    buf.p = np;
    buf.n = ncap;
  };
  grow(dperm_);
  grow(drowIdx_);
  grow(dpkind_);
  grow(ddDiag_);
  grow(ddOff_);
}

void Plan::uploadLayout() {
  dm_.upload(hm_, stream_);
  dnE_.upload(hnE_, stream_);
  dncols_.upload(hncols_, stream_);
  dfirst_.upload(hfirst_, stream_);
  didxOff_.upload(hidxOff_, stream_);
  dfrontIdx_.upload(hfrontIdx_, stream_);
  drelmap_.upload(hrelmap_, stream_);
  dchildPtr_.upload(hchildPtr_, stream_);
  dchildList_.upload(hchildList_, stream_);
  dacolPtr_.upload(hacolPtr_, stream_);
  daRow_.upload(haRow_, stream_);
  daSrc_.upload(haSrc_, stream_);
  std::vector<void*> fp(hF_);
  dF_.upload(fp, stream_);
  const size_t cap = std::max<size_t>(static_cast<size_t>(idxUsed_), 1);
  if (dperm_.n < cap) {
    dperm_.alloc(cap);
    drowIdx_.alloc(cap);
    dpkind_.alloc(cap);
    ddDiag_.alloc(cap);
    ddOff_.alloc(cap);
  }
  computeCut();
  dschurCut_.upload(hcut_, stream_);
  // the solve's gather map, launch shapes and flags depend on the layout
  solveLayoutReady_ = false;
  // the uploads below read these host vectors until the stream sync at the end
  std::vector<int> lst, start(1, 0);
  std::vector<int64_t> off(numSN_ + 1, 0);
  {
    for (int s = 0; s < numSN_; ++s) {
      const int nb = hm_[s] > kTinyFront ? (hnE_[s] + kSolveBlock - 1) / kSolveBlock : 0;
      off[s + 1] = off[s] + static_cast<int64_t>(nb) * kSolveBlock * kSolveBlock;
      if (nb > 0) {
        lst.push_back(s);
        start.push_back(start.back() + nb);
      }
    }
    invCount_ = static_cast<int>(lst.size());
    invBlocks_ = start.back();
    dinvList_.upload(lst, stream_);
    dinvStart_.upload(start, stream_);
    dinvOff_.upload(off, stream_);
    dinv_.ensure(static_cast<size_t>(std::max<int64_t>(off[numSN_], 1)));
  }
  hipCheck(hipStreamSynchronize(stream_), "uploadLayout");
}

void Plan::computeCut() {
  hcut_.assign(numSN_, 0);
  for (int s = 0; s < numSN_; ++s) {
    const bool defer = tune_.cbMinRows >= 0 && hm_[s] > tune_.smallMax && hnE_[s] >= kCbMinCols &&
                       hm_[s] - hnE_[s] >= std::max(tune_.cbMinRows, 1);
    hcut_[s] = defer ? hnE_[s] : hm_[s];
  }
}

CbUpdate Plan::cbUpdate(const std::vector<CbTask>& tasks) {
  CbUpdate c;
  c.F = &hF_;
  c.fp32 = prec_ != 0;
  c.bf16x3 = prec_ != 0 && cbBf16x3_;
  if (!tasks.empty()) {
    size_t need = 0;
    for (const CbTask& t : tasks) need = std::max(need, static_cast<size_t>(t.m - t.nE) * kCbChunk);
    dcbW_.ensure(need);  // doubles: enough for either front type
  }
  c.w = dcbW_.p;
  c.wElems = dcbW_.n * (c.fp32 ? 2 : 1);
  return c;
}

void Plan::launchInverse() {
  launchDiagInverse(dinvList_.p, invCount_, dinvStart_.p, invBlocks_, devFronts(), dinv_.p, dinvOff_.p, stream_);
}

void Plan::buildSchedule() {
  std::vector<int> small, all;
  std::vector<int4> tasks;
  sched_.assign(levelFronts_.size(), {});
  levelFrontOff_.assign(levelFronts_.size() + 1, 0);
  for (size_t L = 0; L < levelFronts_.size(); ++L) {
    auto& ls = sched_[L];
    ls.smallOff = static_cast<int64_t>(small.size());
    ls.taskOff = static_cast<int64_t>(tasks.size());
    levelFrontOff_[L] = static_cast<int64_t>(all.size());
    for (int s : levelFronts_[L]) {
      all.push_back(s);
      // column tiles: a large front gets narrower tiles so that it alone
      // spreads over about 1024 workgroups (the top levels hold few fronts)
      int w = tune_.asmTile;
      while (w > 8 && (hm_[s] + w - 1) / w < 1024 && hm_[s] > 4 * tune_.asmTile) w >>= 1;
      for (int c0 = 0; c0 < hm_[s]; c0 += w) tasks.push_back(make_int4(s, c0, std::min(c0 + w, hm_[s]), 0));
      if (hm_[s] <= tune_.smallMax) {
        small.push_back(s);
        ls.smallMaxM = std::max(ls.smallMaxM, hm_[s]);
      } else {
        ls.large.push_back(s);
      }
    }
    ls.smallCount = static_cast<int>(small.size() - ls.smallOff);
    ls.taskCount = static_cast<int>(tasks.size() - ls.taskOff);
  }
  levelFrontOff_[levelFronts_.size()] = static_cast<int64_t>(all.size());
  LargeSchedHost lh;
  hwsOff_.assign(numSN_, 0);
  int64_t wsNeed = 2;
  computeCut();
  dschurCut_.upload(hcut_, stream_);
  for (auto& ls : sched_) {
    ls.lsched = buildLargeSched(ls.large, hm_, hnE_, hcut_, lh, tune_.specPanel ? tune_.specMinRows : -1);
    (void)cbUpdate(ls.lsched.cb);  // handle and workspace now: a recorded graph cannot allocate
    int64_t off = 0;
    for (int s : ls.large) {
      hwsOff_[s] = off;
      off += panelWorkspace(hm_[s], hnE_[s], tune_.specPanel && hnE_[s] > tune_.specMinRows);
    }
    wsNeed = std::max(wsNeed, off);
  }
  dws_.ensure(static_cast<size_t>(wsNeed));
  dwsOff_.upload(hwsOff_, stream_);
  dlargeLists_.upload(lh.lists, stream_);
  dlargeRowStart_.upload(lh.rowStart, stream_);
  dlargeTileStart_.upload(lh.tileStart, stream_);
  dlargeTileTp_.upload(lh.tileTp, stream_);
  dlargeMwStart_.upload(lh.mwStart, stream_);
  dlargeSpecStart_.upload(lh.specStart, stream_);
  dsmall_.upload(small, stream_);
  dtasks_.upload(tasks, stream_);
  dallFronts_.upload(all, stream_);
  hipCheck(hipStreamSynchronize(stream_), "buildSchedule");
}

// Symmetric CSR (both triangles from the stored lower entries, the
// selfadjointView<Lower> of refine.hpp:116) and the stored entries grouped by
// row (sparseMaxAbsRowSum, refine.hpp:69-81).
void Plan::buildRefineCsr() {
  std::vector<int64_t> symCnt(n_ + 1, 0), stCnt(n_ + 1, 0);
  for (int c = 0; c < n_; ++c)
    for (int64_t t = aColptr_[c]; t < aColptr_[c + 1]; ++t) {
      const int r = aRowidx_[t];
      ++stCnt[r + 1];
      if (r < c) continue;
      ++symCnt[r + 1];
      if (r != c) ++symCnt[c + 1];
    }
  for (int i = 0; i < n_; ++i) {
    symCnt[i + 1] += symCnt[i];
    stCnt[i + 1] += stCnt[i];
  }
  std::vector<int64_t> symSrc(symCnt[n_]), stSrc(stCnt[n_]);
  std::vector<int> symCol(symCnt[n_]);
  std::vector<int64_t> ps(symCnt.begin(), symCnt.end() - 1), pt(stCnt.begin(), stCnt.end() - 1);
  for (int c = 0; c < n_; ++c)
    for (int64_t t = aColptr_[c]; t < aColptr_[c + 1]; ++t) {
      const int r = aRowidx_[t];
      stSrc[pt[r]++] = t;
      if (r < c) continue;
      symCol[ps[r]] = c;
      symSrc[ps[r]++] = t;
      if (r != c) {
        symCol[ps[c]] = r;
        symSrc[ps[c]++] = t;
      }
    }
  nnzSym_ = symCnt[n_];
  dsymPtr_.upload(symCnt, stream_);
  dsymCol_.upload(symCol, stream_);
  dsymSrc_.upload(symSrc, stream_);
  dstPtr_.upload(stCnt, stream_);
  dstSrc_.upload(stSrc, stream_);
}

DevFronts Plan::devFronts() {
  DevFronts f;
  f.m = dm_.p;
  f.nE = dnE_.p;
  f.ncols = dncols_.p;
  f.F = dF_.p;
  f.fp32 = prec_;
  f.idxOff = didxOff_.p;
  f.frontIdx = dfrontIdx_.p;
  f.relmap = drelmap_.p;
  f.childPtr = dchildPtr_.p;
  f.childList = dchildList_.p;
  f.firstCol = dfirst_.p;
  f.acolPtr = dacolPtr_.p;
  f.aRow = daRow_.p;
  f.aSrc = daSrc_.p;
  f.schurCut = dschurCut_.p;
  f.perm = dperm_.p;
  f.rowIdx = drowIdx_.p;
  f.dDiag = ddDiag_.p;
  f.dOff = ddOff_.p;
  f.pkind = dpkind_.p;
  f.nF = dnF_.p;
  f.inertia = dinertia_.p;
  f.origMaxBits = dorigMax_.p;
  f.runMaxBits = drunMax_.p;
  f.kcur = dkcur_.p;
  f.pcount = dpcount_.p;
  f.stopped = dstopped_.p;
  f.cbDone = dcbDone_.p;
  f.pert = dpert_.p;
  f.npert = dnpert_.p;
  f.schurFlops = dschur_.p;
  return f;
}

FactorParams Plan::params(const FactorOptions& o, bool staticMode, double delta, const int* signDev) const {
  FactorParams p;
  p.zeroTol = o.zeroTol;
  p.relFloor = o.relFloor;
  p.staticMode = staticMode ? 1 : 0;
  p.staticDelta = delta;
  p.staticRelFloor = o.staticRelFloor;
  p.staticAbsFloor = o.staticAbsFloor;
  p.expectedSign = signDev;
  p.specPanel = tune_.specPanel ? 1 : 0;
  p.specMinRows = tune_.specMinRows;
  return p;
}

LargeSched buildLargeSched(const std::vector<int>& fronts, const std::vector<int>& m, const std::vector<int>& nE,
                           const std::vector<int>& cut, LargeSchedHost& h, int specMinRows, bool allowMw) {
  LargeSched ls;
  ls.allOff = static_cast<int64_t>(h.lists.size());
  ls.allCount = static_cast<int>(fronts.size());
  int maxIters = 0;
  for (int s : fronts) {
    h.lists.push_back(s);
    ls.maxM = std::max(ls.maxM, m[s]);
    maxIters = std::max(maxIters, (nE[s] + kPanel - 1) / kPanel);
    if (cut[s] < m[s]) ls.cb.push_back(CbTask{s, m[s], nE[s]});
  }
  for (int it = 0; it < maxIters; ++it) {
    LargeIter li;
    li.listOff = static_cast<int64_t>(h.lists.size());
    li.rowOff = static_cast<int64_t>(h.rowStart.size());
    li.tileOff = static_cast<int64_t>(h.tileStart.size());
    int rb = 0, tl = 0, sb = 0;
    li.specOff = static_cast<int64_t>(h.specStart.size());
    std::vector<int> reg, mw, mwW;
    int mwTotal = 0;
    for (int s : fronts) {
      if ((nE[s] + kPanel - 1) / kPanel <= it) continue;
      h.lists.push_back(s);
      h.rowStart.push_back(rb);
      h.tileStart.push_back(tl);
      h.specStart.push_back(sb);
      rb += (m[s] - nE[s] + 255) / 256;
      // rows below this iteration's diagonal block (the block starts at or after it * kPanel)
      const int below = m[s] - it * kPanel - std::min(kPanel, nE[s] - it * kPanel);
      if (specMinRows >= 0 && nE[s] > specMinRows) sb += std::max(1, (below + kSpecRows - 1) / kSpecRows);
      // trailing tiles: T per side from this iteration's panel start, column
      // blocks up to the cut
      const int T = (m[s] - it * kPanel + 63) / 64;
      const int Tc = std::min(T, (cut[s] - it * kPanel + 63) / 64);
      h.tileTp.push_back(T);
      tl += Tc * T - Tc * (Tc - 1) / 2;
      ++li.count;
      // eligible rows left at this iteration's first pivot (no delays: it * kPanel)
      const int rows = nE[s] - it * kPanel;
      const int W = (rows + kMwRows - 1) / kMwRows;
      if (allowMw && rows > kMwThreshold && W <= kMwMaxW && static_cast<int>(mw.size()) < kMwSlots &&
          mwTotal + W <= kMwMaxBlocks) {
        mw.push_back(s);
        mwW.push_back(W);
        mwTotal += W;
      } else {
        reg.push_back(s);
        li.regMaxNE = std::max(li.regMaxNE, nE[s]);
      }
    }
    h.rowStart.push_back(rb);
    h.tileStart.push_back(tl);
    h.tileTp.push_back(0);  // keeps tileTp at the offsets of tileStart
    h.specStart.push_back(sb);
    li.specBlocks = sb;
    li.rowBlocks = rb;
    li.tiles = tl;
    li.regOff = static_cast<int64_t>(h.lists.size());
    li.regCount = static_cast<int>(reg.size());
    h.lists.insert(h.lists.end(), reg.begin(), reg.end());
    li.mwOff = static_cast<int64_t>(h.lists.size());
    li.mwCount = static_cast<int>(mw.size());
    h.lists.insert(h.lists.end(), mw.begin(), mw.end());
    li.mwStartOff = static_cast<int64_t>(h.mwStart.size());
    int acc = 0;
    for (int W : mwW) {
      h.mwStart.push_back(acc);
      acc += W;
    }
    h.mwStart.push_back(acc);
    li.mwBlocks = acc;
    ls.iters.push_back(li);
  }
  return ls;
}

void runLargeSched(const LargeSched& ls, const LargeDev& dv, const DevFronts& fd, const FactorParams& prm, double* ws,
                   const int64_t* wsOff, hipStream_t st, const CbUpdate* cb, Plan* prof) {
  if (ls.allCount == 0) return;
  if (!dv.frontMaxDone) launchFrontMaxAbs(dv.lists + ls.allOff, ls.allCount, ls.maxM, fd, st);
  if (prm.specPanel) launchSpecReset(dv.lists + ls.allOff, ls.allCount, fd, ws, wsOff, st);
  for (size_t t = 0; t < ls.iters.size(); ++t) {
    const LargeIter& it = ls.iters[t];
    const int* lst = dv.lists + it.listOff;
    if (prof) prof->profBegin(Plan::kProfPanel);
    if (prm.specPanel && it.specBlocks > 0)
      launchSpecPanel(lst, it.count, dv.specStart + it.specOff, it.specBlocks, static_cast<int>(t & 1), fd, prm, ws,
                      wsOff, st);
    launchPanel(dv.lists + it.regOff, it.regCount, it.regMaxNE, fd, prm, ws, wsOff, st);
    launchPanelMw(dv.lists + it.mwOff, it.mwCount, dv.mwStart + it.mwStartOff, it.mwBlocks, fd, prm, ws, wsOff,
                  dv.sync, dv.err, st);
    if (prof) prof->profEnd(Plan::kProfPanel);
    if (it.rowBlocks > 0) {
      if (prof) prof->profBegin(Plan::kProfPanelRows);
      launchPanelRows(lst, it.count, dv.rowStart + it.rowOff, it.rowBlocks, fd, st);
      if (prof) prof->profEnd(Plan::kProfPanelRows);
    }
    if (prof) prof->profBegin(Plan::kProfSyrk);
    launchSyrk(lst, it.count, dv.tileStart + it.tileOff, dv.tileTp + it.tileOff, it.tiles, fd, st);
    if (prof) prof->profEnd(Plan::kProfSyrk);
  }
  launchFrontFinish(dv.lists + ls.allOff, ls.allCount, fd, st);
  if (ls.cb.empty()) return;
  if (!cb || !cb->w || !cb->F) throw std::logic_error("symla: deferred contribution blocks without a workspace");
  const size_t esz = cb->fp32 ? sizeof(float) : sizeof(double);
  for (const CbTask& t : ls.cb) {
    const int N = t.m - t.nE;
    if (static_cast<size_t>(N) * kCbChunk > cb->wElems) throw std::logic_error("symla: contribution-block workspace");
    char* F = static_cast<char*>((*cb->F)[t.s]);
    for (int q0 = 0; q0 < t.nE; q0 += kCbChunk) {
      const int kc = std::min(kCbChunk, t.nE - q0);
      // one profiled launch per chunk: the operand kernel and the MFMA SYRK
      if (prof) prof->profBegin(Plan::kProfCb);
      launchFormW(fd, t.s, N, q0, kc, cb->w, st);
      const void* B = F + esz * (static_cast<size_t>(t.nE) + static_cast<size_t>(q0) * t.m);
      void* C = F + esz * (static_cast<size_t>(t.nE) + static_cast<size_t>(t.nE) * t.m);
      if (cb->bf16x3)
        launchCbSyrkBf16x3(N, kc, static_cast<const float*>(cb->w), N, static_cast<const float*>(B), t.m,
                           static_cast<float*>(C), t.m, st);
      else
        launchCbSyrk(cb->fp32, N, kc, cb->w, N, B, t.m, C, t.m, st);
      if (prof) {
        prof->profEnd(Plan::kProfCb);
        prof->addCbFlops(static_cast<double>(kc) * N * (N + 1.0));
      }
    }
  }
}

LargeDev Plan::largeDev(const int* lists, const int* rowStart, const int* tileStart, const int* tileTp,
                        const int* mwStart, const int* specStart) {
  LargeDev d;
  d.frontMaxDone = true;  // k_assemble forms every front's max |F|
  d.tileTp = tileTp;
  d.specStart = specStart;
  d.lists = lists;
  d.rowStart = rowStart;
  d.tileStart = tileStart;
  d.mwStart = mwStart;
  d.sync = dmwSync_.p;
  d.err = dfactErr_.p;
  return d;
}

void Plan::runStatic(const double* aValsDev, const FactorParams& prm) {
  if (layoutDirty_) {
    buildStaticLayout();
    uploadLayout();
    layoutDirty_ = false;
  }
  DevFronts fd = devFronts();
  launchResetFronts(numSN_, fd, stream_);
  launchResetPerm(dallFronts_.p, numSN_, fd, stream_);
  for (size_t L = 0; L < sched_.size(); ++L) {
    const auto& ls = sched_[L];
    profBegin(kProfAssemble);
    launchAssemble(dtasks_.p + ls.taskOff, ls.taskCount, fd, aValsDev, 1, stream_);
    profEnd(kProfAssemble);
    if (ls.smallCount > 0) {
      profBegin(kProfSmall);
      launchFactorSmall(dsmall_.p + ls.smallOff, ls.smallCount, ls.smallMaxM, fd, prm, stream_);
      profEnd(kProfSmall);
    }
    const CbUpdate cb = cbUpdate(ls.lsched.cb);
    runLargeSched(ls.lsched,
                  largeDev(dlargeLists_.p, dlargeRowStart_.p, dlargeTileStart_.p, dlargeTileTp_.p, dlargeMwStart_.p,
                           dlargeSpecStart_.p),
                  fd, prm, dws_.p, dwsOff_.p, stream_, &cb, profiling_ ? this : nullptr);
  }
  launchMakeRowIdx(dallFronts_.p, numSN_, fd, stream_);
  launchReduceStats(numSN_, fd, dstatI_.p, dstatD_.p, stream_);
  launchInverse();
}

void Plan::collectStats(FactorInfo& info, bool& anyDelay) {
  if (profiling_) {
    unsigned long long fl = 0;
    hipCheck(hipMemcpyAsync(&fl, dschur_.p, sizeof(fl), hipMemcpyDeviceToHost, stream_), "schur flops");
    hipCheck(hipStreamSynchronize(stream_), "schur flops");
    profAcc_.schurFlops += static_cast<double>(fl);
  }
  long long si[8];
  double sd[4];
  int ferr = 0;
  hipCheck(hipMemcpyAsync(si, dstatI_.p, 5 * sizeof(long long), hipMemcpyDeviceToHost, stream_), "stats");
  hipCheck(hipMemcpyAsync(sd, dstatD_.p, sizeof(double), hipMemcpyDeviceToHost, stream_), "stats");
  hipCheck(hipMemcpyAsync(&ferr, dfactErr_.p, sizeof(int), hipMemcpyDeviceToHost, stream_), "stats");
  hipCheck(hipStreamSynchronize(stream_), "factorize");
  if (ferr != 0) {
    hipCheck(hipMemset(dfactErr_.p, 0, sizeof(int)), "error word");
    throw DeviceError("symla: a cooperative panel step exceeded its bounded wait (code " + std::to_string(ferr) +
                      ")");
  }
  info.pos = static_cast<int>(si[0]);
  info.neg = static_cast<int>(si[1]);
  info.zero = static_cast<int>(si[2]);
  info.delayedFronts = static_cast<int>(si[3]);
  info.numPerturbed = static_cast<int>(si[4]);
  info.totalPerturbation = sd[0];
  anyDelay = si[3] > 0;
}

// Host-driven level-by-level factorization that sizes every front with the
// delayed columns handed up by its children (multifrontal.hpp:453-479) and
// forwards leftover rows by their actual post-pivoting identity
// (multifrontal.hpp:612, :627-643).
void Plan::runDynamic(const double* aValsDev, const FactorParams& prm) {
  buildStaticLayout();
  layoutDirty_ = true;
  // final-order row of every A entry, for re-deriving local rows
  std::vector<int> inv(n_);
  for (int k = 0; k < n_; ++k) inv[perm_[k]] = k;
  std::vector<int> entryHi(haSrc_.size());
  for (size_t t = 0; t < haSrc_.size(); ++t) {
    const int64_t src = haSrc_[t];
    // original column of src: search colptr
    const int c = static_cast<int>(std::upper_bound(aColptr_.begin(), aColptr_.end(), src) - aColptr_.begin()) - 1;
    const int r = aRowidx_[src];
    entryHi[t] = std::max(inv[r], inv[c]);
  }
  uploadLayout();
  hnF_.assign(numSN_, 0);
  std::vector<std::vector<int>> rowIdxHost(numSN_);
  DevFronts fd = devFronts();
  launchResetFronts(numSN_, fd, stream_);
  DBuf<int> dlist;
  DBuf<int4> dtask;
  for (size_t L = 0; L < levelFronts_.size(); ++L) {
    const auto& fronts = levelFronts_[L];
    for (int s : fronts) {
      const auto& pat = sn_[s].rowPattern;
      std::vector<int> extras;
      for (int c : children_[s])
        for (int i = hnF_[c]; i < hm_[c]; ++i) {
          const int g = rowIdxHost[c][i];
          if (!std::binary_search(pat.begin(), pat.end(), g)) extras.push_back(g);
        }
      std::sort(extras.begin(), extras.end());
      extras.erase(std::unique(extras.begin(), extras.end()), extras.end());
      const int nExtra = static_cast<int>(extras.size());
      const int ncols = sn_[s].ncols;
      if (nExtra > 0) {
        const int m = static_cast<int>(pat.size()) + nExtra;
        std::vector<int> fi(m);
        for (int j = 0; j < ncols; ++j) fi[j] = sn_[s].firstCol + j;
        for (int j = 0; j < nExtra; ++j) fi[ncols + j] = extras[j];
        for (int j = ncols; j < static_cast<int>(pat.size()); ++j) fi[ncols + nExtra + (j - ncols)] = pat[j];
        const int64_t off = idxUsed_;
        ensureIdxCapacity(off + m);
        idxUsed_ += m;
        hfrontIdx_.resize(idxUsed_, 0);
        hrelmap_.resize(idxUsed_, -1);
        std::copy(fi.begin(), fi.end(), hfrontIdx_.begin() + off);
        hidxOff_[s] = off;
        hm_[s] = m;
        hnE_[s] = ncols + nExtra;
        void* np = nullptr;
        hipCheck(hipMalloc(&np, elemSize() * static_cast<size_t>(m) * m), "hipMalloc front");
        ++g_devAllocEpoch;
        dynAllocs_.push_back(np);
        hF_[s] = np;
        // own-column A entries: local rows shift past the extras
        for (int j = 0; j < ncols; ++j) {
          const int gcol = sn_[s].firstCol + j;
          for (int64_t t = hacolPtr_[gcol]; t < hacolPtr_[gcol + 1]; ++t) {
            const int hi = entryHi[t];
            auto it = std::lower_bound(pat.begin(), pat.end(), hi);
            const int pidx = static_cast<int>(it - pat.begin());
            haRow_[t] = pidx < ncols ? pidx : pidx + nExtra;
          }
        }
      }
      // children's leftover rows -> this front's local positions
      const int m = hm_[s];
      const int64_t io = hidxOff_[s];
      for (int c : children_[s]) {
        const int64_t ioc = hidxOff_[c];
        for (int i = 0; i < hm_[c]; ++i) {
          int v = -1;
          if (i >= hnF_[c]) {
            const int g = rowIdxHost[c][i];
            // local position of g in s's front index list
            // front index list = [own columns | extras (sorted) | update rows (sorted)]
            const int* beg = hfrontIdx_.data() + io;
            const int* extrasEnd = beg + hnE_[s];
            if (g >= sn_[s].firstCol && g < sn_[s].firstCol + ncols) {
              v = g - sn_[s].firstCol;
            } else {
              const int* p1 = std::lower_bound(beg + ncols, extrasEnd, g);
              if (p1 != extrasEnd && *p1 == g) {
                v = static_cast<int>(p1 - beg);
              } else {
                const int* p2 = std::lower_bound(extrasEnd, beg + m, g);
                if (p2 != beg + m && *p2 == g) v = static_cast<int>(p2 - beg);
              }
            }
            if (v < 0) throw std::runtime_error("symla: forwarded row not found in parent front");
          }
          hrelmap_[ioc + i] = v;
        }
      }
    }
    uploadLayout();
    fd = devFronts();
    std::vector<int> small, all;
    std::vector<int4> tasks;
    int maxM = 0;
    std::vector<int> large;
    for (int s : fronts) {
      all.push_back(s);
      tasks.push_back(make_int4(s, 0, hm_[s], 0));
      if (hm_[s] <= tune_.smallMax) {
        small.push_back(s);
        maxM = std::max(maxM, hm_[s]);
      } else {
        large.push_back(s);
      }
    }
    std::vector<int> lists(all);
    lists.insert(lists.end(), small.begin(), small.end());
    dlist.upload(lists, stream_);
    dtask.upload(tasks, stream_);
    launchResetPerm(dlist.p, static_cast<int>(all.size()), fd, stream_);
    launchAssemble(dtask.p, static_cast<int>(tasks.size()), fd, aValsDev, 0, stream_);
    launchFactorSmall(dlist.p + all.size(), static_cast<int>(small.size()), maxM, fd, prm, stream_);
    if (!large.empty()) {
      LargeSchedHost lh;
      LargeSched lsd = buildLargeSched(large, hm_, hnE_, hcut_, lh, tune_.specPanel ? tune_.specMinRows : -1);
      std::vector<int64_t> wsOffH(numSN_, 0);
      int64_t off = 0;
      for (int s : large) {
        wsOffH[s] = off;
        off += panelWorkspace(hm_[s], hnE_[s], tune_.specPanel && hnE_[s] > tune_.specMinRows);
      }
      dws_.ensure(static_cast<size_t>(std::max<int64_t>(off, 2)));
      DBuf<int> dl, dr, dt, dtp, dmw, dsp;
      DBuf<int64_t> dw;
      dl.upload(lh.lists, stream_);
      dr.upload(lh.rowStart, stream_);
      dt.upload(lh.tileStart, stream_);
      dtp.upload(lh.tileTp, stream_);
      dmw.upload(lh.mwStart, stream_);
      dsp.upload(lh.specStart, stream_);
      dw.upload(wsOffH, stream_);
      const CbUpdate cb = cbUpdate(lsd.cb);
      runLargeSched(lsd, largeDev(dl.p, dr.p, dt.p, dtp.p, dmw.p, dsp.p), fd, prm, dws_.p, dw.p, stream_, &cb,
                    nullptr);
      hipCheck(hipStreamSynchronize(stream_), "dynamic large fronts");
    }
    launchMakeRowIdx(dlist.p, static_cast<int>(all.size()), fd, stream_);
    // read back pivot counts and row identities of this level
    std::vector<int> nfAll(numSN_);
    hipCheck(hipMemcpyAsync(nfAll.data(), dnF_.p, sizeof(int) * numSN_, hipMemcpyDeviceToHost, stream_), "nF");
    for (int s : fronts) {
      rowIdxHost[s].resize(hm_[s]);
      hipCheck(hipMemcpyAsync(rowIdxHost[s].data(), drowIdx_.p + hidxOff_[s], sizeof(int) * hm_[s],
                              hipMemcpyDeviceToHost, stream_),
               "rowIdx");
    }
    hipCheck(hipStreamSynchronize(stream_), "dynamic level");
    for (int s : fronts) hnF_[s] = nfAll[s];
  }
  // the relmap of every child was rewritten for its parent; upload the final layout
  uploadLayout();
  fd = devFronts();
  launchReduceStats(numSN_, fd, dstatI_.p, dstatD_.p, stream_);
  launchInverse();
  singular_ = false;
  singularColsOrig_.clear();
  std::vector<int> sing;
  for (int r : roots_)
    for (int i = hnF_[r]; i < hm_[r]; ++i) sing.push_back(rowIdxHost[r][i]);
  std::sort(sing.begin(), sing.end());
  for (int g : sing) singularColsOrig_.push_back(perm_[g]);
  singular_ = !sing.empty();
}

void Plan::factorize(const double* aHost, const double* aDev, const FactorOptions& opts, FactorInfo& info) {
  factorized_ = false;
  const double* vals = aDev;
  if (aHost) {
    hipCheck(hipMemcpyAsync(dvals_.p, aHost, sizeof(double) * nnzA_, hipMemcpyHostToDevice, stream_), "values");
    vals = dvals_.p;
  }
  info = FactorInfo{};
  if (opts.precision != 0 && opts.precision != 1) throw std::invalid_argument("symla: unknown precision");
  cbBf16x3_ = opts.bf16x3 && opts.precision == 1;
  if (opts.precision != prec_) {
    // the front storage changes element size: re-plan the pool
    prec_ = opts.precision;
    layoutDirty_ = true;
  }
  hipCheck(hipEventRecord(ev0_, stream_), "event");
  if (opts.mode == 0) {
    FactorParams prm = params(opts, false, 0.0, nullptr);
    singular_ = false;
    singularColsOrig_.clear();
    runStatic(vals, prm);
    bool anyDelay = false;
    collectStats(info, anyDelay);
    if (anyDelay) {
      runDynamic(vals, prm);
      collectStats(info, anyDelay);
      info.repaired = true;
    }
    info.singular = singular_;
    info.nSingular = static_cast<int>(singularColsOrig_.size());
    info.precisionUsed = precisionCode();
    hipCheck(hipEventRecord(ev1_, stream_), "event");
    hipCheck(hipEventSynchronize(ev1_), "event");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0_, ev1_);
    factorMs_ = ms;
    ensureSolveLayout();  // a no-op unless the layout changed
    factorized_ = true;
    return;
  }
  // Mode::StaticRegularized (solver.hpp:328-417)
  const bool haveExp = opts.haveExpectation;
  std::vector<int> signsOrig(n_, 0);
  if (haveExp) {
    if (static_cast<int>(opts.expectedSignOrig.size()) != n_)
      throw std::invalid_argument("symla: expected sign pattern size does not match the matrix size");
    signsOrig = opts.expectedSignOrig;
  }
  std::vector<int> signsFinal(n_);
  for (int t = 0; t < n_; ++t) signsFinal[t] = signsOrig[perm_[t]];
  dsignFinal_.upload(signsFinal, stream_);
  int expectedPos = 0, expectedNeg = 0;
  for (int v : signsOrig) {
    if (v > 0)
      ++expectedPos;
    else if (v < 0)
      ++expectedNeg;
  }
  // max|A| for the default perturbation sqrt(eps) * max|A| (solver.hpp:346-358)
  double maxAbs = 1.0;
  if (opts.initialDelta <= 0.0) {
    launchZeroU64(dnorms_.p, 1, stream_);
    if (nnzA_ > 0) launchAbsMax(static_cast<int>(nnzA_), 1, vals, nnzA_, dnorms_.p, stream_);
    unsigned long long b = 0;
    hipCheck(hipMemcpyAsync(&b, dnorms_.p, sizeof(b), hipMemcpyDeviceToHost, stream_), "maxabs");
    hipCheck(hipStreamSynchronize(stream_), "maxabs");
    maxAbs = bitsToDoubleHost(b);
    if (maxAbs <= 0.0) maxAbs = 1.0;
  }
  const bool fullCoverage = (expectedPos + expectedNeg == n_);
  // one inertia-controlled retry sequence in the current front precision;
  // returns true when the factorization holds the expected inertia
  auto retryLoop = [&](int maxAttempts, double eps) {
    double delta = opts.initialDelta > 0.0 ? opts.initialDelta : std::sqrt(eps) * maxAbs;
    for (int attempt = 0; attempt < maxAttempts; ++attempt) {
      FactorParams prm = params(opts, true, delta, dsignFinal_.p);
      singular_ = false;
      singularColsOrig_.clear();
      runStatic(vals, prm);
      bool anyDelay = false;
      collectStats(info, anyDelay);
      info.lastDelta = delta;
      info.retriesUsed = attempt;
      info.precisionUsed = precisionCode();
      const bool ok = !haveExp || (info.zero == 0 && !singular_ &&
                                   (!fullCoverage || (info.pos == expectedPos && info.neg == expectedNeg)));
      if (ok) return true;
      if (attempt + 1 < maxAttempts) delta *= opts.growth;
    }
    return false;
  };
  auto finish = [&]() {
    hipCheck(hipEventRecord(ev1_, stream_), "event");
    hipCheck(hipEventSynchronize(ev1_), "event");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0_, ev1_);
    factorMs_ = ms;
    ensureSolveLayout();
    factorized_ = true;
  };
  const int maxAttempts = std::max(1, opts.maxRetries + 1);
  if (prec_ == 1) {
    // FP32 fronts with a perturbation sized for FP32 (eps of float); if they
    // do not hold the expected inertia, the element growth of the perturbed
    // pivots is beyond FP32 and the reference's own FP64 sequence runs
    // instead (reported as precision_used = FP64).  A second FP32 attempt
    // with a grown perturbation did not hold it either on the KKT systems
    // measured, and cost a whole factorization.
    if (retryLoop(1, static_cast<double>(std::numeric_limits<float>::epsilon()))) {
      finish();
      return;
    }
    prec_ = 0;
    layoutDirty_ = true;
  }
  if (retryLoop(maxAttempts, std::numeric_limits<double>::epsilon())) {
    finish();
    return;
  }
  std::ostringstream oss;
  oss << "symla::SymLDLT::factorize (Mode::StaticRegularized): inertia-controlled regularization retry budget ("
      << opts.maxRetries << " retries) exhausted without matching the expected inertia. expected (n_pos="
      << expectedPos << ", n_neg=" << expectedNeg << ", n_zero=0), got (n_pos=" << info.pos << ", n_neg=" << info.neg
      << ", n_zero=" << info.zero << ", singular=" << (singular_ ? "true" : "false") << ") at delta=" << info.lastDelta
      << ". Refusing to report a wrong-inertia factorization as successful.";
  throw std::runtime_error(oss.str());
}

void Plan::ensureSolveWs(int nrhs) {
  ensureSolveLayout();
  const size_t nn = static_cast<size_t>(n_) * nrhs;
  dBp_.ensure(nn);
  dZ_.ensure(nn);
  dX_.ensure(nn);
  dDl_.ensure(static_cast<size_t>(workRows_) * nrhs);
}

void Plan::ensureSolveLayout() {
  if (!solveLayoutReady_) {
    // per-front work blocks of m x nrhs start at off[s] * nrhs
    std::vector<int64_t> off(numSN_ + 1, 0);
    for (int s = 0; s < numSN_; ++s) off[s + 1] = off[s] + hm_[s];
    sumM_ = off[numSN_];
    dgOff_.upload(off, stream_);
    // Gather map from the current layout and pivot counts (the static
    // layout takes every eligible pivot; a repaired one has its own counts).
    {
      std::vector<int64_t> cnt(static_cast<size_t>(idxUsed_) + 1, 0);
      auto forEach = [&](auto&& fn) {
        for (int s = 0; s < numSN_; ++s)
          for (int ci = hchildPtr_[s]; ci < hchildPtr_[s + 1]; ++ci) {
            const int c = hchildList_[ci];
            const int nfc = layoutDirty_ ? hnF_[c] : hnE_[c];
            for (int a = nfc; a < hm_[c]; ++a) {
              const int t = hrelmap_[hidxOff_[c] + a];
              if (t >= 0) fn(hidxOff_[s] + t, c, a);
            }
          }
      };
      forEach([&](int64_t pos, int, int) { ++cnt[pos + 1]; });
      for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
      std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
      std::vector<int64_t> src(static_cast<size_t>(std::max<int64_t>(cnt.back(), 1)));
      forEach([&](int64_t pos, int c, int a) { src[fill[pos]++] = off[c] + a; });
      dmapPtr_.upload(cnt, stream_);
      dmapRow_.upload(src, stream_);
    }
    // Per level launch lists (see SolveLevel); the pivots each front took
    // are known here (the static layout takes every eligible pivot).
    const int nl = static_cast<int>(levelFronts_.size());
    solveLv_.assign(nl, {});
    std::vector<int> lists, starts;
    std::vector<int2> pos;
    std::vector<AncPart> parts(numSN_, AncPart{0, 0, 1});
    int64_t partRows = 0;
    auto addLaunch = [&](SolveLaunch& ln, const std::vector<int>& fronts, auto&& blocksOf) {
      ln.list = static_cast<int64_t>(lists.size());
      ln.start = static_cast<int64_t>(starts.size());
      ln.count = static_cast<int>(fronts.size());
      int acc = 0;
      for (int s : fronts) {
        lists.push_back(s);
        starts.push_back(acc);
        acc += blocksOf(s);
      }
      starts.push_back(acc);
      ln.blocks = acc;
    };
    for (int L = 0; L < nl; ++L) {
      SolveLevel& lv = solveLv_[L];
      const auto& fr = levelFronts_[L];
      lv.posOff = static_cast<int64_t>(pos.size());
      for (int s : fr)
        for (int p = 0; p < hm_[s]; ++p) pos.push_back(make_int2(s, p));
      lv.positions = static_cast<int64_t>(pos.size()) - lv.posOff;
      std::vector<int> own, chain, delta, anc, tiny;
      for (int s : fr) {
        const int nfs = layoutDirty_ ? hnF_[s] : hnE_[s];
        if (hm_[s] <= kTinyFront) {
          tiny.push_back(s);
          continue;
        }
        anc.push_back(s);
        if ((nfs + kSolveBlock - 1) / kSolveBlock <= kOwnChainBlocks)
          own.push_back(s);
        else
          chain.push_back(s);
        if (hm_[s] > nfs) delta.push_back(s);
      }
      auto nfOf = [&](int s) { return layoutDirty_ ? hnF_[s] : hnE_[s]; };
      auto ownBlocks = [&](int s) { return (nfOf(s) + kSolveBlock - 1) / kSolveBlock; };
      // The ancestors' product of a level: split each front's ancestor rows
      // in chunks of at least kAncChunkRows until the level has about
      // kAncLevelBlocks workgroups (the top levels have few fronts with
      // many rows); chunk k >= 1 leaves its part in rows past the fronts.
      int levelOwn = 0;
      for (int s : anc) levelOwn += ownBlocks(s);
      for (int s : anc) {
        const int rowsAnc = hm_[s] - nfOf(s);
        int split = 1;
        if (rowsAnc > 0 && levelOwn > 0)
          split = std::max(1, std::min({(rowsAnc + kAncChunkRows - 1) / kAncChunkRows,
                                        (kAncLevelBlocks + levelOwn - 1) / levelOwn, kAncMaxSplit}));
        AncPart& pa = parts[s];
        pa.split = split;
        pa.stride = ownBlocks(s) * kSolveBlock;
        pa.off = sumM_ + partRows;
        partRows += static_cast<int64_t>(split - 1) * pa.stride;
      }
      addLaunch(lv.own, own, [](int) { return 1; });
      addLaunch(lv.chain, chain, ownBlocks);
      addLaunch(lv.delta, delta, [&](int s) { return (hm_[s] - nfOf(s) + 63) / 64; });
      addLaunch(lv.anc, anc, [&](int s) { return ownBlocks(s) * parts[s].split; });
      addLaunch(lv.tiny, tiny, [](int) { return 1; });
    }
    workRows_ = sumM_ + partRows;
    dancPart_.upload(parts, stream_);
    dsolveLists_.upload(lists, stream_);
    dsolveStarts_.upload(starts, stream_);
    dpos_.upload(pos, stream_);
    dposInfo_.ensure(std::max<size_t>(pos.size(), 1));
    posTotal_ = static_cast<int64_t>(pos.size());
    std::vector<int64_t> fo(numSN_ + 1, 0);
    for (int s = 0; s < numSN_; ++s) fo[s + 1] = fo[s] + (hm_[s] + 31) / 32;
    dflagOff_.upload(fo, stream_);
    const size_t nflags = static_cast<size_t>(std::max<int64_t>(fo[numSN_], 1));
    dflagsF_.alloc(nflags);
    dflagsB_.alloc(nflags);
    launchZeroU32(dflagsF_.p, static_cast<int64_t>(nflags), stream_);
    launchZeroU32(dflagsB_.p, static_cast<int64_t>(nflags), stream_);
    if (dgen_.n == 0) {
      dgen_.alloc(1);
      dsolveErr_.alloc(1);
    }
    launchZeroU32(dgen_.p, 1, stream_);
    launchZeroU32(reinterpret_cast<unsigned*>(dsolveErr_.p), 1, stream_);
    // the uploads above read host vectors local to this block
    hipCheck(hipStreamSynchronize(stream_), "solve workspace");
    solveLayoutReady_ = true;
  }
}

void Plan::checkSolveError() {
  int err = 0;
  hipCheck(hipMemcpyAsync(&err, dsolveErr_.p, sizeof(int), hipMemcpyDeviceToHost, stream_), "solve status");
  hipCheck(hipStreamSynchronize(stream_), "solve status");
  if (err != 0) {
    launchZeroU32(reinterpret_cast<unsigned*>(dsolveErr_.p), 1, stream_);
    hipCheck(hipStreamSynchronize(stream_), "solve status");
    throw DeviceError("symla: a solve sweep exceeded its bounded dependency wait");
  }
}

void Plan::solveCore(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx, const SolveWs* ext) {
  SolveWs w;
  if (ext) {
    if (!solveLayoutReady_) throw std::logic_error("symla: the solve layout is not built (factorize first)");
    w = *ext;
  } else {
    ensureSolveWs(nrhs);
    w = SolveWs{dBp_.p, dZ_.p, dX_.p, dDl_.p};
  }
  DevFronts fd = devFronts();
  launchSolveBegin(dgen_.p, stream_);
  launchPermuteIn(n_, nrhs, dpermVec_.p, dB, ldb, w.Bp, stream_);
  launchPosInfo(dpos_.p, posTotal_, fd, dgOff_.p, dmapPtr_.p, dposInfo_.p, stream_);
  const int* lists = dsolveLists_.p;
  const int* starts = dsolveStarts_.p;
  int* err = dsolveErr_.p;
  const int nl = static_cast<int>(solveLv_.size());
  for (int L = 0; L < nl; ++L) {
    const SolveLevel& lv = solveLv_[L];
    profBegin(kProfFwd);
    launchFwdGather(dposInfo_.p + lv.posOff, lv.positions, nrhs, w.Bp, w.Dl, dmapRow_.p, stream_);
    launchFwdOwn(lists + lv.own.list, lv.own.count, fd, nrhs, w.Z, w.Dl, dgOff_.p, dinv_.p, dinvOff_.p, stream_);
    launchFwdChain(lists + lv.chain.list, lv.chain.count, starts + lv.chain.start, lv.chain.blocks, fd, nrhs, w.Z,
                   w.Dl, dgOff_.p, dflagsF_.p, dflagOff_.p, dgen_.p, err, dinv_.p, dinvOff_.p, stream_);
    launchFwdDelta(lists + lv.delta.list, lv.delta.count, starts + lv.delta.start, lv.delta.blocks, fd, nrhs, w.Dl,
                   dgOff_.p, err, stream_);
    launchFwdTiny(lists + lv.tiny.list, lv.tiny.count, fd, nrhs, w.Z, w.Dl, dgOff_.p, stream_);
    profEnd(kProfFwd);
  }
  for (int L = nl - 1; L >= 0; --L) {
    const SolveLevel& lv = solveLv_[L];
    profBegin(kProfBwd);
    launchBwdAnc(lists + lv.anc.list, lv.anc.count, starts + lv.anc.start, lv.anc.blocks, fd, nrhs, w.Z, w.X, w.Dl,
                 dgOff_.p, dancPart_.p, err, stream_);
    launchBwdOwn(lists + lv.own.list, lv.own.count, fd, nrhs, w.X, w.Dl, dgOff_.p, dancPart_.p, dinv_.p, dinvOff_.p,
                 stream_);
    launchBwdChain(lists + lv.chain.list, lv.chain.count, starts + lv.chain.start, lv.chain.blocks, fd, nrhs, w.X,
                   w.Dl, dgOff_.p, dancPart_.p, dflagsB_.p, dflagOff_.p, dgen_.p, err, dinv_.p, dinvOff_.p,
                   stream_);
    launchBwdTiny(lists + lv.tiny.list, lv.tiny.count, fd, nrhs, w.Z, w.X, stream_);
    profEnd(kProfBwd);
  }
  launchPermuteOut(n_, nrhs, dpermVec_.p, w.X, dX, ldx, stream_);
}

void Plan::solve(const double* B, int64_t ldb, int nrhs, double* X, int64_t ldx, bool deviceIo) {
  if (!factorized_) throw std::logic_error("symla::SymLDLT::solve: factorize() must be called before solve()");
  if (singular_)
    throw std::logic_error(
        "symla::SymLDLT::solve: factorization is numerically singular (see isSingular()/singularColumns()); refusing "
        "to solve");
  if (nrhs < 0) throw std::invalid_argument("symla: nrhs must be non-negative");
  if (nrhs == 0 || n_ == 0) return;
  hipCheck(hipEventRecord(ev0_, stream_), "event");
  if (deviceIo) {
    solveCore(B, ldb, nrhs, X, ldx);
  } else {
    const size_t nn = static_cast<size_t>(n_) * nrhs;
    dBin_.ensure(nn);
    dXout_.ensure(nn);
    hipCheck(hipMemcpy2DAsync(dBin_.p, sizeof(double) * n_, B, sizeof(double) * ldb, sizeof(double) * n_, nrhs,
                              hipMemcpyHostToDevice, stream_),
             "B upload");
    solveCore(dBin_.p, n_, nrhs, dXout_.p, n_);
    hipCheck(hipMemcpy2DAsync(X, sizeof(double) * ldx, dXout_.p, sizeof(double) * n_, sizeof(double) * n_, nrhs,
                              hipMemcpyDeviceToHost, stream_),
             "X download");
  }
  hipCheck(hipEventRecord(ev1_, stream_), "event");
  hipCheck(hipEventSynchronize(ev1_), "solve");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  solveMs_ = ms;
  checkSolveError();
}

void Plan::solveNoAlloc(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx, double* scratch,
                        size_t capacity) {
  if (!factorized_) throw std::logic_error("symla::SymLDLT::solve: factorize() must be called before solve()");
  if (singular_)
    throw std::logic_error(
        "symla::SymLDLT::solve: factorization is numerically singular (see isSingular()/singularColumns()); refusing "
        "to solve");
  if (nrhs < 0) throw std::invalid_argument("symla: nrhs must be non-negative");
  if (nrhs == 0 || n_ == 0) return;
  if (!solveLayoutReady_) throw std::logic_error("symla: the solve layout is not built (factorize first)");
  const size_t nn = static_cast<size_t>(n_) * nrhs;
  const size_t gm = static_cast<size_t>(workRows_) * nrhs;
  const size_t need = 3 * nn + gm;
  if (need > capacity)
    throw std::invalid_argument("symla_solve_device_noalloc: this solve needs " + std::to_string(need * 8 >> 20) +
                                " MiB of workspace ((3 n + solve work rows) x nrhs doubles), above the " +
                                std::to_string(capacity * 8 >> 20) +
                                " MiB static scratch; use fewer right-hand sides per call or symla_solve_device");
  const SolveWs w{scratch, scratch + nn, scratch + 2 * nn, scratch + 3 * nn};
  // the event profile would create events on first use: off for this call
  const bool prof = profiling_;
  profiling_ = false;
  hipCheck(hipEventRecord(ev0_, stream_), "event");
  solveCore(dB, ldb, nrhs, dX, ldx, &w);
  hipCheck(hipEventRecord(ev1_, stream_), "event");
  profiling_ = prof;
  hipCheck(hipEventSynchronize(ev1_), "solve");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  solveMs_ = ms;
  checkSolveError();
}

// Refinement in pieces, shared by the eager loop and the recorded graphs:
// refineSetup (norms of A and B, symmetric values, first solve, residual)
// and refineStep (one correction); the host reads the norm slots between
// them.  Norm slots of dnorms_: 0 the reference's ||A|| (refine.hpp:106), 1
// the symmetric ||A||inf, 2 ||B||, 3 ||R||, 4 ||X||, 8.. per-column norms
// of the backward error.
void Plan::ensureRefineWs(int nrhs) {
  const size_t nn = static_cast<size_t>(n_) * nrhs;
  dBr_.ensure(nn);
  dR_.ensure(nn);
  dDx_.ensure(nn);
  dXn_.ensure(nn);
  dRn_.ensure(nn);
  dXout_.ensure(nn);
  if (nrhs >= 16) dXt_.ensure(nn);  // row-major copy of X for the many-right-hand-side residual
  dnorms_.ensure(8 + 3 * static_cast<size_t>(nrhs));
  dsymVal_.ensure(static_cast<size_t>(nnzSym_));
}

void Plan::refineSetup(const double* vals, int nrhs) {
  launchZeroU64(dnorms_.p, 8, stream_);
  launchRowSumMax(n_, dstPtr_.p, dstSrc_.p, vals, dnorms_.p + 0, stream_);
  launchRowSumMax(n_, dsymPtr_.p, dsymSrc_.p, vals, dnorms_.p + 1, stream_);
  launchAbsMax(n_, nrhs, dBr_.p, n_, dnorms_.p + 2, stream_);
  launchGatherVals(nnzSym_, dsymSrc_.p, vals, dsymVal_.p, stream_);
  solveCore(dBr_.p, n_, nrhs, dXout_.p, n_);
  // max |R| and max |X| accumulate into slots 3 and 4
  launchResidual(n_, nrhs, dsymPtr_.p, dsymCol_.p, dsymVal_.p, dBr_.p, dXout_.p, dR_.p, dnorms_.p + 3, stream_,
                 dXt_.p);
  launchAbsMax(n_, nrhs, dXout_.p, n_, dnorms_.p + 4, stream_);
}

void Plan::refineStep(const double* Xc, const double* Rc, double* Xn, double* Rn, int nrhs) {
  solveCore(Rc, n_, nrhs, dDx_.p, n_);
  launchZeroU64(dnorms_.p + 3, 2, stream_);
  launchAxpy(static_cast<int64_t>(n_) * nrhs, Xc, dDx_.p, Xn, dnorms_.p + 4, stream_);
  launchResidual(n_, nrhs, dsymPtr_.p, dsymCol_.p, dsymVal_.p, dBr_.p, Xn, Rn, dnorms_.p + 3, stream_, dXt_.p);
}

Plan::RefineNorms Plan::readNorms() {
  unsigned long long h[5];
  hipCheck(hipMemcpyAsync(h, dnorms_.p, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_), "norms");
  hipCheck(hipStreamSynchronize(stream_), "norms");
  RefineNorms r;
  r.normA = bitsToDoubleHost(h[0]);
  r.normAsym = bitsToDoubleHost(h[1]);
  r.normB = bitsToDoubleHost(h[2]);
  const double nR = bitsToDoubleHost(h[3]), nX = bitsToDoubleHost(h[4]);
  // the device maps NaN entries to +inf: a non-finite residual or solution
  // is an infinite relative residual, never a converged one
  if (!std::isfinite(nR) || !std::isfinite(nX)) {
    r.rel = std::numeric_limits<double>::infinity();
  } else {
    const double denom = r.normA * nX + r.normB;
    r.rel = denom > 0.0 ? nR / denom : nR;
  }
  return r;
}

double Plan::backwardError(const double* R, const double* X, int nrhs, double normAsym) {
  // per column with the symmetric infinity norm
  launchZeroU64(dnorms_.p + 8, 3 * nrhs, stream_);
  launchAbsMaxCols(n_, nrhs, R, n_, dnorms_.p + 8, 3, stream_);
  launchAbsMaxCols(n_, nrhs, X, n_, dnorms_.p + 9, 3, stream_);
  launchAbsMaxCols(n_, nrhs, dBr_.p, n_, dnorms_.p + 10, 3, stream_);
  std::vector<unsigned long long> h(3 * static_cast<size_t>(nrhs));
  hipCheck(hipMemcpyAsync(h.data(), dnorms_.p + 8, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          stream_),
           "be");
  hipCheck(hipStreamSynchronize(stream_), "be");
  double be = 0.0;
  for (int c = 0; c < nrhs; ++c) {
    const double nr = bitsToDoubleHost(h[3 * c]), nx = bitsToDoubleHost(h[3 * c + 1]),
                 nb = bitsToDoubleHost(h[3 * c + 2]);
    const double den = normAsym * nx + nb;
    const double bc = (!std::isfinite(nr) || !std::isfinite(nx)) ? std::numeric_limits<double>::infinity()
                      : den > 0.0                                 ? nr / den
                                                                  : nr;
    be = std::max(be, bc);
  }
  return be;
}

RefineResult Plan::solveRefined(const double* aHost, const double* aDev, const double* B, int64_t ldb, int nrhs,
                                double* X, int64_t ldx, bool deviceIo, int maxIt, double tol) {
  if (!factorized_) throw std::logic_error("symla::SymLDLT::solve: factorize() must be called before solve()");
  if (singular_)
    throw std::logic_error(
        "symla::SymLDLT::solve: factorization is numerically singular (see isSingular()/singularColumns()); refusing "
        "to solve");
  RefineResult res;
  if (nrhs <= 0 || n_ == 0) return res;
  const double* vals = aDev;
  if (aHost) {
    hipCheck(hipMemcpyAsync(dvals_.p, aHost, sizeof(double) * nnzA_, hipMemcpyHostToDevice, stream_), "values");
    vals = dvals_.p;
  }
  ensureRefineWs(nrhs);
  hipCheck(hipMemcpy2DAsync(dBr_.p, sizeof(double) * n_, B, sizeof(double) * ldb, sizeof(double) * n_, nrhs,
                            deviceIo ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_),
           "B copy");
  refineSetup(vals, nrhs);
  const RefineNorms nm = readNorms();
  const double* Xc = refineLoop(nm, nrhs, maxIt, tol, nullptr, res);
  hipCheck(hipMemcpy2DAsync(X, sizeof(double) * ldx, Xc, sizeof(double) * n_, sizeof(double) * n_, nrhs,
                            deviceIo ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream_),
           "X out");
  hipCheck(hipStreamSynchronize(stream_), "refine");
  checkSolveError();
  return res;
}

const double* Plan::refineLoop(const RefineNorms& nm, int nrhs, int maxIt, double tol, hipGraphExec_t* steps,
                               RefineResult& res) {
  double* Xc = dXout_.p;
  double* Rc = dR_.p;
  double* Xn = dXn_.p;
  double* Rn = dRn_.p;
  double r0 = nm.rel;
  res.initialRel = r0;
  res.finalRel = r0;
  int parity = 0;
  for (int it = 0; it < maxIt; ++it) {  // refine.hpp:121-141
    if (r0 <= tol) break;
    if (steps)
      hipCheck(hipGraphLaunch(steps[parity], stream_), "graph launch");
    else
      refineStep(Xc, Rc, Xn, Rn, nrhs);
    const double rNew = readNorms().rel;
    if (!(rNew < r0)) break;
    std::swap(Xc, Xn);
    std::swap(Rc, Rn);
    parity ^= 1;
    r0 = rNew;
    res.iterations = it + 1;
    res.finalRel = r0;
  }
  res.backwardError = backwardError(Rc, Xc, nrhs, nm.normAsym);
  return Xc;
}

std::vector<RefineResult> Plan::solveRefinedGroups(const double* aHost, const double* aDev, const double* B,
                                                   int64_t ldb, int nrhs, double* X, int64_t ldx, bool deviceIo,
                                                   int maxIt, double tol, int ngroups, const int* groupStart) {
  if (!factorized_) throw std::logic_error("symla::SymLDLT::solve: factorize() must be called before solve()");
  if (singular_)
    throw std::logic_error(
        "symla::SymLDLT::solve: factorization is numerically singular (see isSingular()/singularColumns()); refusing "
        "to solve");
  if (ngroups <= 0 || !groupStart || groupStart[0] != 0 || groupStart[ngroups] != n_)
    throw std::invalid_argument("symla: systems must cover the original indices 0..n-1");
  for (int g = 0; g < ngroups; ++g)
    if (groupStart[g + 1] <= groupStart[g]) throw std::invalid_argument("symla: every system needs at least one row");
  std::vector<RefineResult> res(ngroups);
  if (nrhs <= 0 || n_ == 0) return res;
  const double* vals = aDev;
  if (aHost) {
    hipCheck(hipMemcpyAsync(dvals_.p, aHost, sizeof(double) * nnzA_, hipMemcpyHostToDevice, stream_), "values");
    vals = dvals_.p;
  }
  ensureRefineWs(nrhs);
  if (hgroupStart_.size() != static_cast<size_t>(ngroups) + 1 ||
      !std::equal(hgroupStart_.begin(), hgroupStart_.end(), groupStart)) {
    hgroupStart_.assign(groupStart, groupStart + ngroups + 1);
    std::vector<int> rg(n_);
    for (int g = 0; g < ngroups; ++g)
      std::fill(rg.begin() + groupStart[g], rg.begin() + groupStart[g + 1], g);
    drowGroup_.upload(rg, stream_);
  }
  dgAccept_.ensure(ngroups);
  const size_t G = static_cast<size_t>(ngroups);
  dgNorms_.ensure((5 + 3 * static_cast<size_t>(nrhs)) * G);
  unsigned long long* slot = dgNorms_.p;  // slot k of system g at slot[k * G + g]
  const size_t nn = static_cast<size_t>(n_) * nrhs;
  hipCheck(hipMemcpy2DAsync(dBr_.p, sizeof(double) * n_, B, sizeof(double) * ldb, sizeof(double) * n_, nrhs,
                            deviceIo ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_),
           "B copy");
  launchZeroU64(slot, static_cast<int>(5 * G), stream_);
  launchGroupRowSumMax(n_, dstPtr_.p, dstSrc_.p, vals, drowGroup_.p, slot + 0 * G, 1, stream_);
  launchGroupRowSumMax(n_, dsymPtr_.p, dsymSrc_.p, vals, drowGroup_.p, slot + 1 * G, 1, stream_);
  launchGroupAbsMax(n_, nrhs, dBr_.p, n_, drowGroup_.p, slot + 2 * G, 1, stream_);
  launchGatherVals(nnzSym_, dsymSrc_.p, vals, dsymVal_.p, stream_);
  double* Xc = dXout_.p;
  double* Rc = dR_.p;
  solveCore(dBr_.p, n_, nrhs, Xc, n_);
  launchResidual(n_, nrhs, dsymPtr_.p, dsymCol_.p, dsymVal_.p, dBr_.p, Xc, Rc, nullptr, stream_, dXt_.p);
  launchGroupAbsMax(n_, nrhs, Rc, n_, drowGroup_.p, slot + 3 * G, 1, stream_);
  launchGroupAbsMax(n_, nrhs, Xc, n_, drowGroup_.p, slot + 4 * G, 1, stream_);
  std::vector<unsigned long long> h(5 * G);
  auto readSlots = [&](int first, int count) {
    hipCheck(hipMemcpyAsync(h.data() + first * G, slot + first * G, count * G * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, stream_),
             "norms");
    hipCheck(hipStreamSynchronize(stream_), "norms");
  };
  readSlots(0, 5);
  std::vector<double> normA(G), normAsym(G), normB(G);
  for (size_t g = 0; g < G; ++g) {
    normA[g] = bitsToDoubleHost(h[g]);
    normAsym[g] = bitsToDoubleHost(h[G + g]);
    normB[g] = bitsToDoubleHost(h[2 * G + g]);
  }
  auto rel = [&](size_t g) {  // refine.hpp's relative residual of system g from slots 3 and 4
    const double nR = bitsToDoubleHost(h[3 * G + g]), nX = bitsToDoubleHost(h[4 * G + g]);
    if (!std::isfinite(nR) || !std::isfinite(nX)) return std::numeric_limits<double>::infinity();
    const double denom = normA[g] * nX + normB[g];
    return denom > 0.0 ? nR / denom : nR;
  };
  std::vector<double> r0(G);
  std::vector<char> active(G);
  size_t nActive = 0;
  for (size_t g = 0; g < G; ++g) {
    r0[g] = rel(g);
    res[g].initialRel = res[g].finalRel = r0[g];
    active[g] = r0[g] > tol;
    nActive += active[g];
  }
  std::vector<int> accept(G, 0);
  for (int it = 0; it < maxIt && nActive > 0; ++it) {  // refine.hpp:121-141, per system
    solveCore(Rc, n_, nrhs, dDx_.p, n_);
    launchAxpy(static_cast<int64_t>(nn), Xc, dDx_.p, dXn_.p, nullptr, stream_);
    launchResidual(n_, nrhs, dsymPtr_.p, dsymCol_.p, dsymVal_.p, dBr_.p, dXn_.p, dRn_.p, nullptr, stream_,
                   dXt_.p);
    launchZeroU64(slot + 3 * G, static_cast<int>(2 * G), stream_);
    launchGroupAbsMax(n_, nrhs, dRn_.p, n_, drowGroup_.p, slot + 3 * G, 1, stream_);
    launchGroupAbsMax(n_, nrhs, dXn_.p, n_, drowGroup_.p, slot + 4 * G, 1, stream_);
    readSlots(3, 2);
    for (size_t g = 0; g < G; ++g) {
      accept[g] = 0;
      if (!active[g]) continue;
      const double rNew = rel(g);
      if (rNew < r0[g]) {
        accept[g] = 1;
        r0[g] = rNew;
        res[g].iterations = it + 1;
        res[g].finalRel = rNew;
        if (rNew <= tol) active[g] = 0;
      } else {
        active[g] = 0;  // no improvement: this system keeps its current solution
      }
    }
    nActive = 0;
    for (size_t g = 0; g < G; ++g) nActive += active[g];
    hipCheck(hipMemcpyAsync(dgAccept_.p, accept.data(), G * sizeof(int), hipMemcpyHostToDevice, stream_), "accept");
    launchGroupSelect(n_, nrhs, drowGroup_.p, dgAccept_.p, dXn_.p, Xc, n_, stream_);
    launchGroupSelect(n_, nrhs, drowGroup_.p, dgAccept_.p, dRn_.p, Rc, n_, stream_);
    hipCheck(hipStreamSynchronize(stream_), "accept");  // `accept` is rewritten next iteration
  }
  // backward error per system: max over its columns with its symmetric infinity norm
  unsigned long long* cR = slot + 5 * G;
  unsigned long long* cX = cR + static_cast<size_t>(nrhs) * G;
  unsigned long long* cB = cX + static_cast<size_t>(nrhs) * G;
  launchZeroU64(cR, static_cast<int>(3 * nrhs * G), stream_);
  launchGroupAbsMaxCols(n_, nrhs, Rc, n_, drowGroup_.p, cR, 1, ngroups, stream_);
  launchGroupAbsMaxCols(n_, nrhs, Xc, n_, drowGroup_.p, cX, 1, ngroups, stream_);
  launchGroupAbsMaxCols(n_, nrhs, dBr_.p, n_, drowGroup_.p, cB, 1, ngroups, stream_);
  std::vector<unsigned long long> hc(3 * static_cast<size_t>(nrhs) * G);
  hipCheck(hipMemcpyAsync(hc.data(), cR, hc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_),
           "be");
  hipCheck(hipMemcpy2DAsync(X, sizeof(double) * ldx, Xc, sizeof(double) * n_, sizeof(double) * n_, nrhs,
                            deviceIo ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream_),
           "X out");
  hipCheck(hipStreamSynchronize(stream_), "refine");
  for (size_t g = 0; g < G; ++g) {
    double be = 0.0;
    for (int c = 0; c < nrhs; ++c) {
      const size_t k = static_cast<size_t>(c) * G + g;
      const double nr = bitsToDoubleHost(hc[k]), nx = bitsToDoubleHost(hc[nrhs * G + k]),
                   nb = bitsToDoubleHost(hc[2 * nrhs * G + k]);
      const double den = normAsym[g] * nx + nb;
      const double bc = (!std::isfinite(nr) || !std::isfinite(nx)) ? std::numeric_limits<double>::infinity()
                        : den > 0.0                                 ? nr / den
                                                                    : nr;
      be = std::max(be, bc);
    }
    res[g].backwardError = be;
  }
  checkSolveError();
  return res;
}

void Plan::groupInertia(int ngroups, const int* groupStart, int* inertia3, int* singular) {
  if (!factorized_) throw std::logic_error("symla: no factorization");
  if (ngroups <= 0 || !groupStart || !inertia3) throw std::invalid_argument("symla: bad group arguments");
  if (groupStart[0] != 0 || groupStart[ngroups] != n_)
    throw std::invalid_argument("symla: groups must cover the original indices 0..n-1");
  for (int g = 0; g < ngroups; ++g)
    if (groupStart[g + 1] < groupStart[g]) throw std::invalid_argument("symla: group boundaries must not decrease");
  auto groupOf = [&](int orig) {
    return static_cast<int>(std::upper_bound(groupStart, groupStart + ngroups + 1, orig) - groupStart) - 1;
  };
  std::vector<int> fi(3 * static_cast<size_t>(numSN_));
  hipCheck(hipMemcpy(fi.data(), dinertia_.p, sizeof(int) * fi.size(), hipMemcpyDeviceToHost), "front inertia");
  std::fill(inertia3, inertia3 + 3 * static_cast<size_t>(ngroups), 0);
  for (int s = 0; s < numSN_; ++s) {
    const int g = groupOf(perm_[hfirst_[s]]);
    for (int q = 0; q < 3; ++q) inertia3[3 * g + q] += fi[3 * static_cast<size_t>(s) + q];
  }
  if (singular) {
    std::fill(singular, singular + ngroups, 0);
    for (int c : singularColsOrig_) ++singular[groupOf(c)];
  }
}

void Plan::graphRecord() {
  if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graphExec_ = nullptr;
  graph_ = nullptr;
  for (auto& e : stepExec_) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
  const GraphArgs& g = graphArgs_;
  cbBf16x3_ = g.opts.bf16x3 && g.opts.precision == 1;
  if (g.opts.precision != prec_) {
    prec_ = g.opts.precision;
    layoutDirty_ = true;
  }
  if (layoutDirty_) {
    buildStaticLayout();
    uploadLayout();
    layoutDirty_ = false;
  }
  ensureSolveWs(g.nrhs);
  if (g.refine) ensureRefineWs(g.nrhs);
  const bool prof = profiling_;
  profiling_ = false;  // no event records inside the graph
  FactorParams prm = params(g.opts, false, 0.0, nullptr);
  // records what `body` launches on the plan's stream as one graph
  auto record = [&](auto&& body, hipGraph_t* graphOut, hipGraphExec_t* execOut) {
    hipGraph_t gr = nullptr;
    hipCheck(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
      body();
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(stream_, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      profiling_ = prof;
      throw;
    }
    hipCheck(hipStreamEndCapture(stream_, &gr), "end capture");
    const hipError_t e = hipGraphInstantiate(execOut, gr, nullptr, nullptr, 0);
    if (graphOut) {
      *graphOut = gr;
    } else {
      (void)hipGraphDestroy(gr);
    }
    hipCheck(e, "instantiate graph");
  };
  if (!g.refine) {
    record([&] {
      runStatic(g.vals, prm);
      solveCore(g.B, g.ldb, g.nrhs, g.X, g.ldx);
    }, &graph_, &graphExec_);
  } else {
    record([&] {
      runStatic(g.vals, prm);
      hipCheck(hipMemcpy2DAsync(dBr_.p, sizeof(double) * n_, g.B, sizeof(double) * g.ldb, sizeof(double) * n_,
                                g.nrhs, hipMemcpyDeviceToDevice, stream_),
               "B copy");
      refineSetup(g.vals, g.nrhs);
    }, &graph_, &graphExec_);
    record([&] { refineStep(dXout_.p, dR_.p, dXn_.p, dRn_.p, g.nrhs); }, nullptr, &stepExec_[0]);
    record([&] { refineStep(dXn_.p, dRn_.p, dXout_.p, dR_.p, g.nrhs); }, nullptr, &stepExec_[1]);
  }
  profiling_ = prof;
  graphEpoch_ = g_devAllocEpoch.load();
  graphStale_ = false;
}

void Plan::graphCapture(const double* dVals, const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx,
                        const FactorOptions& opts) {
  if (opts.mode != 0) throw std::invalid_argument("symla: graph capture covers threshold-pivoting factorizations");
  if (!dVals || !dB || !dX) throw std::invalid_argument("symla: graph capture needs device buffers");
  if (nrhs <= 0 || ldb < n_ || ldx < n_) throw std::invalid_argument("symla: bad right-hand-side layout");
  if (opts.precision != 0 && opts.precision != 1) throw std::invalid_argument("symla: unknown precision");
  graphArgs_ = GraphArgs{dVals, dB, ldb, nrhs, dX, ldx, opts};
  graphRecord();
}

void Plan::graphCaptureRefined(const double* dVals, const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx,
                               const FactorOptions& opts, int maxIt, double tol) {
  if (opts.mode != 0) throw std::invalid_argument("symla: graph capture covers threshold-pivoting factorizations");
  if (!dVals || !dB || !dX) throw std::invalid_argument("symla: graph capture needs device buffers");
  if (nrhs <= 0 || ldb < n_ || ldx < n_) throw std::invalid_argument("symla: bad right-hand-side layout");
  if (opts.precision != 0 && opts.precision != 1) throw std::invalid_argument("symla: unknown precision");
  if (maxIt < 0) throw std::invalid_argument("symla: max_iterations must be non-negative");
  graphArgs_ = GraphArgs{dVals, dB, ldb, nrhs, dX, ldx, opts, true, maxIt, tol};
  graphRecord();
}

void Plan::graphRelease() {
  if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graphExec_ = nullptr;
  graph_ = nullptr;
  for (auto& e : stepExec_) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
  graphStale_ = true;
}

void Plan::graphReplay(FactorInfo& info, RefineResult* rres) {
  if (!graphExec_) throw std::logic_error("symla: no recorded graph (call symla_graph_capture first)");
  if (rres && !graphArgs_.refine)
    throw std::logic_error("symla: the recorded graph has no refinement (use symla_graph_capture_refined)");
  factorized_ = false;
  // a repair pass, another solve size or a new tuning may have moved or
  // changed what the graph refers to
  if (layoutDirty_ || graphArgs_.opts.precision != prec_ || !graphCurrent()) {
    graphRecord();
  } else {
    ensureSolveWs(graphArgs_.nrhs);
    if (graphArgs_.refine) ensureRefineWs(graphArgs_.nrhs);
    if (!graphCurrent()) graphRecord();
  }
  info = FactorInfo{};
  singular_ = false;
  singularColsOrig_.clear();
  hipCheck(hipEventRecord(ev0_, stream_), "event");
  hipCheck(hipGraphLaunch(graphExec_, stream_), "graph launch");
  bool anyDelay = false;
  collectStats(info, anyDelay);
  info.precisionUsed = precisionCode();
  checkSolveError();
  if (anyDelay) {
    // the recorded schedule assumed no delays: repair and solve again
    FactorParams prm = params(graphArgs_.opts, false, 0.0, nullptr);
    runDynamic(graphArgs_.vals, prm);
    collectStats(info, anyDelay);
    info.repaired = true;
    info.singular = singular_;
    info.nSingular = static_cast<int>(singularColsOrig_.size());
    if (!singular_ && !graphArgs_.refine) {
      solveCore(graphArgs_.B, graphArgs_.ldb, graphArgs_.nrhs, graphArgs_.X, graphArgs_.ldx);
      hipCheck(hipStreamSynchronize(stream_), "repaired solve");
      checkSolveError();
    } else if (!singular_) {
      factorized_ = true;
      RefineResult r = solveRefined(nullptr, graphArgs_.vals, graphArgs_.B, graphArgs_.ldb, graphArgs_.nrhs,
                                    graphArgs_.X, graphArgs_.ldx, true, graphArgs_.maxIt, graphArgs_.tol);
      if (rres) *rres = r;
    }
  } else if (graphArgs_.refine && !singular_) {
    // the recorded steps, while refine.hpp's tests ask for more
    RefineResult r;
    const double* Xc = refineLoop(readNorms(), graphArgs_.nrhs, graphArgs_.maxIt, graphArgs_.tol, stepExec_, r);
    hipCheck(hipMemcpy2DAsync(graphArgs_.X, sizeof(double) * graphArgs_.ldx, Xc, sizeof(double) * n_,
                              sizeof(double) * n_, graphArgs_.nrhs, hipMemcpyDeviceToDevice, stream_),
             "X out");
    hipCheck(hipStreamSynchronize(stream_), "refine");
    checkSolveError();
    if (rres) *rres = r;
  }
  hipCheck(hipEventRecord(ev1_, stream_), "event");
  hipCheck(hipEventSynchronize(ev1_), "event");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0_, ev1_);
  factorMs_ = ms;
  solveMs_ = 0.0;
  factorized_ = true;
}

void Plan::setTuning(const Tuning& t) {
  if (t.smallMax < 0 || t.smallMax > kSmallFrontMax)
    throw std::invalid_argument("symla: small-front bound must lie in [0, 192]");
  if (t.asmTile < 16 || t.asmTile > 1024 || (t.asmTile & (t.asmTile - 1)) != 0)
    throw std::invalid_argument("symla: assembly tile must be a power of two in [16, 1024]");
  if (t.specMinRows < 0) throw std::invalid_argument("symla: speculative-panel row bound must be >= 0");
  tune_ = t;
  buildSchedule();
  graphStale_ = true;  // a recorded graph holds the old schedule: re-record on replay
}

std::string Plan::tuneKey(int nrhs, int precision) const {
  hipDeviceProp_t prop{};
  int dev = 0;
  hipCheck(hipGetDevice(&dev), "hipGetDevice");
  hipCheck(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  // front-order histogram in powers of two: plans with the same histogram share a tuning
  int hist[32] = {};
  for (int s = 0; s < numSN_; ++s) {
    int b = 0;
    while ((2 << b) <= hm_[s] && b < 31) ++b;
    ++hist[b];
  }
  std::string key = std::string(prop.gcnArchName) + "|p" + std::to_string(precision) + "|r" + std::to_string(nrhs) +
                    "|h";
  for (int b = 0; b < 32; ++b)
    if (hist[b]) key += std::to_string(b) + ":" + std::to_string(hist[b]) + ",";
  return key;
}

std::vector<Plan::TuneTrial> Plan::autotune(const double* dVals, int nrhs, int precision, int reps,
                                            const char* cachePath, bool* fromCache) {
  if (!dVals) throw std::invalid_argument("symla: autotune needs device values");
  if (nrhs <= 0 || reps <= 0) throw std::invalid_argument("symla: autotune needs nrhs > 0 and reps > 0");
  if (precision < 0 || precision > 2) throw std::invalid_argument("symla: unknown precision");
  if (fromCache) *fromCache = false;
  const std::string key = cachePath ? tuneKey(nrhs, precision) : std::string();
  if (cachePath) {  // one line per entry: key, then the five tuning fields
    std::ifstream in(cachePath);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string k;
      Tuning t;
      if (ls >> k >> t.smallMax >> t.asmTile >> t.specPanel >> t.specMinRows >> t.cbMinRows && k == key) {
        setTuning(t);
        if (fromCache) *fromCache = true;
        return {};
      }
    }
  }
  std::vector<double> hb(static_cast<size_t>(n_) * nrhs);
  uint64_t st = 0x9E3779B97F4A7C15ull;
  for (double& v : hb) {  // fixed pseudo-random right-hand sides in [-1, 1)
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    v = static_cast<double>(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
  DBuf<double> dB, dXt;
  dB.alloc(hb.size());
  dXt.alloc(hb.size());
  hipCheck(hipMemcpy(dB.p, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice), "autotune rhs");
  FactorOptions o;
  o.precision = precision == 2 ? 1 : precision;  // SYMLA_PREC_BF16X3: FP32 fronts, BF16x3 contribution blocks
  o.bf16x3 = precision == 2;
  std::vector<TuneTrial> trials;
  auto same = [](const Tuning& a, const Tuning& b) {
    return a.smallMax == b.smallMax && a.asmTile == b.asmTile && a.specPanel == b.specPanel &&
           a.specMinRows == b.specMinRows && a.cbMinRows == b.cbMinRows;
  };
  auto timeOf = [&](const Tuning& t) {
    for (const auto& tr : trials)
      if (same(tr.t, t)) return tr.ms;
    setTuning(t);
    FactorInfo info;
    factorize(nullptr, dVals, o, info);  // warm-up: workspaces, kernels
    solve(dB.p, n_, nrhs, dXt.p, n_, true);
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      factorize(nullptr, dVals, o, info);
      solve(dB.p, n_, nrhs, dXt.p, n_, true);
      const auto t1 = std::chrono::steady_clock::now();
      best = std::min(best, std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    trials.push_back(TuneTrial{t, best});
    return best;
  };
  // coordinate descent from the current tuning: each knob in turn over its
  // candidates with the others fixed, the fastest kept; repeated until a
  // sweep changes nothing (at most three sweeps)
  Tuning cur = tune_;
  double curMs = timeOf(cur);
  const int smalls[] = {0, 32, 64, 96, 128, 192};
  const int tiles[] = {32, 64, 128};
  const int specs[] = {-1, 0, 1024, 2048, 4096};  // -1: no speculative step
  const int cbs[] = {-1, 1024, 2048, 4096};
  for (int sweep = 0; sweep < 3; ++sweep) {
    const Tuning before = cur;
    auto tryKnob = [&](auto&& apply, const int* vals, int nv) {
      for (int i = 0; i < nv; ++i) {
        Tuning t = cur;
        apply(t, vals[i]);
        const double ms = timeOf(t);
        if (ms < curMs) {
          curMs = ms;
          cur = t;
        }
      }
    };
    tryKnob([](Tuning& t, int v) { t.smallMax = v; }, smalls, 6);
    tryKnob([](Tuning& t, int v) { t.asmTile = v; }, tiles, 3);
    tryKnob(
        [](Tuning& t, int v) {
          t.specPanel = v >= 0;
          if (v >= 0) t.specMinRows = v;
        },
        specs, 5);
    tryKnob([](Tuning& t, int v) { t.cbMinRows = v; }, cbs, 4);
    if (same(before, cur)) break;
  }
  setTuning(cur);
  if (cachePath) {
    std::ofstream out(cachePath, std::ios::app);
    out << key << ' ' << cur.smallMax << ' ' << cur.asmTile << ' ' << cur.specPanel << ' ' << cur.specMinRows << ' '
        << cur.cbMinRows << '\n';
  }
  return trials;
}

void Plan::frontView(int si, int* m, int* npiv, int* rows, double* L, double* dd, double* dof, int* ps, int* pk,
                     int* nblocks) {
  if (!factorized_) throw std::logic_error("symla: no factorization");
  if (si < 0 || si >= numSN_) throw std::invalid_argument("symla: front index out of range");
  int nf = 0;
  hipCheck(hipMemcpy(&nf, dnF_.p + si, sizeof(int), hipMemcpyDeviceToHost), "view");
  const int mm = hm_[si];
  if (m) *m = mm;
  if (npiv) *npiv = nf;
  const int64_t io = hidxOff_[si];
  if (rows) hipCheck(hipMemcpy(rows, drowIdx_.p + io, sizeof(int) * mm, hipMemcpyDeviceToHost), "view rows");
  if (L && nf > 0) {
    const size_t cnt = static_cast<size_t>(mm) * nf;
    if (prec_) {
      std::vector<float> tmp(cnt);
      hipCheck(hipMemcpy(tmp.data(), hF_[si], sizeof(float) * cnt, hipMemcpyDeviceToHost), "view L");
      for (size_t e = 0; e < cnt; ++e) L[e] = tmp[e];
    } else {
      hipCheck(hipMemcpy(L, hF_[si], sizeof(double) * cnt, hipMemcpyDeviceToHost), "view L");
    }
    for (int j = 0; j < nf; ++j)
      for (int i = 0; i < j && i < mm; ++i) L[i + static_cast<size_t>(j) * mm] = 0.0;  // upper part is not stored
  }
  std::vector<int> kinds(nf);
  std::vector<double> dg(nf), dfo(nf);
  if (nf > 0) {
    hipCheck(hipMemcpy(kinds.data(), dpkind_.p + io, sizeof(int) * nf, hipMemcpyDeviceToHost), "view");
    hipCheck(hipMemcpy(dg.data(), ddDiag_.p + io, sizeof(double) * nf, hipMemcpyDeviceToHost), "view");
    hipCheck(hipMemcpy(dfo.data(), ddOff_.p + io, sizeof(double) * nf, hipMemcpyDeviceToHost), "view");
  }
  int nb = 0;
  for (int i = 0; i < nf; ++i) {
    if (dd) dd[i] = dg[i];
    if (dof) dof[i] = (kinds[i] == 2) ? dfo[i] : 0.0;
    if (kinds[i] == 1 || kinds[i] == 2) {
      if (ps) ps[nb] = i;
      if (pk) pk[nb] = kinds[i];
      ++nb;
    }
  }
  if (nblocks) *nblocks = nb;
}

}  // namespace symla_hip
