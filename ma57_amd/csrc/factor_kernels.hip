// factor_kernels.hip -- gfx950 kernels of the multifrontal numeric
// factorization: front reset, assembly (zero + value scatter + extend-add),
// small-front Bunch-Kaufman in LDS, and the large-front panel path (panel
// pivoting on the eligible rows, row-parallel multipliers for the remaining
// rows, FP64 MFMA symmetric rank-p trailing update).
//
// Pivot decision logic follows the reference's DenseLDLT::factor
// (include/symla/dense_kernel.hpp:585-792) and DenseLDLT::factorStatic
// (dense_kernel.hpp:962-1035); see the comments at each step.
#include <climits>

#include "device.hpp"
#include "kernels.hpp"

#include <cfloat>
#include <cmath>

namespace symla_hip {

namespace {

constexpr double kAlpha = 0.64038820320220756872767623199676;  // (1 + sqrt(17)) / 8, dense_kernel.hpp:400

__device__ inline double bitsToDouble(unsigned long long b) { return __longlong_as_double(static_cast<long long>(b)); }
__device__ inline unsigned long long doubleToBits(double d) {
  return static_cast<unsigned long long>(__double_as_longlong(d));
}
// max of non-negative doubles via the ordering of their bit patterns
__device__ inline void atomicMaxNonNeg(unsigned long long* addr, double v) {
  if (v > 0.0) atomicMax(addr, doubleToBits(v));
}

// Candidate combine for the lambda search: larger value wins; equal values
// keep the smaller row (the reference scans rows upward with a strict '>').
__device__ inline void argmaxCombine(double& v, int& i, double v2, int i2) {
  if (i2 < 0) return;
  if (i < 0 || v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

// Wavefront reductions of the device library (cross-lane DPP steps instead
// of LDS permutes).
extern "C" __device__ __attribute__((const)) double __ockl_wfred_max_f64(double);
extern "C" __device__ __attribute__((const)) int __ockl_wfred_min_i32(int);

// argmaxCombine over the 64 lanes of a wave (every lane active; a lane with
// no candidate holds v = 0, idx = -1): the largest value and, among lanes
// holding it, the smallest row.
__device__ inline void waveArgmax(double& v, int& idx) {
  const double wm = __ockl_wfred_max_f64(v);
  const int wi = __ockl_wfred_min_i32((idx >= 0 && v == wm) ? idx : INT_MAX);
  v = wm;
  idx = (wi == INT_MAX) ? -1 : wi;
}

template <int NT>
__device__ inline void blockArgmax(double& v, int& idx, double* sv, int* si) {
  waveArgmax(v, idx);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sv[w] = v;
    si[w] = idx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bv = sv[0];
    int bi = si[0];
    for (int t = 1; t < NT / 64; ++t) argmaxCombine(bv, bi, sv[t], si[t]);
    sv[NT / 64] = bv;
    si[NT / 64] = bi;
  }
  __syncthreads();
  v = sv[NT / 64];
  idx = si[NT / 64];
  __syncthreads();
}

template <int NT>
__device__ inline double blockMax(double v, double* sv) {
  v = __ockl_wfred_max_f64(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sv[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sv[0];
    for (int t = 1; t < NT / 64; ++t) b = fmax(b, sv[t]);
    sv[NT / 64] = b;
  }
  __syncthreads();
  v = sv[NT / 64];
  __syncthreads();
  return v;
}

__device__ inline void countInertia(double d, int& pos, int& neg, int& zero) {
  if (d > 0)
    ++pos;
  else if (d < 0)
    ++neg;
  else
    ++zero;
}

// Inertia of an accepted 2x2 block (dense_kernel.hpp:753-768).
__device__ inline void countInertia2(double d11, double d22, double det, int& pos, int& neg) {
  if (det < 0) {
    ++pos;
    ++neg;
  } else {
    const double tr = d11 + d22;
    if (tr > 0)
      pos += 2;
    else if (tr < 0)
      neg += 2;
    else {
      ++pos;
      ++neg;
    }
  }
}

// Static pivot (dense_kernel.hpp:982-1011): returns the pivot to use and the
// perturbation applied.
__device__ inline double staticPivot(double a_kk, int sign, double delta, double floorv, double& applied) {
  bool need;
  if (sign > 0)
    need = a_kk < floorv;
  else if (sign < 0)
    need = a_kk > -floorv;
  else
    need = fabs(a_kk) < floorv;
  double d = a_kk;
  applied = 0.0;
  if (need) {
    const double s = (sign != 0) ? static_cast<double>(sign) : (a_kk >= 0.0 ? 1.0 : -1.0);
    d = a_kk + s * delta;
    applied = d - a_kk;
  }
  return d;
}

// Front owning flat block `b` of a batched launch: start[f] <= b < start[f+1].
__device__ inline int findSegment(const int* start, int count, int b) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

}  // namespace

// ---------------------------------------------------------------------------
// Reset per-front state for a factorization (no memset nodes: safe to record).
// ---------------------------------------------------------------------------
__global__ void k_reset_fronts(int numSN, DevFronts fd) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < numSN; s += gridDim.x * blockDim.x) {
    fd.nF[s] = 0;
    fd.inertia[3 * s] = 0;
    fd.inertia[3 * s + 1] = 0;
    fd.inertia[3 * s + 2] = 0;
    fd.origMaxBits[s] = 0ull;
    fd.runMaxBits[s] = 0ull;
    fd.kcur[s] = 0;
    fd.pcount[s] = 0;
    fd.stopped[s] = 0;
    fd.cbDone[s] = 0;
    fd.pert[s] = 0.0;
    fd.npert[s] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && fd.schurFlops) *fd.schurFlops = 0ull;
}

__global__ void k_reset_perm(const int* list, int count, DevFronts fd) {
  if (static_cast<int>(blockIdx.x) >= count) return;
  const int s = list[blockIdx.x];
  const int m = fd.m[s];
  const int64_t io = fd.idxOff[s];
  for (int i = threadIdx.x; i < m; i += blockDim.x) fd.perm[io + i] = i;
}

// ---------------------------------------------------------------------------
// Assembly of one column tile [c0, c1) of one front (multifrontal.hpp:498-535):
// zero the lower triangle, add A's own-column entries, then extend-add every
// child's contribution block in child order.  One workgroup owns each
// (front, column tile), so every target entry has exactly one writer and the
// summation order is fixed.  The tile's max |F| then goes into the front's
// orig_max (dense_kernel.hpp:413-415: the max over the assembled front is the
// max over its tiles).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_assemble(const int4* tasks, int ntasks, DevFronts fd,
                                                   const double* __restrict__ aVals, int monotone) {
  const int task = blockIdx.x;
  if (task >= ntasks) return;
  const int s = tasks[task].x;
  const int c0 = tasks[task].y;
  const int m = fd.m[s];
  const int c1 = min(tasks[task].z, m);
  T* F = static_cast<T*>(fd.F[s]);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;

  for (int j = c0 + wv; j < c1; j += nw) {
    T* col = F + static_cast<size_t>(j) * m;
    for (int i = j + lane; i < m; i += 64) col[i] = T(0);
  }
  __syncthreads();

  const int nc = fd.ncols[s];
  const int fc = fd.firstCol[s];
  // A's entries of the own columns, four at a time: the rows and values of a
  // batch are loaded together, then added in entry order (a caller's
  // repeated entry lands on the same row and adds in turn)
  for (int j = c0 + threadIdx.x; j < min(c1, nc); j += blockDim.x) {
    const int gcol = fc + j;
    T* col = F + static_cast<size_t>(j) * m;
    const int64_t tb = fd.acolPtr[gcol], te = fd.acolPtr[gcol + 1];
    for (int64_t t0 = tb; t0 < te; t0 += 4) {
      int row[4];
      double av[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t t = min(t0 + u, te - 1);
        row[u] = fd.aRow[t];
        av[u] = aVals[fd.aSrc[t]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (t0 + u < te) col[row[u]] += static_cast<T>(av[u]);
    }
  }
  __syncthreads();

  for (int ci = fd.childPtr[s]; ci < fd.childPtr[s + 1]; ++ci) {
    const int c = fd.childList[ci];
    const int mc = fd.m[c];
    const int nfc = fd.nF[c];
    const T* Fc = static_cast<const T*>(fd.F[c]);
    const int* rel = fd.relmap + fd.idxOff[c];
    int bLo = nfc, bHi = mc;
    if (monotone) {
      // rel[] is non-decreasing over [nfc, mc): find the columns landing in [c0, c1)
      int lo = nfc, hi = mc;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rel[mid] < c0)
          lo = mid + 1;
        else
          hi = mid;
      }
      bLo = lo;
      hi = mc;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rel[mid] < c1)
          lo = mid + 1;
        else
          hi = mid;
      }
      bHi = lo;
    }
    for (int b = bLo + wv; b < bHi; b += nw) {
      const int lb = rel[b];
      if (lb < 0) continue;
      const T* ccol = Fc + static_cast<size_t>(b) * mc;
      // four rows per lane at a time, every load of a batch in flight before
      // any add (a skipped row reads the child's own entry (b, b) instead);
      // the relabelling is one to one, so no two rows meet in the parent
      for (int a0 = b + lane; a0 < mc; a0 += 256) {
        int la[4];
        T cv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int a = min(a0 + 64 * u, mc - 1);
          la[u] = rel[a];
          cv[u] = ccol[a];
        }
        size_t idx[4];
        bool ok[4];
        T old[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int tr = la[u] > lb ? la[u] : lb;
          const int tc = la[u] > lb ? lb : la[u];
          ok[u] = a0 + 64 * u < mc && la[u] >= 0 && tc >= c0 && tc < c1;
          idx[u] = tr + static_cast<size_t>(tc) * m;
          old[u] = ok[u] ? F[idx[u]] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (ok[u]) F[idx[u]] = old[u] + cv[u];
      }
    }
    __syncthreads();
  }

  __shared__ double wmax[4];
  double v = 0.0;
  for (int j = c0 + wv; j < c1; j += nw) {
    const T* col = F + static_cast<size_t>(j) * m;
    for (int i = j + lane; i < m; i += 64) v = fmax(v, fabs(static_cast<double>(col[i])));
  }
  v = __ockl_wfred_max_f64(v);
  if (lane == 0) wmax[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) v = fmax(v, wmax[w]);
    atomicMaxNonNeg(fd.origMaxBits + s, v);
  }
}

// ---------------------------------------------------------------------------
// Small fronts: the packed lower triangle lives in LDS; right-looking
// Bunch-Kaufman with the rank-1/rank-2 update applied eagerly after every
// pivot, which is the reference's panel_size = 1 schedule for fronts with
// fewer than 256 eligible columns (multifrontal.hpp:564-566).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kSmallThreads) void k_factor_small(const int* list, int count, DevFronts fd,
                                                                 FactorParams prm) {
  constexpr int NT = kSmallThreads;
  extern __shared__ double smem[];
  if (static_cast<int>(blockIdx.x) >= count) return;
  const int s = list[blockIdx.x];
  const int m = fd.m[s];
  const int nE = fd.nE[s];
  T* F = static_cast<T*>(fd.F[s]);
  const int64_t io = fd.idxOff[s];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = NT / 64;

  double* redV = smem;                                    // NT/64 + 1
  int* redI = reinterpret_cast<int*>(smem + 32);          // NT/64 + 1
  double* A = smem + 64;                                  // packed lower, m(m+1)/2
  const int npk = m * (m + 1) / 2;
  int* perm = reinterpret_cast<int*>(A + npk);            // m

  auto P = [m](int i, int j) { return j * m - (j * (j - 1)) / 2 + (i - j); };

  double om = 0.0;
  for (int j = wv; j < m; j += nw) {
    const T* col = F + static_cast<size_t>(j) * m;
    const int base = P(j, j) - j;
    for (int i = j + lane; i < m; i += 64) {
      const double v = col[i];
      A[base + i] = v;
      om = fmax(om, fabs(v));
    }
  }
  for (int i = tid; i < m; i += NT) perm[i] = i;
  __syncthreads();
  om = blockMax<NT>(om, redV);

  const bool stat = prm.staticMode != 0;
  double origMax = om;
  if (!stat && origMax < prm.zeroTol) origMax = 1.0;        // dense_kernel.hpp:414
  if (stat && origMax < prm.staticAbsFloor) origMax = 1.0;  // dense_kernel.hpp:935
  double tmax = origMax;  // thread-local contribution to running_max

  const double delta = prm.staticDelta > 0.0 ? prm.staticDelta : sqrt(DBL_EPSILON) * origMax;
  const double sfloor = fmax(prm.staticAbsFloor, prm.staticRelFloor * origMax);

  int npos = 0, nneg = 0, nzero = 0, npert = 0;
  double pert = 0.0;
  int k = 0;
  bool stopped = false;

  // symmetric interchange of positions a < b over the whole front
  auto swapSym = [&](int a, int b) {
    for (int t = tid; t < m; t += NT) {
      int x = -1, y = -1;
      if (t < a) {
        x = P(a, t);
        y = P(b, t);
      } else if (t == a) {
        x = P(a, a);
        y = P(b, b);
      } else if (t < b) {
        x = P(t, a);
        y = P(b, t);
      } else if (t > b) {
        x = P(t, a);
        y = P(t, b);
      }
      if (x >= 0) {
        const double tv = A[x];
        A[x] = A[y];
        A[y] = tv;
      }
    }
    if (tid == 0) {
      const int tp = perm[a];
      perm[a] = perm[b];
      perm[b] = tp;
    }
    __syncthreads();
  };

  // accept a 1x1 pivot at k: scale column, eager rank-1 update
  auto accept1 = [&](double d) {
    if (tid == 0) {
      fd.dDiag[io + k] = d;
      fd.dOff[io + k] = 0.0;
      fd.pkind[io + k] = 1;
    }
    countInertia(d, npos, nneg, nzero);
    tmax = fmax(tmax, fabs(d));
    const int ck = P(k, k) - k;
    for (int i = k + 1 + tid; i < m; i += NT) {
      const double c = A[ck + i];
      tmax = fmax(tmax, fabs(c));
      A[ck + i] = c / d;
    }
    __syncthreads();
    for (int j = k + 1 + wv; j < m; j += nw) {
      const double lj = A[ck + j];
      const int cj = P(j, j) - j;
      for (int i = j + lane; i < m; i += 64) {
        const double li = A[ck + i];
        const double v = fma(-(li * d), lj, A[cj + i]);
        A[cj + i] = v;
        tmax = fmax(tmax, fabs(v));
      }
    }
    __syncthreads();
    ++k;
  };

  if (stat) {
    const int* sg = prm.expectedSign;
    while (k < nE) {
      const double akk = A[P(k, k)];
      const int sign = sg ? sg[fd.frontIdx[io + k]] : 0;
      double applied;
      const double d = staticPivot(akk, sign, delta, sfloor, applied);
      if (applied != 0.0) {
        pert += fabs(applied);
        ++npert;
      }
      accept1(d);
    }
  } else {
    while (k < nE) {
      // lambda = max_{k<i<nE} |A(i,k)| at row r (dense_kernel.hpp:588-596)
      double lv = 0.0;
      int li = -1;
      {
        const int ck = P(k, k) - k;
        for (int i = k + 1 + tid; i < nE; i += NT) {
          const double v = fabs(A[ck + i]);
          if (v > 0.0) argmaxCombine(lv, li, v, i);
        }
      }
      blockArgmax<NT>(lv, li, redV, redI);
      const double akk = A[P(k, k)];
      int action;  // 0 stop, 1 1x1 at k, 2 1x1 after swap k<->r, 3 2x2 at (k, r)
      const int r = li;
      if (lv <= prm.zeroTol) {  // dense_kernel.hpp:601-616
        const double isolatedFloor = fmax(prm.zeroTol, prm.relFloor * origMax);
        action = (fabs(akk) <= isolatedFloor) ? 0 : 1;
      } else if (fabs(akk) >= kAlpha * lv) {  // :617
        action = 1;
      } else {
        double sg = 0.0;  // sigma over the eligible rows of column r, i != r (:628-633)
        for (int i = k + tid; i < nE; i += NT) {
          if (i == r) continue;
          const double v = fabs(i > r ? A[P(i, r)] : A[P(r, i)]);
          sg = fmax(sg, v);
        }
        sg = blockMax<NT>(sg, redV);
        if (fabs(akk) * sg >= kAlpha * lv * lv)  // :641-645
          action = 1;
        else if (fabs(A[P(r, r)]) >= kAlpha * sg)
          action = 2;
        else
          action = 3;
      }
      if (action == 0) {
        stopped = true;
        break;
      }
      if (action == 2) swapSym(k, r);
      if (action == 1 || action == 2) {
        accept1(A[P(k, k)]);
        continue;
      }
      // 2x2 pivot (dense_kernel.hpp:721-792)
      if (r != k + 1) swapSym(k + 1, r);
      const double d11 = A[P(k, k)];
      const double d21 = A[P(k + 1, k)];
      const double d22 = A[P(k + 1, k + 1)];
      const double det = d11 * d22 - d21 * d21;
      const double rmax = blockMax<NT>(tmax, redV);
      if (fabs(det) <= prm.zeroTol * fmax(1.0, rmax) * fmax(1.0, rmax)) {
        stopped = true;
        break;
      }
      tmax = fmax(tmax, fmax(fabs(d11), fmax(fabs(d21), fabs(d22))));
      if (tid == 0) {
        fd.dDiag[io + k] = d11;
        fd.dOff[io + k] = d21;
        fd.dDiag[io + k + 1] = d22;
        fd.dOff[io + k + 1] = 0.0;
        fd.pkind[io + k] = 2;
        fd.pkind[io + k + 1] = 0;
      }
      countInertia2(d11, d22, det, npos, nneg);
      const double inv11 = d22 / det, inv21 = -d21 / det, inv22 = d11 / det;
      const int c0 = P(k, k) - k, c1 = P(k + 1, k + 1) - (k + 1);
      for (int i = k + 2 + tid; i < m; i += NT) {
        const double a0 = A[c0 + i], a1 = A[c1 + i];
        tmax = fmax(tmax, fmax(fabs(a0), fabs(a1)));
        A[c0 + i] = a0 * inv11 + a1 * inv21;
        A[c1 + i] = a0 * inv21 + a1 * inv22;
      }
      __syncthreads();
      for (int j = k + 2 + wv; j < m; j += nw) {
        const double lj0 = A[c0 + j], lj1 = A[c1 + j];
        const int cj = P(j, j) - j;
        for (int i = j + lane; i < m; i += 64) {
          const double li0 = A[c0 + i], li1 = A[c1 + i];
          const double w0 = li0 * d11 + li1 * d21;
          const double w1 = li0 * d21 + li1 * d22;
          const double v = A[cj + i] - (w0 * lj0 + w1 * lj1);
          A[cj + i] = v;
          tmax = fmax(tmax, fabs(v));
        }
      }
      __syncthreads();
      k += 2;
    }
  }
  (void)stopped;

  // write back
  for (int j = wv; j < m; j += nw) {
    T* col = F + static_cast<size_t>(j) * m;
    const int base = P(j, j) - j;
    for (int i = j + lane; i < m; i += 64) col[i] = static_cast<T>(A[base + i]);
  }
  for (int i = tid; i < m; i += NT) fd.perm[io + i] = perm[i];
  const double rm = blockMax<NT>(tmax, redV);
  if (tid == 0) {
    fd.nF[s] = k;
    fd.inertia[3 * s] = npos;
    fd.inertia[3 * s + 1] = nneg;
    fd.inertia[3 * s + 2] = nzero;
    fd.origMaxBits[s] = doubleToBits(origMax);
    fd.runMaxBits[s] = doubleToBits(rm);
    fd.pert[s] = pert;
    fd.npert[s] = npert;
    fd.stopped[s] = (k < nE) ? 1 : 0;
    fd.kcur[s] = k;
  }
}

size_t smallFrontSmemBytes(int m) {
  return sizeof(double) * (64 + static_cast<size_t>(m) * (m + 1) / 2) + sizeof(int) * (m + 2);
}

// ---------------------------------------------------------------------------
// Large fronts.
// ---------------------------------------------------------------------------

// max|F| over the lower triangle (dense_kernel.hpp:413-415); `bpf` blocks
// per listed front.
template <typename T>
__global__ __launch_bounds__(256) void k_front_maxabs(const int* list, int count, int bpf, DevFronts fd) {
  const int f = blockIdx.x / bpf;
  if (f >= count) return;
  const int s = list[f];
  const int part = blockIdx.x % bpf;
  const int m = fd.m[s];
  const T* F = static_cast<const T*>(fd.F[s]);
  double v = 0.0;
  const int lane = threadIdx.x & 63;
  const int gw = part * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nwTot = bpf * (blockDim.x >> 6);
  for (int j = gw; j < m; j += nwTot) {
    const T* col = F + static_cast<size_t>(j) * m;
    for (int i = j + lane; i < m; i += 64) v = fmax(v, fabs(static_cast<double>(col[i])));
  }
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if (lane == 0) atomicMaxNonNeg(fd.origMaxBits + s, v);
}

// Panel pivoting on the eligible rows [k, nE) of one front.  Columns not yet
// pivoted are never modified in place in global memory: the current value of
// a candidate column is its stored (pre-panel) value minus the correction of
// this panel's pivots, like the reference's setColumnRaw
// (dense_kernel.hpp:516-545).  Accepted pivots write their multipliers for
// the eligible rows; the remaining rows are solved by k_panel_rows.
//
// When every eligible row fits RM slots per thread (RM > 0), the next SW
// columns from the current pivot (a sub-panel) are loaded into LDS with the
// correction of the panel's earlier pivots applied once, then factored
// right-looking inside LDS: a pivot accepted by the first Bunch-Kaufman test
// (every static pivot) costs one barrier and no global reads.  Any other
// decision runs the general code, which works on column buffers in global
// memory, and the sub-panel is reloaded after it.
template <int RM>
struct PanelShape {
  // sub-panel width (a wider one saves reloads but costs LDS traffic on every pivot)
  static constexpr int SW = RM <= 4 ? 8 : (RM <= 8 ? 4 : (RM <= 16 ? 2 : 1));
  static constexpr int ROWS = RM > 0 ? RM * kPanelThreads : 0;
  static constexpr size_t lds() { return sizeof(double) * static_cast<size_t>(ROWS) * SW; }
};

template <int RM, typename T>
__global__ __launch_bounds__(kPanelThreads) void k_panel(const int* list, int count, DevFronts fd, FactorParams prm,
                                                          double* wsBase, const int64_t* wsOff) {
  constexpr int NT = kPanelThreads;
  constexpr int SW = PanelShape<RM>::SW;
  constexpr int CSR = PanelShape<RM>::ROWS;
  extern __shared__ double Cs[];  // [SW][CSR]: sub-panel column j, row k0 + r at Cs[j * CSR + r]
  __shared__ double Wsub[SW][kPanel + 2];
  __shared__ double wvec[kPanel + 2];
  __shared__ double pdd[kPanel + 2], pdo[kPanel + 2];
  __shared__ int pkd[kPanel + 2];
  __shared__ double redV[NT / 64 + 1];
  __shared__ int redI[NT / 64 + 1];
  __shared__ double fastV[2][NT / 64];
  __shared__ int fastI[2][NT / 64];
  __shared__ double blkV[NT / 64][SW];  // per-wave column maxima of a blocked step

  const int tid = threadIdx.x;
  if (static_cast<int>(blockIdx.x) >= count) return;
  const int s = list[blockIdx.x];
  double* ws = wsBase + wsOff[s];
  if (fd.stopped[s]) {
    if (tid == 0) fd.pcount[s] = 0;
    return;
  }
  const int m = fd.m[s];
  const int nE = fd.nE[s];
  T* F = static_cast<T*>(fd.F[s]);
  const int64_t io = fd.idxOff[s];
  // after a speculative step the panel started p0 pivots before kcur
  const int kc = fd.kcur[s];
  const bool sp = specFront(prm, nE);
  const int p0 = sp ? fd.pcount[s] : 0;
  const int k0 = kc - p0;
  if (kc >= nE) {
    if (tid == 0 && !sp) fd.pcount[s] = 0;
    return;
  }
  const int nbEnd = min(k0 + kPanel, nE);
  if (kc >= nbEnd) return;
  for (int q = tid; q < p0; q += NT) {
    pdd[q] = fd.dDiag[io + k0 + q];
    pdo[q] = fd.dOff[io + k0 + q];
    pkd[q] = fd.pkind[io + k0 + q];
  }
  __syncthreads();
  double* curK = ws;
  double* curR = ws + nE;
  auto Fat = [&](int i, int j) -> T& { return F[i + static_cast<size_t>(j) * m]; };

  const bool stat = prm.staticMode != 0;
  double origMax = bitsToDouble(fd.origMaxBits[s]);
  if (!stat && origMax < prm.zeroTol) origMax = 1.0;
  if (stat && origMax < prm.staticAbsFloor) origMax = 1.0;
  const double runMax0 = fmax(origMax, bitsToDouble(fd.runMaxBits[s]));
  double tmax = 0.0;
  const double delta = prm.staticDelta > 0.0 ? prm.staticDelta : sqrt(DBL_EPSILON) * origMax;
  const double sfloor = fmax(prm.staticAbsFloor, prm.staticRelFloor * origMax);

  int npos = 0, nneg = 0, nzero = 0, npert = 0;
  double pert = 0.0;
  int k = kc, p = p0;
  bool stop = false;

  // (D L^T)(q, row) for this panel's pivot q and a given row
  auto weight = [&](int q, int row) -> double {
    const int kind = pkd[q];
    if (kind == 1) return pdd[q] * static_cast<double>(Fat(row, k0 + q));
    if (kind == 2) return pdd[q] * static_cast<double>(Fat(row, k0 + q)) + pdo[q] * static_cast<double>(Fat(row, k0 + q + 1));
    return pdo[q - 1] * static_cast<double>(Fat(row, k0 + q - 1)) + pdd[q] * static_cast<double>(Fat(row, k0 + q));
  };
  // weights of this panel's p pivots for one row (general path)
  auto computeW = [&](int row) {
    for (int q = tid; q < p; q += NT) wvec[q] = weight(q, row);
    __syncthreads();
  };
  // out[i] = stored column `col` (rows >= k) minus the panel correction
  auto currentColumn = [&](int col, double* out) {
    for (int i = k + tid; i < nE; i += NT) {
      const double raw = (i >= col) ? Fat(i, col) : Fat(col, i);
      double acc = 0.0;
      for (int q = 0; q < p; ++q) acc = fma(static_cast<double>(Fat(i, k0 + q)), wvec[q], acc);
      out[i] = raw - acc;
    }
    __syncthreads();
  };
  // symmetric interchange of positions a < b in storage (all rows, incl. the
  // factored part and the not-fully-summed rows) and in the column buffers
  auto swapSym = [&](int a, int b) {
    for (int t = tid; t < m; t += NT) {
      if (t < a) {
        const T x = Fat(a, t);
        Fat(a, t) = Fat(b, t);
        Fat(b, t) = x;
      } else if (t == a) {
        const T x = Fat(a, a);
        Fat(a, a) = Fat(b, b);
        Fat(b, b) = x;
      } else if (t < b) {
        const T x = Fat(t, a);
        Fat(t, a) = Fat(b, t);
        Fat(b, t) = x;
      } else if (t > b) {
        const T x = Fat(t, a);
        Fat(t, a) = Fat(t, b);
        Fat(t, b) = x;
      }
    }
    if (tid == 0) {
      const int tp = fd.perm[io + a];
      fd.perm[io + a] = fd.perm[io + b];
      fd.perm[io + b] = tp;
      double x = curK[a];
      curK[a] = curK[b];
      curK[b] = x;
      x = curR[a];
      curR[a] = curR[b];
      curR[b] = x;
    }
    __syncthreads();
  };
  auto accept1 = [&](const double* col, double d) {
    if (tid == 0) {
      fd.dDiag[io + k] = d;
      fd.dOff[io + k] = 0.0;
      fd.pkind[io + k] = 1;
      pdd[p] = d;
      pdo[p] = 0.0;
      pkd[p] = 1;
      Fat(k, k) = static_cast<T>(d);
    }
    countInertia(d, npos, nneg, nzero);
    tmax = fmax(tmax, fabs(d));
    for (int i = k + 1 + tid; i < nE; i += NT) {
      const double c = col[i];
      tmax = fmax(tmax, fabs(c));
      Fat(i, k) = static_cast<T>(c / d);
    }
    __syncthreads();
    ++k;
    ++p;
  };

  // Sub-panel [c0, c0 + sw): current values of the eligible rows >= k.
  int c0 = k, sw = 0;
  auto loadSub = [&]() {
    c0 = k;
    sw = min(SW, nbEnd - k);
    for (int e = tid; e < sw * p; e += NT) {
      const int j = e % sw, q = e / sw;
      Wsub[j][q] = weight(q, c0 + j);
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < (RM > 0 ? RM : 1); ++jj) {
      const int i = k0 + tid + jj * NT;
      if (i < k || i >= nE) continue;
      double acc[SW];
      loadWiden<SW>(acc, static_cast<const T*>(F), [&](int j) -> int64_t {
        const int col = c0 + j;
        if (j >= sw) return -1;
        return i >= col ? i + static_cast<int64_t>(col) * m : col + static_cast<int64_t>(i) * m;
      });
      constexpr int U = 8;
      int q = 0;
      for (; q + U <= p; q += U) {
        double lv[U];
        loadWiden<U>(lv, static_cast<const T*>(F), [&](int u) -> int64_t { return i + static_cast<int64_t>(k0 + q + u) * m; });
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < SW; ++j) acc[j] = fma(-lv[u], Wsub[j][q + u], acc[j]);
      }
      for (; q < p; ++q) {
        const double l = Fat(i, k0 + q);
#pragma unroll
        for (int j = 0; j < SW; ++j) acc[j] = fma(-l, Wsub[j][q], acc[j]);
      }
#pragma unroll
      for (int j = 0; j < SW; ++j)
        if (j < sw) Cs[j * CSR + (i - k0)] = acc[j];
    }
    __syncthreads();
  };

  int parity = 0;
  if (RM > 0) loadSub();

  // Blocked first test over the sub-panel's remaining columns [k, c0 + sw):
  // every thread factors that block (at most SW x SW, from LDS) with 1x1
  // pivots in order and runs the same recurrence over its own rows,
  // c_ij = a_ij - sum_{q<j} l_iq d_q l_(k+j)q, so one reduction gives
  // lambda for every column at once.  The pivots before the first column
  // whose test fails are kept (the same decisions as pivot by pivot, up to
  // the order of rounding); that column goes through the single-pivot path
  // below.  Returns the number of pivots kept.
  auto blockStep = [&]() -> int {
    const int kk0 = k - c0, bw = sw - kk0;
    double Wb[SW][SW], dq[SW], aq[SW];
    {
      double B[SW][SW];
#pragma unroll
      for (int a = 0; a < SW; ++a)
#pragma unroll
        for (int b = 0; b <= a; ++b) B[a][b] = (a < bw) ? Cs[(kk0 + b) * CSR + (k + a - k0)] : 0.0;
#pragma unroll
      for (int q = 0; q < SW; ++q) {
        dq[q] = 1.0;
        aq[q] = 0.0;
        if (q < bw) {
          aq[q] = B[q][q];
          double d = aq[q];
          if (stat) {
            const int sign = prm.expectedSign ? prm.expectedSign[fd.frontIdx[io + k + q]] : 0;
            double applied;
            d = staticPivot(d, sign, delta, sfloor, applied);
          }
          dq[q] = d;
          double la[SW];
#pragma unroll
          for (int a = q + 1; a < SW; ++a) {
            la[a] = (a < bw) ? static_cast<double>(static_cast<T>(B[a][q] / d)) : 0.0;
            Wb[a][q] = d * la[a];
          }
#pragma unroll
          for (int a = q + 1; a < SW; ++a)
#pragma unroll
            for (int b = q + 1; b <= a; ++b) B[a][b] = fma(-la[a], Wb[b][q], B[a][b]);
        }
      }
    }
    // pass 1: column maxima over this thread's rows below each pivot
    double vmax[SW];
#pragma unroll
    for (int j = 0; j < SW; ++j) vmax[j] = 0.0;
#pragma unroll
    for (int jj = 0; jj < (RM > 0 ? RM : 1); ++jj) {
      const int i = k0 + tid + jj * NT;
      if (i < k || i >= nE) continue;
      double l[SW];
#pragma unroll
      for (int j = 0; j < SW; ++j) {
        if (j < bw) {
          double c = Cs[(kk0 + j) * CSR + (i - k0)];
#pragma unroll
          for (int q = 0; q < j; ++q) c = fma(-l[q], Wb[j][q], c);
          if (i > k + j) vmax[j] = fmax(vmax[j], fabs(c));
          l[j] = static_cast<double>(static_cast<T>(c / dq[j]));
        }
      }
    }
#pragma unroll
    for (int j = 0; j < SW; ++j) {
      const double v = __ockl_wfred_max_f64(vmax[j]);
      if ((tid & 63) == 0) blkV[tid >> 6][j] = v;
    }
    __syncthreads();
    int qs = bw;
    {
      const double isolatedFloor = fmax(prm.zeroTol, prm.relFloor * origMax);
      for (int j = 0; j < bw; ++j) {
        double lam = 0.0;
        for (int w = 0; w < NT / 64; ++w) lam = fmax(lam, blkV[w][j]);
        const bool ok = stat || ((lam <= prm.zeroTol) ? fabs(aq[j]) > isolatedFloor : fabs(aq[j]) >= kAlpha * lam);
        if (!ok) {
          qs = j;
          break;
        }
      }
    }
    if (qs == 0) return 0;  // blkV is rewritten only after the next block step's barrier
    // pass 2: multipliers of the kept columns, current values of the others
#pragma unroll
    for (int jj = 0; jj < (RM > 0 ? RM : 1); ++jj) {
      const int i = k0 + tid + jj * NT;
      if (i < k || i >= nE) continue;
      double x[SW], l[SW];
#pragma unroll
      for (int j = 0; j < SW; ++j) x[j] = (j < bw) ? Cs[(kk0 + j) * CSR + (i - k0)] : 0.0;
#pragma unroll
      for (int j = 0; j < SW; ++j) {
        if (j < qs) {
          double c = x[j];
#pragma unroll
          for (int q = 0; q < j; ++q) c = fma(-l[q], Wb[j][q], c);
          l[j] = static_cast<double>(static_cast<T>(c / dq[j]));
          if (i > k + j) {
            tmax = fmax(tmax, fabs(c));
            Fat(i, k + j) = static_cast<T>(l[j]);
          }
        } else if (j < bw) {
          double y = x[j];
#pragma unroll
          for (int q = 0; q < SW; ++q)
            if (q < qs) y = fma(-l[q], Wb[j][q], y);
          Cs[(kk0 + j) * CSR + (i - k0)] = y;
        }
      }
    }
    for (int q = 0; q < qs; ++q) {
      pdd[p + q] = dq[q];
      pdo[p + q] = 0.0;
      pkd[p + q] = 1;
      countInertia(dq[q], npos, nneg, nzero);
      tmax = fmax(tmax, fabs(dq[q]));
      if (stat && dq[q] != aq[q]) {
        pert += fabs(dq[q] - aq[q]);
        ++npert;
      }
    }
    if (tid == 0)
      for (int q = 0; q < qs; ++q) {
        fd.dDiag[io + k + q] = dq[q];
        fd.dOff[io + k + q] = 0.0;
        fd.pkind[io + k + q] = 1;
        Fat(k + q, k + q) = static_cast<T>(dq[q]);
      }
    __syncthreads();  // the kept rows' Cs values are read by other threads next
    return qs;
  };
  int skipAt = -1;

  while (k < nbEnd) {
    if (RM > 0 && SW > 1 && k != skipAt && k - c0 + 1 < sw) {
      const int kept = blockStep();
      if (kept == 0) {
        skipAt = k;
      } else {
        k += kept;
        p += kept;
        if (k - c0 >= sw && k < nbEnd) {
          __syncthreads();
          loadSub();
        }
      }
      continue;
    }
    double lv = 0.0;
    int li = -1;
    double akk;
    const int kk = k - c0;
    if (RM > 0) {
      const double* col = Cs + kk * CSR;
#pragma unroll
      for (int jj = 0; jj < (RM > 0 ? RM : 1); ++jj) {
        const int i = k0 + tid + jj * NT;
        if (i > k && i < nE) {
          const double v = fabs(col[i - k0]);
          if (v > 0.0) argmaxCombine(lv, li, v, i);
        }
      }
      waveArgmax(lv, li);
      if ((tid & 63) == 0) {
        fastV[parity][tid >> 6] = lv;
        fastI[parity][tid >> 6] = li;
      }
      __syncthreads();
      lv = 0.0;
      li = -1;
      for (int w = 0; w < NT / 64; ++w) argmaxCombine(lv, li, fastV[parity][w], fastI[parity][w]);
      parity ^= 1;
      akk = col[k - k0];
    } else {
      computeW(k);
      currentColumn(k, curK);
      if (!stat) {
        for (int i = k + 1 + tid; i < nE; i += NT) {
          const double v = fabs(curK[i]);
          if (v > 0.0) argmaxCombine(lv, li, v, i);
        }
        blockArgmax<NT>(lv, li, redV, redI);
      }
      akk = curK[k];
    }

    // first Bunch-Kaufman tests (dense_kernel.hpp:601-618); static pivots
    // are always accepted at k (dense_kernel.hpp:982-1011)
    int action = -1;  // -1: sigma test needed
    double dpiv = akk;
    if (stat) {
      const int sign = prm.expectedSign ? prm.expectedSign[fd.frontIdx[io + k]] : 0;
      double applied;
      dpiv = staticPivot(akk, sign, delta, sfloor, applied);
      if (applied != 0.0) {
        pert += fabs(applied);
        ++npert;
      }
      action = 1;
    } else if (lv <= prm.zeroTol) {
      const double isolatedFloor = fmax(prm.zeroTol, prm.relFloor * origMax);
      action = (fabs(akk) <= isolatedFloor) ? 0 : 1;
    } else if (fabs(akk) >= kAlpha * lv) {
      action = 1;
    }
    if (action == 0) {
      stop = true;
      break;
    }
    if (action == 1 && RM > 0) {
      const double d = dpiv;
      const double* col = Cs + kk * CSR;
      // (D L^T)(k, c0 + j) for the sub-panel columns right of k, as the
      // general path would form it from the stored multiplier
      double wj[SW];
#pragma unroll
      for (int j = 0; j < SW; ++j)
        wj[j] = (j > kk && j < sw) ? d * static_cast<double>(static_cast<T>(col[c0 + j - k0] / d)) : 0.0;
      pdd[p] = d;  // every thread stores the same value
      pdo[p] = 0.0;
      pkd[p] = 1;
      if (tid == 0) {
        fd.dDiag[io + k] = d;
        fd.dOff[io + k] = 0.0;
        fd.pkind[io + k] = 1;
        Fat(k, k) = static_cast<T>(d);
      }
      countInertia(d, npos, nneg, nzero);
      tmax = fmax(tmax, fabs(d));
#pragma unroll
      for (int jj = 0; jj < (RM > 0 ? RM : 1); ++jj) {
        const int i = k0 + tid + jj * NT;
        if (i > k && i < nE) {
          const double c = col[i - k0];
          tmax = fmax(tmax, fabs(c));
          const T lt = static_cast<T>(c / d);
          Fat(i, k) = lt;
          const double l = lt;
#pragma unroll
          for (int j = 0; j < SW; ++j)
            if (j > kk && j < sw) Cs[j * CSR + (i - k0)] = fma(-l, wj[j], Cs[j * CSR + (i - k0)]);
        }
      }
      // No barrier here: a thread updated only its own rows, and the values
      // the next step reads from other threads' rows (a_kk, the sub-panel
      // rows of the pivot column) are read after the next reduction's
      // barrier.  A reload reads the multipliers just stored: it waits.
      ++k;
      ++p;
      if (k - c0 >= sw && k < nbEnd) {
        __syncthreads();
        loadSub();
      }
      continue;
    }
    if (RM > 0) {
      // the general path below works on the column buffers
      const double* col = Cs + kk * CSR;
      for (int i = k + tid; i < nE; i += NT) curK[i] = col[i - k0];
      __syncthreads();
    }
    if (action == 1) {
      accept1(curK, dpiv);
      if (RM > 0 && k < nbEnd) loadSub();
      continue;
    }
    const int r = li;
    computeW(r);
    currentColumn(r, curR);
    // the reference mirrors column r into row r, so column k's entry at
    // row r becomes column r's value at row k (dense_kernel.hpp:543)
    if (tid == 0) curK[r] = curR[k];
    __syncthreads();
    double sg = 0.0;
    for (int i = k + tid; i < nE; i += NT)
      if (i != r) sg = fmax(sg, fabs(curR[i]));
    sg = blockMax<NT>(sg, redV);
    if (fabs(akk) * sg >= kAlpha * lv * lv)
      action = 1;
    else if (fabs(curR[r]) >= kAlpha * sg)
      action = 2;
    else
      action = 3;
    if (action == 1) {
      accept1(curK, curK[k]);
      if (RM > 0 && k < nbEnd) loadSub();
      continue;
    }
    if (action == 2) {
      swapSym(k, r);
      accept1(curR, curR[k]);
      if (RM > 0 && k < nbEnd) loadSub();
      continue;
    }
    // 2x2 at (k, r): bring r to k+1
    if (r != k + 1) swapSym(k + 1, r);
    const double d11 = curK[k];
    const double d21 = curK[k + 1];
    const double d22 = curR[k + 1];
    const double det = d11 * d22 - d21 * d21;
    const double rmax = fmax(runMax0, blockMax<NT>(tmax, redV));
    if (fabs(det) <= prm.zeroTol * fmax(1.0, rmax) * fmax(1.0, rmax)) {
      stop = true;
      break;
    }
    tmax = fmax(tmax, fmax(fabs(d11), fmax(fabs(d21), fabs(d22))));
    if (tid == 0) {
      fd.dDiag[io + k] = d11;
      fd.dOff[io + k] = d21;
      fd.dDiag[io + k + 1] = d22;
      fd.dOff[io + k + 1] = 0.0;
      fd.pkind[io + k] = 2;
      fd.pkind[io + k + 1] = 0;
      pdd[p] = d11;
      pdo[p] = d21;
      pkd[p] = 2;
      pdd[p + 1] = d22;
      pdo[p + 1] = 0.0;
      pkd[p + 1] = 0;
      Fat(k, k) = static_cast<T>(d11);
      Fat(k + 1, k) = static_cast<T>(d21);
      Fat(k + 1, k + 1) = static_cast<T>(d22);
    }
    countInertia2(d11, d22, det, npos, nneg);
    const double inv11 = d22 / det, inv21 = -d21 / det, inv22 = d11 / det;
    for (int i = k + 2 + tid; i < nE; i += NT) {
      const double a0 = curK[i], a1 = curR[i];
      tmax = fmax(tmax, fmax(fabs(a0), fabs(a1)));
      Fat(i, k) = static_cast<T>(a0 * inv11 + a1 * inv21);
      Fat(i, k + 1) = static_cast<T>(a0 * inv21 + a1 * inv22);
    }
    __syncthreads();
    k += 2;
    p += 2;
    if (RM > 0 && k < nbEnd) loadSub();
  }
  __syncthreads();

  const double bm = blockMax<NT>(tmax, redV);
  if (tid == 0) {
    fd.kcur[s] = k;
    fd.pcount[s] = p;
    fd.stopped[s] = stop ? 1 : 0;
    fd.inertia[3 * s] += npos;
    fd.inertia[3 * s + 1] += nneg;
    fd.inertia[3 * s + 2] += nzero;
    fd.pert[s] += pert;
    fd.npert[s] += npert;
    atomicMaxNonNeg(fd.runMaxBits + s, fmax(bm, origMax));
  }
}

// Device-coherent accesses for values exchanged between workgroups of one
// launch (no cache write-back or invalidation needed around them).
template <typename T>
__device__ inline void storeCoherent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline T loadCoherent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cooperative panel of a front with more eligible rows than one workgroup
// holds (the same decisions as k_panel, dense_kernel.hpp:585-792).  The W
// workgroups of a front own kMwRows consecutive eligible rows each; the
// sub-panel (SW columns from the current pivot) of those rows lives in each
// workgroup's LDS.  Per pivot every workgroup posts its local max |a_ik|
// (workgroup 0 also a_kk and the pivot column's sub-panel rows) and reads
// everyone's, so all reach the same decision.  A decision other than "1x1
// at k" is carried out by workgroup 0 alone on the column buffers in global
// memory (the general path of k_panel) while the others wait; then all
// reload their sub-panel.  Waits are bounded (the error word is raised and
// the wait abandoned), so every workgroup runs to its end.
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void k_panel_mw(const int* list, int count, const int* wgStart,
                                                             DevFronts fd, FactorParams prm, double* wsBase,
                                                             const int64_t* wsOff, MwSync* syncBase, int* err) {
  constexpr int NT = kPanelThreads;
  constexpr int SW = 8;
  constexpr int CSR = kMwRows;
  extern __shared__ double Cs[];  // [SW][CSR]
  __shared__ double Wsub[SW][kPanel + 2];
  __shared__ double wvec[kPanel + 2];
  __shared__ double pdd[kPanel + 2], pdo[kPanel + 2];
  __shared__ int pkd[kPanel + 2];
  __shared__ double redV[NT / 64 + 1];
  __shared__ int redI[NT / 64 + 1];
  __shared__ double bcast[16];
  __shared__ int bcastI[4];
  __shared__ double lastL[SW + 1];  // multipliers of the latest fast pivot at rows k + 1 .. k + SW
  __shared__ int abandon;

  const int tid = threadIdx.x;
  if (static_cast<int>(blockIdx.x) >= wgStart[count]) return;
  const int f = findSegment(wgStart, count, blockIdx.x);
  const int s = list[f];
  const int w = blockIdx.x - wgStart[f];
  const int W = wgStart[f + 1] - wgStart[f];
  MwSync* sy = syncBase + f;
  const unsigned long long base = static_cast<unsigned long long>(loadCoherent(&sy->epoch)) << 16;
  const bool lead = (w == 0);
  double* ws = wsBase + wsOff[s];
  const int m = fd.m[s];
  const int nE = fd.nE[s];
  T* F = static_cast<T*>(fd.F[s]);
  const int64_t io = fd.idxOff[s];
  // after a speculative step the panel started p0 pivots before kcur
  const int kc = fd.kcur[s];
  const bool sp = specFront(prm, nE);
  const int p0 = sp ? fd.pcount[s] : 0;
  const int k0 = kc - p0;
  const int nbEnd = min(k0 + kPanel, nE);
  const bool skip = fd.stopped[s] != 0 || kc >= nbEnd;
  if (skip || nE - k0 > W * kMwRows) {
    if (!skip && tid == 0) atomicExch(err, 2);  // the launch does not cover the rows (never planned so)
    if (lead && tid == 0 && !sp) fd.pcount[s] = 0;
    if (lead && tid == 0) storeCoherent(&sy->epoch, static_cast<unsigned>((base >> 16) + 1));
    return;
  }
  const int r0 = k0 + w * kMwRows;  // first row of this workgroup
  double* curK = ws;
  double* curR = ws + nE;
  auto Fat = [&](int i, int j) -> T& { return F[i + static_cast<size_t>(j) * m]; };

  const bool stat = prm.staticMode != 0;
  double origMax = bitsToDouble(fd.origMaxBits[s]);
  if (!stat && origMax < prm.zeroTol) origMax = 1.0;
  if (stat && origMax < prm.staticAbsFloor) origMax = 1.0;
  const double runMax0 = fmax(origMax, bitsToDouble(fd.runMaxBits[s]));
  double tmax = 0.0;
  const double delta = prm.staticDelta > 0.0 ? prm.staticDelta : sqrt(DBL_EPSILON) * origMax;
  const double sfloor = fmax(prm.staticAbsFloor, prm.staticRelFloor * origMax);
  int npos = 0, nneg = 0, nzero = 0, npert = 0;
  double pert = 0.0;
  int k = kc, p = p0;
  bool stop = false;
  unsigned step = 0;
  if (tid == 0) abandon = 0;
  for (int q = tid; q < p0; q += NT) {
    pdd[q] = fd.dDiag[io + k0 + q];
    pdo[q] = fd.dOff[io + k0 + q];
    pkd[q] = fd.pkind[io + k0 + q];
  }
  __syncthreads();

  // post this workgroup's step, then wait until every workgroup posted it
  auto post = [&](unsigned st) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) storeCoherent(&sy->flag[w], base | st);
  };
  auto waitFor = [&](int from, int to, unsigned st) {
    if (tid >= from && tid < to) {
      unsigned spins = 0;
      while (loadCoherent(&sy->flag[tid]) < (base | st)) {
        if (loadCoherent(err) != 0) {
          abandon = 1;
          break;
        }
        if (++spins > (1u << 22)) {
          atomicExch(err, 3);
          abandon = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
  };

  // (D L^T)(q, row) of this panel's pivot q for a row
  auto weightCoherent = [&](int q, int row) -> double {
    auto L = [&](int c) { return static_cast<double>(loadCoherent(&Fat(row, k0 + c))); };
    const int kind = pkd[q];
    if (kind == 1) return pdd[q] * L(q);
    if (kind == 2) return pdd[q] * L(q) + pdo[q] * L(q + 1);
    return pdo[q - 1] * L(q - 1) + pdd[q] * L(q);
  };
  // lastFast: the pivot just taken (p - 1, at row c0 - 1) was a fast 1x1
  // whose multipliers workgroup 0 may still be storing, so its weights come
  // from lastL (the same values, formed from the broadcast pivot column);
  // every earlier pivot's multipliers were stored before workgroup 0 posted
  // the step that followed it.
  int c0 = k, sw = 0;
  auto loadSub = [&](bool lastFast) {
    c0 = k;
    sw = min(SW, nbEnd - k);
    for (int e = tid; e < sw * p; e += NT) {
      const int j = e % sw, q = e / sw;
      Wsub[j][q] = (lastFast && q == p - 1) ? pdd[q] * lastL[1 + j] : weightCoherent(q, c0 + j);
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int i = r0 + tid + jj * NT;
      if (i < k || i >= nE) continue;
      double acc[SW];
      loadWiden<SW>(acc, static_cast<const T*>(F), [&](int j) -> int64_t {
        const int col = c0 + j;
        if (j >= sw) return -1;
        return i >= col ? i + static_cast<int64_t>(col) * m : col + static_cast<int64_t>(i) * m;
      });
      // the panel's own rows were stored device-coherent: read them the same way
      const bool panelRow = i < k0 + kPanel + 2;
      constexpr int U = 8;
      int q = 0;
      for (; q + U <= p; q += U) {
        double lv[U];
        if (panelRow) {
          T raw[U];
#pragma unroll
          for (int u = 0; u < U; ++u) raw[u] = loadCoherent(&Fat(i, k0 + q + u));
#pragma unroll
          for (int u = 0; u < U; ++u) lv[u] = static_cast<double>(raw[u]);
        } else {
          loadWiden<U>(lv, static_cast<const T*>(F), [&](int u) -> int64_t { return i + static_cast<int64_t>(k0 + q + u) * m; });
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < SW; ++j) acc[j] = fma(-lv[u], Wsub[j][q + u], acc[j]);
      }
      for (; q < p; ++q) {
        const double l = panelRow ? loadCoherent(&Fat(i, k0 + q)) : Fat(i, k0 + q);
#pragma unroll
        for (int j = 0; j < SW; ++j) acc[j] = fma(-l, Wsub[j][q], acc[j]);
      }
#pragma unroll
      for (int j = 0; j < SW; ++j)
        if (j < sw) Cs[j * CSR + (i - r0)] = acc[j];
    }
    __syncthreads();
  };

  // ---- general step (workgroup 0 alone; column buffers in global memory) ----
  auto computeW = [&](int row) {
    for (int q = tid; q < p; q += NT) {
      const int kind = pkd[q];
      double wq;
      if (kind == 1)
        wq = pdd[q] * static_cast<double>(Fat(row, k0 + q));
      else if (kind == 2)
        wq = pdd[q] * static_cast<double>(Fat(row, k0 + q)) + pdo[q] * static_cast<double>(Fat(row, k0 + q + 1));
      else
        wq = pdo[q - 1] * static_cast<double>(Fat(row, k0 + q - 1)) + pdd[q] * static_cast<double>(Fat(row, k0 + q));
      wvec[q] = wq;
    }
    __syncthreads();
  };
  auto currentColumn = [&](int col, double* out) {
    for (int i = k + tid; i < nE; i += NT) {
      const double raw = (i >= col) ? Fat(i, col) : Fat(col, i);
      double acc = 0.0;
      for (int q = 0; q < p; ++q) acc = fma(static_cast<double>(Fat(i, k0 + q)), wvec[q], acc);
      out[i] = raw - acc;
    }
    __syncthreads();
  };
  auto swapSym = [&](int a, int b) {
    for (int t = tid; t < m; t += NT) {
      if (t < a) {
        const T x = Fat(a, t);
        Fat(a, t) = Fat(b, t);
        Fat(b, t) = x;
      } else if (t == a) {
        const T x = Fat(a, a);
        Fat(a, a) = Fat(b, b);
        Fat(b, b) = x;
      } else if (t < b) {
        const T x = Fat(t, a);
        Fat(t, a) = Fat(b, t);
        Fat(b, t) = x;
      } else if (t > b) {
        const T x = Fat(t, a);
        Fat(t, a) = Fat(t, b);
        Fat(t, b) = x;
      }
    }
    if (tid == 0) {
      const int tp = fd.perm[io + a];
      fd.perm[io + a] = fd.perm[io + b];
      fd.perm[io + b] = tp;
      double x = curK[a];
      curK[a] = curK[b];
      curK[b] = x;
      x = curR[a];
      curR[a] = curR[b];
      curR[b] = x;
    }
    __syncthreads();
  };
  auto accept1 = [&](const double* col, double d) {
    if (tid == 0) {
      fd.dDiag[io + k] = d;
      fd.dOff[io + k] = 0.0;
      fd.pkind[io + k] = 1;
      pdd[p] = d;
      pdo[p] = 0.0;
      pkd[p] = 1;
      Fat(k, k) = static_cast<T>(d);
    }
    countInertia(d, npos, nneg, nzero);
    tmax = fmax(tmax, fabs(d));
    for (int i = k + 1 + tid; i < nE; i += NT) {
      const double c = col[i];
      tmax = fmax(tmax, fabs(c));
      Fat(i, k) = static_cast<T>(c / d);
    }
    __syncthreads();
    ++k;
    ++p;
  };
  // one general step at pivot k (sigma test, swaps, 1x1 or 2x2, or a stop);
  // curK holds the current column k for the rows [k, nE)
  auto generalStep = [&](double lv, int li, double akk) {
    const int r = li;
    computeW(r);
    currentColumn(r, curR);
    if (tid == 0) curK[r] = curR[k];  // dense_kernel.hpp:543
    __syncthreads();
    double sg = 0.0;
    for (int i = k + tid; i < nE; i += NT)
      if (i != r) sg = fmax(sg, fabs(curR[i]));
    sg = blockMax<NT>(sg, redV);
    int action;
    if (fabs(akk) * sg >= kAlpha * lv * lv)
      action = 1;
    else if (fabs(curR[r]) >= kAlpha * sg)
      action = 2;
    else
      action = 3;
    if (action == 1) {
      accept1(curK, curK[k]);
      return;
    }
    if (action == 2) {
      swapSym(k, r);
      accept1(curR, curR[k]);
      return;
    }
    if (r != k + 1) swapSym(k + 1, r);
    const double d11 = curK[k];
    const double d21 = curK[k + 1];
    const double d22 = curR[k + 1];
    const double det = d11 * d22 - d21 * d21;
    const double rmax = fmax(runMax0, blockMax<NT>(tmax, redV));
    if (fabs(det) <= prm.zeroTol * fmax(1.0, rmax) * fmax(1.0, rmax)) {
      stop = true;
      return;
    }
    tmax = fmax(tmax, fmax(fabs(d11), fmax(fabs(d21), fabs(d22))));
    if (tid == 0) {
      fd.dDiag[io + k] = d11;
      fd.dOff[io + k] = d21;
      fd.dDiag[io + k + 1] = d22;
      fd.dOff[io + k + 1] = 0.0;
      fd.pkind[io + k] = 2;
      fd.pkind[io + k + 1] = 0;
      pdd[p] = d11;
      pdo[p] = d21;
      pkd[p] = 2;
      pdd[p + 1] = d22;
      pdo[p + 1] = 0.0;
      pkd[p + 1] = 0;
      Fat(k, k) = static_cast<T>(d11);
      Fat(k + 1, k) = static_cast<T>(d21);
      Fat(k + 1, k + 1) = static_cast<T>(d22);
    }
    countInertia2(d11, d22, det, npos, nneg);
    const double inv11 = d22 / det, inv21 = -d21 / det, inv22 = d11 / det;
    for (int i = k + 2 + tid; i < nE; i += NT) {
      const double a0 = curK[i], a1 = curR[i];
      tmax = fmax(tmax, fmax(fabs(a0), fabs(a1)));
      Fat(i, k) = static_cast<T>(a0 * inv11 + a1 * inv21);
      Fat(i, k + 1) = static_cast<T>(a0 * inv21 + a1 * inv22);
    }
    __syncthreads();
    k += 2;
    p += 2;
  };

  loadSub(false);
  while (k < nbEnd && !abandon) {
    const int kk = k - c0;
    const double* col = Cs + kk * CSR;
    // local candidates of column k
    double lv = 0.0;
    int li = -1;
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int i = r0 + tid + jj * NT;
      if (i > k && i < nE) {
        const double v = fabs(col[i - r0]);
        if (v > 0.0) argmaxCombine(lv, li, v, i);
      }
    }
    blockArgmax<NT>(lv, li, redV, redI);
    const int par = step & 1;
    ++step;
    if (tid == 0) {
      storeCoherent(&sy->lv[par][w], lv);
      storeCoherent(&sy->li[par][w], li);
      if (lead) {
        // a_kk and the pivot column's next SW rows (rows k + t < nbEnd)
        storeCoherent(&sy->extra[par][0], col[k - r0]);
        for (int t = 1; t <= SW && k + t < nbEnd; ++t) storeCoherent(&sy->extra[par][t], col[k + t - r0]);
      }
    }
    post(step);
    waitFor(0, W, step);
    if (abandon) break;
    if (tid == 0) {
      double bv = 0.0;
      int bi = -1;
      for (int v = 0; v < W; ++v) argmaxCombine(bv, bi, loadCoherent(&sy->lv[par][v]), loadCoherent(&sy->li[par][v]));
      bcast[0] = bv;
      bcastI[0] = bi;
      bcast[1] = loadCoherent(&sy->extra[par][0]);
      for (int t = 1; t <= SW && k + t < nbEnd; ++t) bcast[1 + t] = loadCoherent(&sy->extra[par][t]);
    }
    __syncthreads();
    lv = bcast[0];
    li = bcastI[0];
    const double akk = bcast[1];

    // first Bunch-Kaufman tests (dense_kernel.hpp:601-618); static pivots
    // are always accepted at k (dense_kernel.hpp:982-1011)
    int action = -1;
    double dpiv = akk;
    if (stat) {
      const int sign = prm.expectedSign ? prm.expectedSign[fd.frontIdx[io + k]] : 0;
      double applied;
      dpiv = staticPivot(akk, sign, delta, sfloor, applied);
      if (applied != 0.0) {
        pert += fabs(applied);
        ++npert;
      }
      action = 1;
    } else if (lv <= prm.zeroTol) {
      const double isolatedFloor = fmax(prm.zeroTol, prm.relFloor * origMax);
      action = (fabs(akk) <= isolatedFloor) ? 0 : 1;
    } else if (fabs(akk) >= kAlpha * lv) {
      action = 1;
    }
    if (action == 0) {
      stop = true;
      break;
    }
    if (action == 1) {
      const double d = dpiv;
      double wj[SW];
#pragma unroll
      for (int j = 0; j < SW; ++j)
        wj[j] = (j > kk && j < sw) ? d * static_cast<double>(static_cast<T>(bcast[1 + j - kk] / d)) : 0.0;
      pdd[p] = d;
      pdo[p] = 0.0;
      pkd[p] = 1;
      if (tid >= 1 && tid <= SW && k + tid < nbEnd) lastL[tid] = static_cast<double>(static_cast<T>(bcast[1 + tid] / d));
      if (lead && tid == 0) {
        fd.dDiag[io + k] = d;
        fd.dOff[io + k] = 0.0;
        fd.pkind[io + k] = 1;
        Fat(k, k) = static_cast<T>(d);
      }
      if (lead) countInertia(d, npos, nneg, nzero);
      tmax = fmax(tmax, fabs(d));
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const int i = r0 + tid + jj * NT;
        if (i > k && i < nE) {
          const double c = col[i - r0];
          tmax = fmax(tmax, fabs(c));
          const T lt = static_cast<T>(c / d);
          // rows of the panel are read by every workgroup (their weights)
          if (i < k0 + kPanel + 2)
            storeCoherent(&Fat(i, k), lt);
          else
            Fat(i, k) = lt;
          const double l = lt;
#pragma unroll
          for (int j = 0; j < SW; ++j)
            if (j > kk && j < sw) Cs[j * CSR + (i - r0)] = fma(-l, wj[j], Cs[j * CSR + (i - r0)]);
        }
      }
      __syncthreads();
      ++k;
      ++p;
      if (k - c0 >= sw && k < nbEnd) loadSub(true);
      continue;
    }
    // general step: everyone hands its rows of column k to curK, workgroup 0
    // decides and updates, then everyone reloads its sub-panel
    for (int jj = 0; jj < 2; ++jj) {
      const int i = r0 + tid + jj * NT;
      if (i >= k && i < nE) curK[i] = col[i - r0];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    ++step;
    post(step);
    if (lead) {
      waitFor(0, W, step);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      if (!abandon) generalStep(lv, li, akk);
      if (tid == 0) {
        storeCoherent(&sy->state[0], k);
        storeCoherent(&sy->state[1], p);
        storeCoherent(&sy->state[2], stop ? 1 : 0);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      ++step;
      post(step);
    } else {
      ++step;
      waitFor(0, 1, step);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      const int nk = loadCoherent(&sy->state[0]);
      const int np = loadCoherent(&sy->state[1]);
      stop = loadCoherent(&sy->state[2]) != 0;
      // the pivots workgroup 0 took in the general step
      for (int q = p + tid; q < np; q += NT) {
        pdd[q] = fd.dDiag[io + k0 + q];
        pdo[q] = fd.dOff[io + k0 + q];
        pkd[q] = fd.pkind[io + k0 + q];
      }
      k = nk;
      p = np;
      __syncthreads();
    }
    if (stop || abandon) break;
    if (k < nbEnd) loadSub(false);
  }
  __syncthreads();

  const double bm = blockMax<NT>(tmax, redV);
  if (tid == 0) atomicMaxNonNeg(fd.runMaxBits + s, fmax(bm, origMax));
  if (lead && tid == 0) {
    fd.kcur[s] = k;
    fd.pcount[s] = p;
    fd.stopped[s] = stop ? 1 : 0;
    fd.inertia[3 * s] += npos;
    fd.inertia[3 * s + 1] += nneg;
    fd.inertia[3 * s + 2] += nzero;
    fd.pert[s] += pert;
    fd.npert[s] += npert;
    storeCoherent(&sy->epoch, static_cast<unsigned>((base >> 16) + 1));
  }
}

// ---------------------------------------------------------------------------
// Speculative panel step.  The threshold Bunch-Kaufman panel
// (dense_kernel.hpp:585-792) takes pivot k as a 1x1 without an interchange
// whenever |a_kk| >= alpha * lambda_k (lambda_k = max over the eligible rows
// below k of |a_ik|, current values), or lambda_k <= zero_tolerance and
// |a_kk| is above the isolated-pivot floor; static pivots are always taken
// at k (dense_kernel.hpp:982-1011).  When every pivot of a panel is taken
// that way, the panel is an unpivoted LDL^T of its diagonal block followed by
// the substitution of the rows below, and lambda_k is max |c_ik| over the
// values c_ik the substitution forms.  k_spec_factor does that factorization
// and the column maxima without writing the front; k_spec_apply finds the
// first column whose test fails, keeps the columns before it, and leaves the
// rest (still holding current values) to the exact panel kernel, which
// continues from there.
// ---------------------------------------------------------------------------
// Row values loaded in their storage type, every load issued before any is
// used: each load is unconditional (a skipped element reads the front's first
// entry, a valid address, and is zeroed afterwards), since a load under a
// branch completes before the branch joins.
template <int N, typename T, typename Off>
__device__ inline void loadWiden(double (&x)[N], const T* base, Off off) {
  T raw[N];
  bool ok[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const int64_t o = off(q);
    ok[q] = o >= 0;
    raw[q] = base[ok[q] ? o : 0];
  }
#pragma unroll
  for (int q = 0; q < N; ++q) x[q] = ok[q] ? static_cast<double>(raw[q]) : 0.0;
}

// 1 / d for a pivot on a dependent chain: the hardware reciprocal refined by
// two Newton steps (within an ulp of the quotient), an IEEE division for
// zero, denormal and non-finite d.
__device__ inline double pivotRecip(double d) {
  const double a = fabs(d);
  if (!(a >= DBL_MIN && a <= DBL_MAX)) return 1.0 / d;
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  return r;
}

struct SpecRec {
  unsigned long long* colmax;  // [2][kPanel]
  unsigned long long* emax;
  unsigned long long* nmax;
  double* cdiag;
  double* dval;
  double* bmax;
  double* rdiv;
  double* lblk;
  double* wmat;
  double* backup;
};
__device__ inline SpecRec specRec(double* ws, int nE) {
  double* r = ws + 2 * static_cast<int64_t>(nE);
  SpecRec x;
  x.colmax = reinterpret_cast<unsigned long long*>(r);
  x.emax = reinterpret_cast<unsigned long long*>(r + 2 * kPanel);
  x.nmax = reinterpret_cast<unsigned long long*>(r + 3 * kPanel);
  x.cdiag = r + 4 * kPanel;
  x.dval = r + 5 * kPanel;
  x.bmax = r + 6 * kPanel;
  x.rdiv = r + 7 * kPanel;
  x.lblk = r + 8 * kPanel;
  x.wmat = r + 8 * kPanel + kPanel * kPanel;
  x.backup = r + kSpecRec;
  return x;
}

// One row below the diagonal block: x holds its raw values in the panel
// columns; for q < lim in order, c_q = x_q - sum_{j<q} l_j W[q][j] with
// W[q][j] = d_j l_qj, then x[q] = l_q = T(c_q r_q), r_q = 1 / d_q (rdv: the
// pivots' reciprocals, so the chain over q carries no division).  onC(q, c_q)
// sees every current value; onGroup(g) runs after columns 16 g .. 16 g + 15
// (every thread of the workgroup reaches it).
//
// The coefficients are broadcast LDS reads, 16 at a time: a chunk's reads are
// all issued before its products (the empty asm statements keep the compiler
// from interleaving them), so their latencies overlap.
template <typename T, typename OnC, typename OnGroup>
__device__ inline void specRow(double (&x)[kPanel], const double (*W)[kPanel], const double* rdv, int lim, OnC onC,
                               OnGroup onGroup) {
#pragma unroll
  for (int g = 0; g < kPanel / 16; ++g) {
#pragma unroll
    for (int qq = 0; qq < 16; ++qq) {
      const int q = 16 * g + qq;
      if (q < lim) {
        double c0 = x[q], c1 = 0.0;
#pragma unroll
        for (int j0 = 0; j0 < q; j0 += 16) {
          double w[16];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            if (j0 + u < q) w[u] = W[q][j0 + u];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            if (j0 + u < q) asm volatile("" : "+v"(w[u]));
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            const int j = j0 + u;
            if (j < q) {
              if (j & 1)
                c1 = fma(-x[j], w[u], c1);
              else
                c0 = fma(-x[j], w[u], c0);
            }
          }
        }
        const double c = c0 + c1;
        onC(q, c);
        x[q] = static_cast<double>(static_cast<T>(c * rdv[q]));
      }
    }
    if (16 * g < lim) onGroup(g);
  }
}

// Smallest float at or above v (v >= 0): a bound that is never below the value.
__device__ inline double roundUpToFloat(double v) {
  float f = static_cast<float>(v);
  if (static_cast<double>(f) < v) f = nextafterf(f, INFINITY);
  return static_cast<double>(f);
}

// Pass 0: one workgroup per front factors the diagonal block (pw x pw,
// pw = min(kPanel, nE - k0)) in registers, 1x1 pivots in order, and writes
// the record: multipliers, pivots and their reciprocals, a_qq, the in-block
// column maxima and W = (d_j l_qj) for the row passes.
template <typename T>
__global__ __launch_bounds__(256) void k_spec_block(const int* list, int count, DevFronts fd, FactorParams prm,
                                                    double* wsBase, const int64_t* wsOff) {
  constexpr int P = kPanel;
  __shared__ double cvec[2][2][P];  // [step parity][column q, q + 1]
  __shared__ double dv[P], rdv[P];
  __shared__ int sgn[P];
  __shared__ double red[4][P];
  if (static_cast<int>(blockIdx.x) >= count) return;
  const int s = list[blockIdx.x];
  const int nE = fd.nE[s];
  if (!specFront(prm, nE) || fd.stopped[s]) return;
  const int m = fd.m[s];
  const int k0 = fd.kcur[s];
  if (k0 >= nE) return;
  const int pw = min(P, nE - k0);
  const T* F = static_cast<const T*>(fd.F[s]);
  const int64_t io = fd.idxOff[s];
  const SpecRec rec = specRec(wsBase + wsOff[s], nE);
  const int tid = threadIdx.x;
  const bool stat = prm.staticMode != 0;
  double origMax = bitsToDouble(fd.origMaxBits[s]);
  if (stat && origMax < prm.staticAbsFloor) origMax = 1.0;
  const double delta = prm.staticDelta > 0.0 ? prm.staticDelta : sqrt(DBL_EPSILON) * origMax;
  const double sfloor = fmax(prm.staticAbsFloor, prm.staticRelFloor * origMax);
  for (int q = tid; q < pw; q += 256)
    sgn[q] = (stat && prm.expectedSign) ? prm.expectedSign[fd.frontIdx[io + k0 + q]] : 0;
  // block entry (i, j), i >= j: column j = tid % 64, rows i = tid / 64 + 4 r
  const int j = tid & 63, ib = tid >> 6;
  double a[16];
  loadWiden<16>(a, F, [&](int r) -> int64_t {
    const int i = ib + 4 * r;
    return (i < pw && j < pw && i >= j) ? (k0 + i) + static_cast<int64_t>(k0 + j) * m : -1;
  });
  __syncthreads();
  // Right-looking: a_ij -= c_iq l_jq (c_iq = d_q l_iq up to the rounding of
  // the stored multiplier), l_jq = T(c_jq (1 / d_q)), two pivots per step.
  // Columns q and q + 1 go through LDS whole after one barrier; every lane
  // applies pivot q to its copy of column q + 1 (the same operations lane
  // q + 1 applies to its own) before taking pivot q + 1, so the result is that
  // of one pivot per step.  A column's lane updates all its rows: entries
  // above the diagonal or past the block take values that nothing reads, and
  // the lower-triangle entries combine lower-triangle entries only.
  for (int q = 0; q < pw; q += 2) {
    double* cv0 = cvec[(q >> 1) & 1][0];
    double* cv1 = cvec[(q >> 1) & 1][1];
    const bool two = q + 1 < pw;
    if (j == q) {
#pragma unroll
      for (int r = 0; r < 16; ++r) cv0[ib + 4 * r] = a[r];
    }
    if (j == q + 1) {
#pragma unroll
      for (int r = 0; r < 16; ++r) cv1[ib + 4 * r] = a[r];
    }
    __syncthreads();
    double ci0[16], ci1[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      ci0[r] = cv0[ib + 4 * r];
      ci1[r] = cv1[ib + 4 * r];
    }
    const double cj0 = cv0[j];
    double cj1 = cv1[j];
    double d0 = cv0[q];
    const double b10 = cv0[q + 1];  // c_{q+1,q}
    double d1 = cv1[q + 1];
    if (stat) {
      double applied;
      d0 = staticPivot(d0, sgn[q], delta, sfloor, applied);
    }
    const double rd0 = pivotRecip(d0);
    const double lj0 = static_cast<double>(static_cast<T>(cj0 * rd0));
    if (j > q && j < pw) {
#pragma unroll
      for (int r = 0; r < 16; ++r) a[r] = fma(-ci0[r], lj0, a[r]);
    }
    if (tid == 0) {
      dv[q] = d0;
      rdv[q] = rd0;
    }
    if (two) {
      // column q + 1 after pivot q
      const double l10 = static_cast<double>(static_cast<T>(b10 * rd0));
#pragma unroll
      for (int r = 0; r < 16; ++r) ci1[r] = fma(-ci0[r], l10, ci1[r]);
      cj1 = fma(-cj0, l10, cj1);
      d1 = fma(-b10, l10, d1);
      if (stat) {
        double applied;
        d1 = staticPivot(d1, sgn[q + 1], delta, sfloor, applied);
      }
      const double rd1 = pivotRecip(d1);
      const double lj1 = static_cast<double>(static_cast<T>(cj1 * rd1));
      if (j > q + 1 && j < pw) {
#pragma unroll
        for (int r = 0; r < 16; ++r) a[r] = fma(-ci1[r], lj1, a[r]);
      }
      if (tid == 0) {
        dv[q + 1] = d1;
        rdv[q + 1] = rd1;
      }
    }
  }
  // columns past the block: a neutral pivot for the branch-free row solve
  for (int q = pw + tid; q < P; q += 256) {
    dv[q] = 1.0;
    rdv[q] = 1.0;
  }
  __syncthreads();
  // a now holds c_ij (current value at pivot j) for i > j and a_jj^(j) on the diagonal
  double bm = 0.0;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = ib + 4 * r;
    if (j < pw && i < pw) {
      if (i > j) {
        const double l = static_cast<double>(static_cast<T>(a[r] * rdv[j]));
        rec.wmat[i * P + j] = dv[j] * l;
        rec.lblk[i * P + j] = l;
        bm = fmax(bm, fabs(a[r]));
      } else if (i == j) {
        rec.cdiag[j] = a[r];
      }
    }
  }
  red[ib][j] = bm;
  __syncthreads();
  if (tid < P) {
    rec.dval[tid] = dv[tid];
    rec.rdiv[tid] = rdv[tid];
    rec.bmax[tid] = fmax(fmax(red[0][tid], red[1][tid]), fmax(red[2][tid], red[3][tid]));
  }
}

// Pass 1: every workgroup takes kSpecRows rows below the block of a front
// (eligible rows [k0 + pw, nE) and rows [nE, m) alike), saves their raw panel
// values in the record's backup, forms their multipliers for the whole block
// and stores them in the front.  The column maxima of the current values go
// to the record: over the eligible rows rounded up to float (colmax[parity],
// the threshold test) and exact (emax), over the rows [nE, m) exact (nmax).
template <typename T>
__global__ __launch_bounds__(kSpecRows) void k_spec_factor(const int* list, int count, const int* wgStart, int parity,
                                                     DevFronts fd, FactorParams prm, double* wsBase,
                                                     const int64_t* wsOff) {
  constexpr int P = kPanel;
  constexpr int NT = kSpecRows;  // one row per thread
  constexpr int NW = NT / 64;
  __shared__ __align__(16) double W[P][P];  // whole 16-byte rows: broadcast reads of two coefficients at once
  __shared__ double rdv[P];
  __shared__ double cm[16][kSpecRows + 1];  // |c| (eligible rows) or -|c| (rows [nE, m)) of 16 columns
  __shared__ double partE[NW][P], partN[NW][P];
  if (static_cast<int>(blockIdx.x) >= wgStart[count]) return;
  const int f = findSegment(wgStart, count, blockIdx.x);
  const int s = list[f];
  const int w = blockIdx.x - wgStart[f];
  const int nE = fd.nE[s];
  if (!specFront(prm, nE) || fd.stopped[s]) return;
  const int m = fd.m[s];
  const int k0 = fd.kcur[s];
  if (k0 >= nE) return;
  const int pw = min(P, nE - k0);
  const int rBeg = k0 + pw + w * kSpecRows;
  if (rBeg >= m) return;
  T* F = static_cast<T*>(fd.F[s]);
  const SpecRec rec = specRec(wsBase + wsOff[s], nE);
  const int tid = threadIdx.x;
  const int i = rBeg + tid;
  const bool valid = i < m;
  const bool elig = i < nE;
  // the block's substitution coefficients, the reciprocal pivots and the
  // row's raw values all in flight together (coefficients outside the
  // block's lower triangle and skipped row entries are read and dropped)
  double wv[P * P / NT];
#pragma unroll
  for (int u = 0; u < P * P / NT; ++u) wv[u] = rec.wmat[tid + NT * u];
  const double rv = rec.rdiv[tid & (P - 1)];
  T raw[P];
  bool ok[P];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    ok[q] = valid && q < pw;
    raw[q] = F[ok[q] ? i + static_cast<int64_t>(k0 + q) * m : 0];
  }
#pragma unroll
  for (int u = 0; u < P * P / NT; ++u) {
    const int e = tid + NT * u, q = e / P, j = e % P;
    W[q][j] = (j < q && q < pw) ? wv[u] : 0.0;
  }
  if (tid < P) rdv[tid] = rv;
  __syncthreads();
  double x[P];
#pragma unroll
  for (int q = 0; q < P; ++q) x[q] = ok[q] ? static_cast<double>(raw[q]) : 0.0;
  // the raw values saved for the exact kernel (completed by the first
  // column group's barrier, long before anything reads them)
  T* bk = reinterpret_cast<T*>(rec.backup);
#pragma unroll
  for (int q = 0; q < P; ++q)
    if (ok[q]) bk[i + static_cast<int64_t>(q) * m] = raw[q];
  // column maxima through LDS, 16 columns at a time
  specRow<T>(
      x, W, rdv, pw,
      [&](int q, double c) {
        const double ac = fabs(c);
        cm[q & 15][tid] = elig ? ac : -ac;
      },
      [&](int g) {
        __syncthreads();
        const int col = tid & 15, seg = tid >> 4;  // 16 rows per thread
        double ve = 0.0, vn = 0.0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const double v = cm[col][seg * 16 + r];
          ve = fmax(ve, v);
          vn = fmax(vn, -v);
        }
        ve = fmax(ve, __shfl_xor(ve, 16, 64));
        ve = fmax(ve, __shfl_xor(ve, 32, 64));
        vn = fmax(vn, __shfl_xor(vn, 16, 64));
        vn = fmax(vn, __shfl_xor(vn, 32, 64));
        if ((tid & 63) < 16) {
          partE[tid >> 6][16 * g + col] = ve;
          partN[tid >> 6][16 * g + col] = vn;
        }
        __syncthreads();
      });
  if (valid) {
#pragma unroll
    for (int q = 0; q < P; ++q)
      if (q < pw) F[i + static_cast<int64_t>(k0 + q) * m] = static_cast<T>(x[q]);
  }
  if (tid < pw) {
    double ve = partE[0][tid], vn = partN[0][tid];
#pragma unroll
    for (int k = 1; k < NW; ++k) {
      ve = fmax(ve, partE[k][tid]);
      vn = fmax(vn, partN[k][tid]);
    }
    if (ve > 0.0) {
      atomicMaxNonNeg(rec.colmax + parity * P + tid, roundUpToFloat(ve));
      atomicMaxNonNeg(rec.emax + tid, ve);
    }
    if (vn > 0.0) atomicMaxNonNeg(rec.nmax + tid, vn);
  }
}

// Pass 2: every workgroup finds the first column q* whose test fails (all of
// them pass when static); when q* < pw it restores, for its rows, the raw
// values the exact kernel continues from (columns [q*, pw) of the eligible
// rows, the whole block of the rows [nE, m)).  Workgroup 0 commits the block's
// columns before q*, the pivots, inertia, perturbation counts and the running
// maximum, sets kcur = k0 + q*, pcount = q*, and clears the other colmax
// buffer and the exact maxima for the next panel.
template <typename T>
__global__ __launch_bounds__(kSpecRows) void k_spec_apply(const int* list, int count, const int* wgStart, int parity,
                                                    DevFronts fd, FactorParams prm, double* wsBase,
                                                    const int64_t* wsOff) {
  constexpr int P = kPanel;
  __shared__ int qsh;
  if (static_cast<int>(blockIdx.x) >= wgStart[count]) return;
  const int f = findSegment(wgStart, count, blockIdx.x);
  const int s = list[f];
  const int w = blockIdx.x - wgStart[f];
  const int tid = threadIdx.x;
  const int m = fd.m[s];
  const int nE = fd.nE[s];
  if (!specFront(prm, nE)) return;  // this front's panels are all exact
  const int k0 = fd.kcur[s];
  if (fd.stopped[s] || k0 >= nE) {
    if (w == 0 && tid == 0) {
      fd.pcount[s] = 0;
      fd.cbDone[s] = 0;
    }
    return;
  }
  const int pw = min(P, nE - k0);
  const int rBeg = k0 + pw + w * kSpecRows;
  if (w > 0 && rBeg >= m) return;
  T* F = static_cast<T*>(fd.F[s]);
  auto Fat = [&](int i, int j) -> T& { return F[i + static_cast<size_t>(j) * m]; };
  const int64_t io = fd.idxOff[s];
  const SpecRec rec = specRec(wsBase + wsOff[s], nE);
  const bool stat = prm.staticMode != 0;
  double origMax = bitsToDouble(fd.origMaxBits[s]);
  if (!stat && origMax < prm.zeroTol) origMax = 1.0;
  if (stat && origMax < prm.staticAbsFloor) origMax = 1.0;
  // first failing column: the first wave tests one column per lane
  if (tid < 64) {
    const int q = tid;
    bool fail = false;
    if (!stat && q < pw) {
      const double isolatedFloor = fmax(prm.zeroTol, prm.relFloor * origMax);
      const double lam = fmax(rec.bmax[q], bitsToDouble(rec.colmax[parity * P + q]));
      const double akk = rec.cdiag[q];
      fail = !((lam <= prm.zeroTol) ? fabs(akk) > isolatedFloor : fabs(akk) >= kAlpha * lam);
    }
    const unsigned long long bad = __ballot(fail);
    if (tid == 0) qsh = bad ? min(pw, __ffsll(static_cast<long long>(bad)) - 1) : pw;
  }
  __syncthreads();
  const int qs = qsh;
  if (qs < pw) {
    const int i = rBeg + tid;
    if (i < m) {
      const T* bk = reinterpret_cast<const T*>(rec.backup);
      for (int q = (i < nE) ? qs : 0; q < pw; ++q) Fat(i, k0 + q) = bk[i + static_cast<int64_t>(q) * m];
    }
  }
  if (w != 0) return;
  double tmax = 0.0;
  for (int e = tid; e < pw * pw; e += kSpecRows) {
    const int q = e / pw, r = e % pw;  // column q, row r of the block
    if (q < qs && r >= q) Fat(k0 + r, k0 + q) = static_cast<T>(r == q ? rec.dval[q] : rec.lblk[r * P + q]);
  }
  if (tid < P) rec.colmax[(parity ^ 1) * P + tid] = 0ull;
  if (tid < 64) {
    // pivots, inertia, perturbation and the maxima of the kept columns: one column per lane
    const int q = tid;
    const double ev = bitsToDouble(rec.emax[q]), nv = bitsToDouble(rec.nmax[q]);
    rec.emax[q] = 0ull;
    rec.nmax[q] = 0ull;
    int npos = 0, nneg = 0, nzero = 0, npert = 0;
    double pert = 0.0;
    if (q < qs) {
      const double d = rec.dval[q];
      fd.dDiag[io + k0 + q] = d;
      fd.dOff[io + k0 + q] = 0.0;
      fd.pkind[io + k0 + q] = 1;
      tmax = fmax(fmax(fabs(d), rec.bmax[q]), ev);
      countInertia(d, npos, nneg, nzero);
      if (stat) {
        const double delta = prm.staticDelta > 0.0 ? prm.staticDelta : sqrt(DBL_EPSILON) * origMax;
        const double sfloor = fmax(prm.staticAbsFloor, prm.staticRelFloor * origMax);
        const int sign = prm.expectedSign ? prm.expectedSign[fd.frontIdx[io + k0 + q]] : 0;
        double applied;
        (void)staticPivot(rec.cdiag[q], sign, delta, sfloor, applied);
        if (applied != 0.0) {
          pert = fabs(applied);
          npert = 1;
        }
      }
    }
    if (qs == pw && q < pw) tmax = fmax(tmax, nv);
    for (int off = 32; off > 0; off >>= 1) {
      npos += __shfl_xor(npos, off, 64);
      nneg += __shfl_xor(nneg, off, 64);
      nzero += __shfl_xor(nzero, off, 64);
      npert += __shfl_xor(npert, off, 64);
      pert += __shfl_xor(pert, off, 64);
      tmax = fmax(tmax, __shfl_xor(tmax, off, 64));
    }
    if (tid == 0) {
      fd.inertia[3 * s] += npos;
      fd.inertia[3 * s + 1] += nneg;
      fd.inertia[3 * s + 2] += nzero;
      fd.pert[s] += pert;
      fd.npert[s] += npert;
      fd.kcur[s] = k0 + qs;
      fd.pcount[s] = qs;
      fd.cbDone[s] = (qs == pw) ? 1 : 0;
      atomicMaxNonNeg(fd.runMaxBits + s, fmax(tmax, origMax));
    }
  }
}

// Clears the colmax buffers and exact maxima of the speculative records of a
// level's fronts.
__global__ void k_spec_reset(const int* list, int count, DevFronts fd, double* wsBase, const int64_t* wsOff) {
  const int f = blockIdx.x;
  if (f >= count) return;
  const int s = list[f];
  const SpecRec rec = specRec(wsBase + wsOff[s], fd.nE[s]);
  for (int t = threadIdx.x; t < 4 * kPanel; t += blockDim.x) rec.colmax[t] = 0ull;
}

// Multipliers of the rows that are not fully summed ([nE, m)) for the last
// panel.  With U = D L^T of the panel, the multipliers of a row are
// f U^{-1} = (f L^{-T}) D^{-1}: every workgroup forms M = L^{-1} of the
// panel's unit lower block in LDS (one column per thread, no barriers), then
// each thread maps one row: y = f M^T in registers, the running-max check on
// y, and l = y D^{-1} for the 1x1 and 2x2 pivot blocks (dense_kernel.hpp:
// 772-787 applied to the whole panel at once).
template <typename T>
__global__ __launch_bounds__(256) void k_panel_rows(const int* list, int count, const int* blockStart,
                                                     DevFronts fd) {
  constexpr int P = kPanel + 1;  // a 2x2 pivot at the panel's last column makes it 65
  __shared__ double Lp[P][P + 1];
  __shared__ double Mi[P][P + 1];
  __shared__ double Tt[P - 32][33];  // L21 M11 of the blocked inverse
  __shared__ double pdd[kPanel + 2], pdo[kPanel + 2];
  __shared__ int pkd[kPanel + 2];
  if (static_cast<int>(blockIdx.x) >= blockStart[count]) return;
  const int f = findSegment(blockStart, count, blockIdx.x);
  const int s = list[f];
  const int rowBlock = blockIdx.x - blockStart[f];
  const int p = fd.pcount[s];
  if (p == 0 || fd.cbDone[s]) return;  // the speculative step already wrote these rows
  const int m = fd.m[s];
  const int nE = fd.nE[s];
  const int k1 = fd.kcur[s];
  const int k0 = k1 - p;
  T* F = static_cast<T*>(fd.F[s]);
  const int64_t io = fd.idxOff[s];
  auto Fat = [&](int i, int j) -> T& { return F[i + static_cast<size_t>(j) * m]; };
  for (int q = threadIdx.x; q < p; q += blockDim.x) {
    pdd[q] = fd.dDiag[io + k0 + q];
    pdo[q] = fd.dOff[io + k0 + q];
    pkd[q] = fd.pkind[io + k0 + q];
  }
  __syncthreads();
  // unit lower L of the panel; the (s+1, s) slot of a 2x2 block holds d21
  for (int t = threadIdx.x; t < p * p; t += blockDim.x) {
    const int r = t % p, c = t / p;
    double v = 0.0;
    if (r > c && !(r == c + 1 && pkd[c] == 2)) v = Fat(k0 + r, k0 + c);
    Lp[r][c] = v;
  }
  __syncthreads();
  // M = L^{-1} by 2 x 2 blocks (the first H columns and the rest): both
  // diagonal blocks are inverted at once, column by column (column c:
  // M[c][c] = 1, M[r][c] = -sum_{c <= t < r} L[r][t] M[t][c]) by threads 0.. and
  // 64.. (two waves), then M21 = -M22 (L21 M11).
  constexpr int H = 32;
  const int h = min(H, p), nb = p - h;
  auto invCol = [&](int base, int nrow, int c) {
    Mi[base + c][base + c] = 1.0;
    for (int r = c + 1; r < nrow; ++r) {
      double a0 = Lp[base + r][base + c], a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int t = c + 1;
      for (; t + 4 <= r; t += 4) {
        a0 = fma(Lp[base + r][base + t], Mi[base + t][base + c], a0);
        a1 = fma(Lp[base + r][base + t + 1], Mi[base + t + 1][base + c], a1);
        a2 = fma(Lp[base + r][base + t + 2], Mi[base + t + 2][base + c], a2);
        a3 = fma(Lp[base + r][base + t + 3], Mi[base + t + 3][base + c], a3);
      }
      for (; t < r; ++t) a0 = fma(Lp[base + r][base + t], Mi[base + t][base + c], a0);
      Mi[base + r][base + c] = -((a0 + a1) + (a2 + a3));
    }
  };
  const int tid = threadIdx.x;
  if (tid < h)
    invCol(0, h, tid);
  else if (tid >= 64 && tid < 64 + nb)
    invCol(h, nb, tid - 64);
  __syncthreads();
  if (nb > 0) {
    // T = L21 M11 (M11 unit lower)
    for (int e = tid; e < nb * h; e += blockDim.x) {
      const int i = e % nb, j = e / nb;
      double a = 0.0;
      for (int k = j; k < h; ++k) a = fma(Lp[h + i][k], Mi[k][j], a);
      Tt[i][j] = a;
    }
    __syncthreads();
    // M21 = -(M22 T) (M22 unit lower)
    for (int e = tid; e < nb * h; e += blockDim.x) {
      const int i = e % nb, j = e / nb;
      double a = Tt[i][j];
      for (int k = 0; k < i; ++k) a = fma(Mi[h + i][h + k], Tt[k][j], a);
      Mi[h + i][j] = -a;
    }
  }
  __syncthreads();
  const int i = nE + rowBlock * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double x[P];
  loadWiden<P>(x, static_cast<const T*>(F), [&](int q) -> int64_t {
    return (q < p) ? i + static_cast<int64_t>(k0 + q) * m : -1;
  });
  // y = x M^T in place, last column first (y_q needs x_0 .. x_q only)
#pragma unroll
  for (int q = P - 1; q > 0; --q) {
    if (q < p) {
      double a0 = x[q], a1 = 0.0;
#pragma unroll
      for (int qp = 0; qp < q; ++qp) {
        if (qp & 1)
          a1 = fma(x[qp], Mi[q][qp], a1);
        else
          a0 = fma(x[qp], Mi[q][qp], a0);
      }
      x[q] = a0 + a1;
    }
  }
  double mx = 0.0;
#pragma unroll
  for (int q = 0; q < P; ++q) {
    if (q < p) {
      const int kind = pkd[q];
      if (kind == 1) {
        mx = fmax(mx, fabs(x[q]));
        Fat(i, k0 + q) = static_cast<T>(x[q] / pdd[q]);
      } else if (kind == 2 && q + 1 < P) {
        const double x0 = x[q], x1 = x[q + 1];
        mx = fmax(mx, fmax(fabs(x0), fabs(x1)));
        const double d11 = pdd[q], d21 = pdo[q], d22 = pdd[q + 1];
        const double det = d11 * d22 - d21 * d21;
        const double inv11 = d22 / det, inv21 = -d21 / det, inv22 = d11 / det;
        Fat(i, k0 + q) = static_cast<T>(x0 * inv11 + x1 * inv21);
        Fat(i, k0 + q + 1) = static_cast<T>(x0 * inv21 + x1 * inv22);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMaxNonNeg(fd.runMaxBits + s, mx);
}

// Trailing update T(k1:m, k1:m) -= (L D) L^T over the lower triangle, for
// the last panel's p pivots, on the matrix cores: v_mfma_f64_16x16x4f64 for
// FP64 fronts, v_mfma_f32_16x16x4f32 on FP32 operands for FP32 fronts (L D
// formed in FP64 from the FP64 pivots, then rounded).  Operand layout of the
// instructions on gfx950 (measured): A[i][k] from lane i + 16k, B[k][j] from
// lane j + 16k; the result D[i][j] is in lane j + 16*(i%4), register i/4
// (f64) or lane j + 16*(i/4), register i%4 (f32).
//
// 64x64 lower-triangle output tiles, four waves of 32x32.  The workgroups are
// persistent: each walks the tiles t = blockIdx.x, + gridDim.x, ...; while a
// tile's products and write-back run, the next tile's operand rows and C
// values are already being loaded into registers, and a front's pivots are
// read only when the walk enters a new front.  The product is formed
// transposed (L operand first), so a register holds 16 consecutive rows of a
// column: C is read and written in whole column segments.
template <typename T>
struct SchurMfma;
template <>
struct SchurMfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // column (within the 16 x 16 block) of register r of this lane; the row is lane % 16
  static __device__ inline int col(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <>
struct SchurMfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ inline int col(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// A front pointer read from the front table is generic; generic (flat)
// accesses count against the LDS wait counter too, so every LDS wait would
// also wait for them.  Front storage is global memory: this view makes the
// accesses global ones.
template <typename T>
using GlobalPtr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ inline GlobalPtr<T> globalView(void* p) {
  return (GlobalPtr<T>)(p);
}

template <typename T>
__global__ __launch_bounds__(256, 4) void k_schur(const int* list, int count, const int* tileStart, const int* tileTp,
                                               DevFronts fd) {
  using M = SchurMfma<T>;
  constexpr int TS = 64;
  constexpr int KP = kPanel + 4;  // padded K (a 2x2 at the panel end makes p = 65)
  constexpr int KU = KP / 4;      // staged columns per thread
  __shared__ T As[TS][KP + 1];    // L D rows of the tile's row block
  __shared__ T Bs[TS][KP + 1];    // L rows of the tile's column block
  __shared__ double pdd[kPanel + 2], pdo[kPanel + 2];
  __shared__ int pkd[kPanel + 2];
  const int total = tileStart[count];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  const int sr = tid & 63, qb = tid >> 6;  // staging: row sr of both operand tiles, columns qb + 4u

  // Geometry of tile t.  A front's tiles are numbered column block by
  // column block (tj, then ti = tj .. Tp-1 down the column), so the
  // workgroups running at the same time read and write consecutive pieces of
  // the same columns.  Tp is the planned tiles per side; the launch covers
  // the column blocks that start before the front's cut (all Tp of them, or
  // fewer when its contribution block is updated separately).  The front's
  // metadata is read again only when the walk enters another front.
  struct Tile {
    int f, s, p, m, k0, k1, Tp, cut, lo, hi, ti, tj;
    GlobalPtr<T> F;
  };
  auto tileOf = [&](int t, const Tile* prev) -> Tile {
    Tile x;
    if (prev && t >= prev->lo && t < prev->hi) {
      x = *prev;
    } else {
      x.f = findSegment(tileStart, count, t);
      x.lo = tileStart[x.f];
      x.hi = tileStart[x.f + 1];
      x.s = list[x.f];
      x.p = fd.pcount[x.s];
      x.m = fd.m[x.s];
      x.k1 = fd.kcur[x.s];
      x.k0 = x.k1 - x.p;
      x.F = globalView<T>(fd.F[x.s]);
      x.Tp = tileTp[x.f];
      x.cut = fd.schurCut[x.s];
    }
    // column block tj starts at S(tj) = tj Tp - tj (tj - 1) / 2
    const int lt = t - x.lo;
    const double b = x.Tp + 0.5;
    int tj = static_cast<int>(b - sqrt(fmax(b * b - 2.0 * lt, 0.0)));
    auto S = [&](int c) { return c * x.Tp - c * (c - 1) / 2; };
    while (tj > 0 && S(tj) > lt) --tj;
    while (tj + 1 < x.Tp && S(tj + 1) <= lt) ++tj;
    x.tj = tj;
    x.ti = tj + (lt - S(tj));
    return x;
  };
  // a tile does work when its front took pivots and the tile lies in the
  // trailing block, left of the cut
  auto live = [&](const Tile& x) { return x.p > 0 && x.k1 + x.ti * TS < x.m && x.k1 + x.tj * TS < x.cut; };
  T la[KU], lb[KU], cv[2][2][4];
  auto loadTile = [&](const Tile& x) {
    const int i0 = x.k1 + x.ti * TS, j0 = x.k1 + x.tj * TS;
    const int ri = i0 + sr, rj = j0 + sr;
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int q = qb + 4 * u;
      la[u] = (q < x.p && ri < x.m) ? x.F[ri + static_cast<size_t>(x.k0 + q) * x.m] : T(0);
      lb[u] = (q < x.p && rj < x.m) ? x.F[rj + static_cast<size_t>(x.k0 + q) * x.m] : T(0);
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int row = i0 + wr + bi * 16 + lr;
          const int col = j0 + wc + bj * 16 + M::col(lane, rr);
          cv[bi][bj][rr] = (row < x.m && col < x.cut && (x.ti != x.tj || row >= col))
                               ? x.F[row + static_cast<size_t>(col) * x.m]
                               : T(0);
        }
  };

  // XCD-aware walk: workgroups go to the eight XCDs round-robin (blockIdx
  // modulo 8), so each XCD's workgroups take a contiguous eighth of the tiles
  // and neighbouring tiles, which share operand rows, meet in one L2.
  // (Any other placement only costs L2 hits.)
  constexpr int kXcd = 8;
  int t, stride, tEnd;
  if (static_cast<int>(gridDim.x) >= kXcd) {
    const int xcd = blockIdx.x % kXcd;
    const int chunk = (total + kXcd - 1) / kXcd;
    t = xcd * chunk + static_cast<int>(blockIdx.x) / kXcd;
    stride = (static_cast<int>(gridDim.x) - xcd + kXcd - 1) / kXcd;  // workgroups with this residue
    tEnd = min(total, (xcd + 1) * chunk);
  } else {
    t = blockIdx.x;
    stride = gridDim.x;
    tEnd = total;
  }
  if (t >= tEnd) return;
  Tile cur = tileOf(t, nullptr);
  int pivFront = -1;  // front whose pivots are in pdd / pdo / pkd
  // running max |F| of the updated entries: per lane over the tiles of one
  // front, folded into the front's record when the walk leaves it
  int mxFront = -1;
  double lmx = 0.0;
  auto flushMax = [&]() {
    if (mxFront < 0) return;
    const double wmx = __ockl_wfred_max_f64(lmx);
    if (lane == 0) atomicMaxNonNeg(fd.runMaxBits + mxFront, wmx);
  };
  if (live(cur)) loadTile(cur);
  while (t < tEnd) {
    const int tn = t + stride;
    if (live(cur)) {
      if (cur.s != pivFront) {
        for (int q = tid; q < cur.p; q += blockDim.x) {
          const int64_t io = fd.idxOff[cur.s];
          pdd[q] = fd.dDiag[io + cur.k0 + q];
          pdo[q] = fd.dOff[io + cur.k0 + q];
          pkd[q] = fd.pkind[io + cur.k0 + q];
        }
        pivFront = cur.s;
      }
      if (cur.ti == 0 && cur.tj == 0 && tid == 0 && fd.schurFlops) {
        // p r (r + 1) over the trailing lower triangle, less the deferred contribution block
        const unsigned long long N = static_cast<unsigned long long>(cur.m - cur.k1);
        const unsigned long long C = static_cast<unsigned long long>(cur.m - max(cur.cut, cur.k1));
        atomicAdd(fd.schurFlops, static_cast<unsigned long long>(cur.p) * (N * (N + 1) - C * (C + 1)));
      }
      // stage the operands: raw L rows, then the L D scaling (a 2x2 block
      // needs its neighbouring column, read back from LDS)
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        As[sr][qb + 4 * u] = la[u];
        Bs[sr][qb + 4 * u] = lb[u];
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        const int q = qb + 4 * u;
        double a = 0.0;
        if (q < cur.p) {
          const int kind = pkd[q];
          const double l = static_cast<double>(la[u]);
          if (kind == 1)
            a = l * pdd[q];
          else if (kind == 2)
            a = l * pdd[q] + static_cast<double>(As[sr][q + 1]) * pdo[q];
          else if (q > 0)
            a = static_cast<double>(As[sr][q - 1]) * pdo[q - 1] + l * pdd[q];
        }
        la[u] = static_cast<T>(a);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < KU; ++u) As[sr][qb + 4 * u] = la[u];
      __syncthreads();
    }
    // the next tile's operands and C travel while this tile computes
    T cn[2][2][4];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) cn[bi][bj][rr] = cv[bi][bj][rr];
    Tile nxt;
    if (tn < tEnd) {
      nxt = tileOf(tn, &cur);
      if (live(nxt)) loadTile(nxt);
    }
    if (live(cur)) {
      const int pp = (cur.p + 3) & ~3;
      typename M::v4 acc[2][2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = typename M::v4{0, 0, 0, 0};
#pragma unroll 4
      for (int kk = 0; kk < pp; kk += 4) {
        const T a0 = As[wr + lr][kk + lk];
        const T a1 = As[wr + 16 + lr][kk + lk];
        const T b0 = Bs[wc + lr][kk + lk];
        const T b1 = Bs[wc + 16 + lr][kk + lk];
        acc[0][0] = M::mma(b0, a0, acc[0][0]);
        acc[0][1] = M::mma(b1, a0, acc[0][1]);
        acc[1][0] = M::mma(b0, a1, acc[1][0]);
        acc[1][1] = M::mma(b1, a1, acc[1][1]);
      }
      const int i0 = cur.k1 + cur.ti * TS, j0 = cur.k1 + cur.tj * TS;
      if (cur.s != mxFront) {
        flushMax();
        mxFront = cur.s;
        lmx = 0.0;
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            const int row = i0 + wr + bi * 16 + lr;
            const int col = j0 + wc + bj * 16 + M::col(lane, rr);
            if (row < cur.m && col < cur.cut && (cur.ti != cur.tj || row >= col)) {
              const T v = static_cast<T>(cn[bi][bj][rr] - acc[bi][bj][rr]);
              cur.F[row + static_cast<size_t>(col) * cur.m] = v;
              lmx = fmax(lmx, fabs(static_cast<double>(v)));
            }
          }
      __syncthreads();  // the LDS tiles are restaged for the next tile
    }
    t = tn;
    if (t < tEnd) cur = nxt;
  }
  flushMax();
}

// Operand of the deferred contribution-block update of front s (a front
// whose panels' trailing updates stopped at column nE): for columns
// q0 .. q0 + kc - 1, W = L(nE:m, q) D, the 2x2 blocks coupling neighbouring
// columns as in the tile staging above; zero in the columns at or past the
// pivots the front took, so the rank-kc product CB -= W L(nE:m, q0:q0+kc)^T
// that follows adds nothing for them.  W is column-major with ld = m - nE.
constexpr int kFormCols = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_form_w(DevFronts fd, int s, int q0, int kc, T* __restrict__ W) {
  __shared__ double pdd[kFormCols + 2], pdo[kFormCols + 2];
  __shared__ int pkd[kFormCols + 2];
  const int m = fd.m[s], nE = fd.nE[s], nF = fd.nF[s];
  const int N = m - nE;
  const int c0 = q0 + blockIdx.y * kFormCols;
  const int c1 = min(c0 + kFormCols, q0 + kc);
  const int64_t io = fd.idxOff[s];
  // slot t + 1 holds column c0 + t; slot 0 column c0 - 1
  for (int t = threadIdx.x; t < kFormCols + 2; t += blockDim.x) {
    const int q = c0 - 1 + t;
    const bool ok = q >= 0 && q < nF;
    pdd[t] = ok ? fd.dDiag[io + q] : 0.0;
    pdo[t] = ok ? fd.dOff[io + q] : 0.0;
    pkd[t] = ok ? fd.pkind[io + q] : -1;
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const T* F = static_cast<const T*>(fd.F[s]);
  const T* Lr = F + nE + i;  // row nE + i of the front, column q at Lr[q * m]
  for (int q = c0; q < c1; ++q) {
    const int t = q - c0 + 1;
    const int kind = pkd[t];
    double a = 0.0;
    if (kind == 1) {
      a = static_cast<double>(Lr[static_cast<size_t>(q) * m]) * pdd[t];
    } else if (kind == 2) {
      a = static_cast<double>(Lr[static_cast<size_t>(q) * m]) * pdd[t] +
          static_cast<double>(Lr[static_cast<size_t>(q + 1) * m]) * pdo[t];
    } else if (kind == 0) {
      a = static_cast<double>(Lr[static_cast<size_t>(q - 1) * m]) * pdo[t - 1] +
          static_cast<double>(Lr[static_cast<size_t>(q) * m]) * pdd[t];
    }
    W[i + static_cast<size_t>(q - q0) * N] = static_cast<T>(a);
  }
}

__global__ void k_front_finish(const int* list, int count, DevFronts fd) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < count) fd.nF[list[f]] = fd.kcur[list[f]];
}

// rowIdx = frontIdx[perm[i]] for every position of the listed fronts.
__global__ void k_make_rowidx(const int* list, int count, DevFronts fd) {
  if (static_cast<int>(blockIdx.x) >= count) return;
  const int s = list[blockIdx.x];
  const int m = fd.m[s];
  const int64_t io = fd.idxOff[s];
  for (int i = threadIdx.x; i < m; i += blockDim.x) fd.rowIdx[io + i] = fd.frontIdx[io + fd.perm[io + i]];
}

// Sum of the per-front statistics (multifrontal.hpp:795-801); one block.
// Each thread sums a contiguous range of fronts in order and thread 0 adds
// the partial sums in thread order, so the perturbation total does not
// depend on timing.
constexpr int kStatsThreads = 1024;
__global__ __launch_bounds__(kStatsThreads) void k_reduce_stats(int numSN, DevFronts fd, long long* outI,
                                                                double* outD) {
  __shared__ long long sI[5][kStatsThreads];
  __shared__ double sD[kStatsThreads];
  long long pos = 0, neg = 0, zero = 0, delayedFronts = 0, npert = 0;
  double pert = 0.0;
  const int chunk = (numSN + blockDim.x - 1) / blockDim.x;
  const int s0 = threadIdx.x * chunk;
  const int s1 = min(numSN, s0 + chunk);
  for (int s = s0; s < s1; ++s) {
    pos += fd.inertia[3 * s];
    neg += fd.inertia[3 * s + 1];
    zero += fd.inertia[3 * s + 2];
    delayedFronts += (fd.nF[s] < fd.nE[s]) ? 1 : 0;
    npert += fd.npert[s];
    pert += fd.pert[s];
  }
  sI[0][threadIdx.x] = pos;
  sI[1][threadIdx.x] = neg;
  sI[2][threadIdx.x] = zero;
  sI[3][threadIdx.x] = delayedFronts;
  sI[4][threadIdx.x] = npert;
  sD[threadIdx.x] = pert;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long r[5] = {0, 0, 0, 0, 0};
    double pt = 0.0;
    for (int t = 0; t < static_cast<int>(blockDim.x); ++t) {
      for (int q = 0; q < 5; ++q) r[q] += sI[q][t];
      pt += sD[t];
    }
    for (int q = 0; q < 5; ++q) outI[q] = r[q];
    outD[0] = pt;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launchResetFronts(int numSN, const DevFronts& fd, hipStream_t st) {
  if (numSN <= 0) return;
  const int blocks = (numSN + 255) / 256;
  k_reset_fronts<<<blocks, 256, 0, st>>>(numSN, fd);
  hipCheck(hipGetLastError(), "launchResetFronts");
}
void launchResetPerm(const int* list, int count, const DevFronts& fd, hipStream_t st) {
  if (count <= 0) return;
  k_reset_perm<<<count, 256, 0, st>>>(list, count, fd);
  hipCheck(hipGetLastError(), "launchResetPerm");
}
void launchAssemble(const int4* tasks, int ntasks, const DevFronts& fd, const double* aVals, int monotone,
                    hipStream_t st) {
  if (ntasks <= 0) return;
  if (fd.fp32)
    k_assemble<float><<<ntasks, 256, 0, st>>>(tasks, ntasks, fd, aVals, monotone);
  else
    k_assemble<double><<<ntasks, 256, 0, st>>>(tasks, ntasks, fd, aVals, monotone);
  hipCheck(hipGetLastError(), "launchAssemble");
}
template <typename T>
static void factorSmallT(const int* list, int count, int maxM, const DevFronts& fd, const FactorParams& prm,
                         hipStream_t st) {
  const size_t smem = smallFrontSmemBytes(maxM);
  static bool attrSet = false;
  if (!attrSet) {
    hipCheck(hipFuncSetAttribute(reinterpret_cast<const void*>(k_factor_small<T>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(smallFrontSmemBytes(kSmallFrontMax))),
             "k_factor_small LDS size");
    attrSet = true;
  }
  k_factor_small<T><<<count, kSmallThreads, smem, st>>>(list, count, fd, prm);
  hipCheck(hipGetLastError(), "factorSmallT");
}
void launchFactorSmall(const int* list, int count, int maxM, const DevFronts& fd, const FactorParams& prm,
                       hipStream_t st) {
  if (count <= 0) return;
  if (fd.fp32)
    factorSmallT<float>(list, count, maxM, fd, prm, st);
  else
    factorSmallT<double>(list, count, maxM, fd, prm, st);
  hipCheck(hipGetLastError(), "launchFactorSmall");
}
void launchFrontMaxAbs(const int* list, int count, int maxM, const DevFronts& fd, hipStream_t st) {
  if (count <= 0) return;
  int bpf = (maxM + 15) / 16;
  if (bpf > 64) bpf = 64;
  if (bpf < 1) bpf = 1;
  if (fd.fp32)
    k_front_maxabs<float><<<count * bpf, 256, 0, st>>>(list, count, bpf, fd);
  else
    k_front_maxabs<double><<<count * bpf, 256, 0, st>>>(list, count, bpf, fd);
  hipCheck(hipGetLastError(), "launchFrontMaxAbs");
}
template <int RM, typename T>
static void panelLaunch(const int* list, int count, const DevFronts& fd, const FactorParams& prm, double* ws,
                        const int64_t* wsOff, hipStream_t st) {
  const size_t lds = PanelShape<RM>::lds();
  static bool attrSet = false;
  if (!attrSet && lds > 0) {
    hipCheck(hipFuncSetAttribute(reinterpret_cast<const void*>(k_panel<RM, T>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)),
             "k_panel LDS size");
    attrSet = true;
  }
  k_panel<RM, T><<<count, kPanelThreads, lds, st>>>(list, count, fd, prm, ws, wsOff);
  hipCheck(hipGetLastError(), "panelLaunch");
}
template <typename T>
static void panelT(const int* list, int count, int maxRows, const DevFronts& fd, const FactorParams& prm, double* ws,
                   const int64_t* wsOff, hipStream_t st) {
  // register slots per thread for the eligible rows of the widest front
  const int slots = (maxRows + kPanelThreads - 1) / kPanelThreads;
  if (slots <= 1)
    panelLaunch<1, T>(list, count, fd, prm, ws, wsOff, st);
  else if (slots <= 2)
    panelLaunch<2, T>(list, count, fd, prm, ws, wsOff, st);
  else if (slots <= 4)
    panelLaunch<4, T>(list, count, fd, prm, ws, wsOff, st);
  else if (slots <= 8)
    panelLaunch<8, T>(list, count, fd, prm, ws, wsOff, st);
  else if (slots <= 16)
    panelLaunch<16, T>(list, count, fd, prm, ws, wsOff, st);
  else if (slots <= 32)
    panelLaunch<32, T>(list, count, fd, prm, ws, wsOff, st);
  else
    panelLaunch<0, T>(list, count, fd, prm, ws, wsOff, st);
}
void launchPanel(const int* list, int count, int maxRows, const DevFronts& fd, const FactorParams& prm, double* ws,
                 const int64_t* wsOff, hipStream_t st) {
  if (count <= 0) return;
  if (fd.fp32)
    panelT<float>(list, count, maxRows, fd, prm, ws, wsOff, st);
  else
    panelT<double>(list, count, maxRows, fd, prm, ws, wsOff, st);
  hipCheck(hipGetLastError(), "launchPanel");
}
void launchPanelMw(const int* list, int count, const int* wgStart, int blocks, const DevFronts& fd,
                   const FactorParams& prm, double* ws, const int64_t* wsOff, MwSync* sync, int* err,
                   hipStream_t st) {
  if (count <= 0 || blocks <= 0) return;
  constexpr size_t lds = sizeof(double) * 8 * kMwRows;
  static bool attrSet[2] = {false, false};
  if (fd.fp32) {
    if (!attrSet[1]) {
      hipCheck(hipFuncSetAttribute(reinterpret_cast<const void*>(k_panel_mw<float>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)),
               "k_panel_mw LDS size");
      attrSet[1] = true;
    }
    k_panel_mw<float><<<blocks, kPanelThreads, lds, st>>>(list, count, wgStart, fd, prm, ws, wsOff, sync, err);
  } else {
    if (!attrSet[0]) {
      hipCheck(hipFuncSetAttribute(reinterpret_cast<const void*>(k_panel_mw<double>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)),
               "k_panel_mw LDS size");
      attrSet[0] = true;
    }
    k_panel_mw<double><<<blocks, kPanelThreads, lds, st>>>(list, count, wgStart, fd, prm, ws, wsOff, sync, err);
  }
  hipCheck(hipGetLastError(), "launchPanelMw");
}
void launchSpecReset(const int* list, int count, const DevFronts& fd, double* ws, const int64_t* wsOff,
                     hipStream_t st) {
  if (count <= 0) return;
  k_spec_reset<<<count, 128, 0, st>>>(list, count, fd, ws, wsOff);
  hipCheck(hipGetLastError(), "launchSpecReset");
}
void launchSpecPanel(const int* list, int count, const int* wgStart, int blocks, int parity, const DevFronts& fd,
                     const FactorParams& prm, double* ws, const int64_t* wsOff, hipStream_t st) {
  if (count <= 0 || blocks <= 0) return;
  if (fd.fp32) {
    k_spec_block<float><<<count, 256, 0, st>>>(list, count, fd, prm, ws, wsOff);
    hipCheck(hipGetLastError(), "k_spec_block");
    k_spec_factor<float><<<blocks, kSpecRows, 0, st>>>(list, count, wgStart, parity, fd, prm, ws, wsOff);
    hipCheck(hipGetLastError(), "k_spec_factor");
    k_spec_apply<float><<<blocks, kSpecRows, 0, st>>>(list, count, wgStart, parity, fd, prm, ws, wsOff);
  } else {
    k_spec_block<double><<<count, 256, 0, st>>>(list, count, fd, prm, ws, wsOff);
    hipCheck(hipGetLastError(), "k_spec_block");
    k_spec_factor<double><<<blocks, kSpecRows, 0, st>>>(list, count, wgStart, parity, fd, prm, ws, wsOff);
    hipCheck(hipGetLastError(), "k_spec_factor");
    k_spec_apply<double><<<blocks, kSpecRows, 0, st>>>(list, count, wgStart, parity, fd, prm, ws, wsOff);
  }
  hipCheck(hipGetLastError(), "launchSpecPanel");
}
void launchPanelRows(const int* list, int count, const int* blockStart, int totalBlocks, const DevFronts& fd,
                     hipStream_t st) {
  if (count <= 0 || totalBlocks <= 0) return;
  if (fd.fp32)
    k_panel_rows<float><<<totalBlocks, 256, 0, st>>>(list, count, blockStart, fd);
  else
    k_panel_rows<double><<<totalBlocks, 256, 0, st>>>(list, count, blockStart, fd);
  hipCheck(hipGetLastError(), "launchPanelRows");
}
void launchSyrk(const int* list, int count, const int* tileStart, const int* tileTp, int totalTiles,
                const DevFronts& fd, hipStream_t st) {
  if (count <= 0 || totalTiles <= 0) return;
  // persistent workgroups: as many as are resident at once (LDS and
  // registers allow two per CU in FP64, four in FP32 with the kernel's
  // register bound), so that none starts after the others have taken their
  // share of tiles
  static int resident[2] = {0, 0};
  int& res = resident[fd.fp32 ? 1 : 0];
  if (res == 0) {
    int dev = 0, cus = 0, per = 0;
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    hipCheck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    hipCheck(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                 &per, fd.fp32 ? reinterpret_cast<const void*>(k_schur<float>) : reinterpret_cast<const void*>(k_schur<double>),
                 256, 0),
             "k_schur occupancy");
    res = std::max(1, per) * std::max(1, cus);
  }
  const int blocks = std::min(totalTiles, res);
  if (fd.fp32)
    k_schur<float><<<blocks, 256, 0, st>>>(list, count, tileStart, tileTp, fd);
  else
    k_schur<double><<<blocks, 256, 0, st>>>(list, count, tileStart, tileTp, fd);
  hipCheck(hipGetLastError(), "launchSyrk");
}
void launchFormW(const DevFronts& fd, int s, int rows, int q0, int kc, void* W, hipStream_t st) {
  if (rows <= 0 || kc <= 0) return;
  const dim3 grid((rows + 255) / 256, (kc + kFormCols - 1) / kFormCols);
  if (fd.fp32)
    k_form_w<float><<<grid, 256, 0, st>>>(fd, s, q0, kc, static_cast<float*>(W));
  else
    k_form_w<double><<<grid, 256, 0, st>>>(fd, s, q0, kc, static_cast<double*>(W));
  hipCheck(hipGetLastError(), "launchFormW");
}
void launchFrontFinish(const int* list, int count, const DevFronts& fd, hipStream_t st) {
  if (count <= 0) return;
  k_front_finish<<<(count + 255) / 256, 256, 0, st>>>(list, count, fd);
  hipCheck(hipGetLastError(), "launchFrontFinish");
}
void launchMakeRowIdx(const int* list, int count, const DevFronts& fd, hipStream_t st) {
  if (count <= 0) return;
  k_make_rowidx<<<count, 256, 0, st>>>(list, count, fd);
  hipCheck(hipGetLastError(), "launchMakeRowIdx");
}
void launchReduceStats(int numSN, const DevFronts& fd, long long* outI, double* outD, hipStream_t st) {
  k_reduce_stats<<<1, kStatsThreads, 0, st>>>(numSN, fd, outI, outD);
  hipCheck(hipGetLastError(), "launchReduceStats");
}

}  // namespace symla_hip
