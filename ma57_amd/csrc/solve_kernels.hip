// solve_kernels.hip -- gfx950 kernels of the multifrontal triangular solve
// (forward L, block-diagonal D, backward L^T) with many right-hand sides, and
// of the FP64 residual / norm computations used by iterative refinement.
//
// The forward pass follows the reference's task-parallel formulation
// (include/symla/solve.hpp:303-362): each front sums its children's
// contributions (fixed child order, no atomics), solves its own rows, and
// hands a private delta for its remaining rows to its parent.  The backward
// pass is the reference's per-front step (solve.hpp:364-383) run top-down.
// Per tree level, the forward pass gathers the children's deltas, solves the
// fronts' own rows in 32-row blocks (one workgroup per front when its chain
// of blocks is short; else one workgroup per block, a block waiting on a
// per-block flag for the blocks it depends on), then forms the deltas of the
// remaining rows with no waits.  The backward pass first takes the
// ancestors' part of every own block, then runs the own-row chains the same
// two ways.
#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"

namespace symla_hip {

// Right-hand sides inside a solve are row-major: Bp, Z and X are n x nrhs
// with row r's right-hand sides contiguous at r * nrhs, and a front's work
// block is m x nrhs, row i at (gOff[s] + i) * nrhs.  A row reached through a
// front's index map is then one contiguous read of nrhs words.

// Caller's column-major B (ld ldb) into the solve's row order through a
// 64 x 64 tile in LDS: read with lanes over rows (the gather through perm
// stays inside one column), write with lanes over right-hand sides.
__global__ __launch_bounds__(256) void k_permute_in(int n, int nrhs, const int* __restrict__ perm,
                                                    const double* __restrict__ B, int64_t ldb,
                                                    double* __restrict__ Bp) {
  __shared__ double tile[64][65];
  const int t0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tn = min(64, n - t0), cn = min(64, nrhs - c0);
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int t = e & 63, c = e >> 6;
    if (t < tn && c < cn) tile[t][c] = B[perm[t0 + t] + static_cast<int64_t>(c0 + c) * ldb];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int c = e & 63, t = e >> 6;
    if (t < tn && c < cn) Bp[static_cast<int64_t>(t0 + t) * nrhs + c0 + c] = tile[t][c];
  }
}

__global__ __launch_bounds__(256) void k_permute_out(int n, int nrhs, const int* __restrict__ perm,
                                                     const double* __restrict__ X, double* __restrict__ Xo,
                                                     int64_t ldx) {
  __shared__ double tile[64][65];
  const int t0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tn = min(64, n - t0), cn = min(64, nrhs - c0);
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int c = e & 63, t = e >> 6;
    if (t < tn && c < cn) tile[t][c] = X[static_cast<int64_t>(t0 + t) * nrhs + c0 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int t = e & 63, c = e >> 6;
    if (t < tn && c < cn) Xo[perm[t0 + t] + static_cast<int64_t>(c0 + c) * ldx] = tile[t][c];
  }
}

namespace {

constexpr int SB = kSolveBlock;  // own rows per solve block; also the column tile of the own-row products
constexpr int DT = 64;           // row tile of the delta and ancestor products
constexpr int YW = 65;           // row stride of the right-hand-side tiles in LDS (64 columns + 1)
// Polls of a dependency flag before a chain block gives up (about a second).
constexpr unsigned kSpinLimit = 1u << 22;

// Front owning flat item b: start[f] <= b < start[f+1].
template <typename I>
__device__ inline int segmentOf(const I* start, int count, I b) {
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

__device__ inline bool flagReached(const unsigned* f, unsigned gen) {
  return static_cast<int>(__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - gen) >= 0;
}

// Values handed between workgroups of one launch (the solved rows of a
// chain block) are written and read with device-coherent accesses, so
// publishing a block needs no cache write-back and waiting for one no cache
// invalidation.
__device__ inline void storeShared(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double loadShared(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void raiseError(int* err, int code) {
  __hip_atomic_store(err, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until another workgroup has published block f of this solve.  Thread 0
// polls with a bounded count; past the bound it raises the error word and
// stops waiting, so every workgroup still runs to its end.
__device__ inline void waitBlock(const unsigned* f, unsigned gen, int* err) {
  if (threadIdx.x == 0) {
    // one device-coherent load per poll; another workgroup's time-out and
    // the spin limit are looked at every 32 polls
    unsigned spins = 0;
    while (!flagReached(f, gen)) {
      if ((++spins & 31u) == 0) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (spins > kSpinLimit) {
          raiseError(err, 1);
          break;
        }
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

// Publish this workgroup's block: every wave waits until its device-coherent
// stores have completed, then one flag store releases the block.
__device__ inline void publishBlock(unsigned* f, unsigned gen) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(f, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// L(row, col) of a front with the (s+1, s) slot of a 2x2 pivot block read as
// zero: that slot stores d21, not a multiplier (multifrontal.hpp:618-628).
template <typename T>
__device__ inline double lval(const T* F, int m, const int* pk, int row, int col) {
  const double v = F[row + static_cast<size_t>(col) * m];
  return (row == col + 1 && pk[col] == 2) ? 0.0 : v;
}

// Elements that may lie outside their block, loaded unconditionally: outside,
// the block's first element (a valid address) is read and zero returned.  A
// load under a branch completes before the branch joins, so conditional
// loads would wait for one another; these are issued back to back.
template <typename T>
__device__ inline T ldOr0(const T* base, int64_t off, bool ok) {
  const T v = base[ok ? off : 0];
  return ok ? v : T(0);
}
template <typename T>
__device__ inline double lvalOr0(const T* F, int m, const int* pk, int row, int col, bool ok) {
  const int r = ok ? row : 0, c = ok ? col : 0;
  const double v = F[r + static_cast<size_t>(c) * m];
  const int k = pk[c];
  return (ok && !(r == c + 1 && k == 2)) ? v : 0.0;
}

// The solve's tile products on the matrix cores, in the storage type of the
// factor: v_mfma_f64_16x16x4f64 for FP64 fronts, v_mfma_f32_16x16x4f32 for
// FP32 fronts (operands rounded to FP32).  A[i][k] comes from lane i + 16k,
// B[k][j] from lane j + 16k; row(lane, r) is the row i of the result held in
// register r of a lane (its column j is lane % 16).
template <typename T>
struct SweepMfma;
template <>
struct SweepMfma<double> {
  using E = double;
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <>
struct SweepMfma<float> {
  using E = float;
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// FP64 sums of a run of tile products: FP64 operands accumulate in the MFMA
// result registers directly; the FP32 products of one tile are summed in
// FP32 and then added in FP64.  begin() gives the tile's accumulators,
// end() closes the tile.
template <typename T, int N>
struct TileAcc;
template <int N>
struct TileAcc<double, N> {
  typename SweepMfma<double>::v4 a[N];
  __device__ void zero() {
#pragma unroll
    for (int n = 0; n < N; ++n) a[n] = typename SweepMfma<double>::v4{0, 0, 0, 0};
  }
  __device__ typename SweepMfma<double>::v4* begin() { return a; }
  __device__ void end() {}
  __device__ double get(int n, int r) const { return a[n][r]; }
};
template <int N>
struct TileAcc<float, N> {
  double a[N][4];
  typename SweepMfma<float>::v4 p[N];
  __device__ void zero() {
#pragma unroll
    for (int n = 0; n < N; ++n)
#pragma unroll
      for (int r = 0; r < 4; ++r) a[n][r] = 0.0;
  }
  __device__ typename SweepMfma<float>::v4* begin() {
#pragma unroll
    for (int n = 0; n < N; ++n) p[n] = typename SweepMfma<float>::v4{0, 0, 0, 0};
    return p;
  }
  __device__ void end() {
#pragma unroll
    for (int n = 0; n < N; ++n)
#pragma unroll
      for (int r = 0; r < 4; ++r) a[n][r] += static_cast<double>(p[n][r]);
  }
  __device__ double get(int n, int r) const { return a[n][r]; }
};

// LDS of one own-row block step: the L operand tile, the right-hand-side
// operand tile (reused for the block's rows in FP64 once the products are
// done), the inverse of its unit diagonal block and its pivot data.
template <typename E>
struct BlockLds {
  E Lt[SB][SB + 1];
  union {
    E yt[SB][YW];       // the right-hand-side operand tile of the products
    double ys[SB][YW];  // then the block's own rows (after the products)
  };
  double Ld[SB][SB + 1];
  double dA[SB + 1], dB[SB + 1];  // d_diag, d_off of positions r0 - 1 .. r0 + SB - 1
  int pkA[SB + 1], rixA[SB + 1];  // their pivot kinds and final-order rows
};

// Work block rows of other blocks: read back by the workgroup that wrote
// them (kShared false: ordered by its barriers) or published by another
// workgroup of the launch (kShared true: device-coherent).
template <bool kShared>
__device__ inline double readRow(const double* p) {
  if constexpr (kShared)
    return loadShared(p);
  else
    return *p;
}
template <bool kShared>
__device__ inline void writeRow(double* p, double v) {
  if constexpr (kShared)
    storeShared(p, v);
  else
    *p = v;
}

// Forward step of own block b of front s for right-hand sides rg .. rg + gw
// (solve.hpp:349-361, :434-455): y_b = Ltop(b, b)^{-1} (y0_b - sum_{j<b}
// L(b, j) y_j), y0_b in the block's work rows (from the gather), y_b written
// over it, and z = D^{-1} y of the pivot blocks that end in this block into
// Z.  With kShared the y_j come from other workgroups, waited for by flag
// (on the first right-hand-side group only), and y_b is published when done.
template <typename T, bool kShared>
__device__ void fwdOwnBlock(BlockLds<typename SweepMfma<T>::E>& S, const DevFronts& fd, int s, int b, int rg,
                            int gw, int nrhs, double* __restrict__ Z, double* Ds, const double* __restrict__ Mi,
                            unsigned* fl, unsigned gen, int* err, bool publish) {
  using M = SweepMfma<T>;
  using E = typename M::E;
  using MD = SweepMfma<double>;
  constexpr int NT = 256;
  const int m = fd.m[s];
  const int nf = fd.nF[s];
  const int64_t io = fd.idxOff[s];
  const int* pk = fd.pkind + io;
  const T* F = static_cast<const T*>(fd.F[s]);
  const int r0 = b * SB, nr = min(SB, nf - r0);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // wave wv owns output rows tr .. tr + 15 and right-hand sides tc .. tc + 31
  const int tr = (wv & 1) * 16, tc = (wv >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  for (int e = tid; e < SB * SB; e += NT) S.Ld[e / SB][e % SB] = Mi[e];
  for (int r = tid; r <= SB; r += NT) {
    const int p = r0 - 1 + r;
    const bool ok = p >= 0 && p < r0 + nr;
    S.dA[r] = ok ? fd.dDiag[io + p] : 0.0;
    S.dB[r] = ok ? fd.dOff[io + p] : 0.0;
    S.pkA[r] = ok ? pk[p] : 0;
    S.rixA[r] = ok ? fd.rowIdx[io + p] : 0;
  }
  double y0p[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int r = tr + M::row(lane, rr), c = tc + 16 * t + lr;
      y0p[t][rr] = ldOr0(static_cast<const double*>(Ds), static_cast<int64_t>(r0 + r) * nrhs + rg + c, r < nr && c < gw);
    }
  // L(b, j) tiles four elements per thread, loaded one tile ahead
  E lpre[4];
  auto loadL = [&](int j) {
    const int c0 = j * SB, cn = min(SB, nf - c0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + u * NT;
      const int r = e % SB, qq = e / SB;
      lpre[u] = static_cast<E>(lvalOr0(F, m, pk, r0 + r, c0 + qq, r < nr && qq < cn));
    }
  };
  double acc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};  // FP64 sums of the per-tile products
  if (b > 0) loadL(0);
  for (int j = 0; j < b; ++j) {
    if (kShared && rg == 0) waitBlock(fl + j, gen, err);
    const int c0 = j * SB, cn = min(SB, nf - c0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + u * NT;
      S.Lt[e % SB][e / SB] = lpre[u];
    }
    for (int e = tid; e < SB * 64; e += NT) {
      const int c = e & 63, qq = e >> 6;
      const bool ok = qq < cn && c < gw;
      const double v = readRow<kShared>(Ds + (ok ? static_cast<int64_t>(c0 + qq) * nrhs + rg + c : 0));
      S.yt[qq][c] = ok ? static_cast<E>(v) : E(0);
    }
    if (j + 1 < b) loadL(j + 1);
    __syncthreads();
    typename M::v4 part[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int kk = 0; kk < SB; kk += 4) {
      const E av = S.Lt[tr + lr][kk + lk];
      part[0] = M::mma(av, S.yt[kk + lk][tc + lr], part[0]);
      part[1] = M::mma(av, S.yt[kk + lk][tc + 16 + lr], part[1]);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) acc[t][rr] += static_cast<double>(part[t][rr]);
    __syncthreads();
  }
  // y_b = Ltop(b, b)^{-1} (y0_b - sum_j L(b, j) y_j), the inverse applied as
  // an FP64 tile product (rows past nr are zero so they add nothing)
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int r = tr + M::row(lane, rr), c = tc + 16 * t + lr;
      S.ys[r][c] = (r < nr && c < gw) ? y0p[t][rr] - acc[t][rr] : 0.0;
    }
  __syncthreads();
  {
    typename MD::v4 yv[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int kk = 0; kk < SB; kk += 4) {
      const double av = S.Ld[tr + lr][kk + lk];
      yv[0] = MD::mma(av, S.ys[kk + lk][tc + lr], yv[0]);
      yv[1] = MD::mma(av, S.ys[kk + lk][tc + 16 + lr], yv[1]);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int r = tr + MD::row(lane, rr), c = tc + 16 * t + lr;
        if (r < nr && c < gw) {
          S.ys[r][c] = yv[t][rr];
          writeRow<kShared>(Ds + static_cast<int64_t>(r0 + r) * nrhs + rg + c, yv[t][rr]);
        }
      }
  }
  __syncthreads();
  // the later blocks need y only: release it before forming z
  if (kShared && publish) publishBlock(fl + b, gen);
  // z = D^{-1} y for the pivot blocks that end in this block (slot r + 1 of
  // the staged pivot data is position r0 + r)
  for (int e = tid; e < nr * gw; e += NT) {
    const int c = e % gw, r = e / gw;
    const int p = r0 + r;
    const int64_t cc = rg + c;
    const int kind = S.pkA[r + 1];
    if (kind == 1) {
      Z[static_cast<int64_t>(S.rixA[r + 1]) * nrhs + cc] = S.ys[r][c] / S.dA[r + 1];
    } else if (kind == 2 && r + 1 < nr) {
      const double d11 = S.dA[r + 1], d21 = S.dB[r + 1], d22 = S.dA[r + 2];
      const double det = d11 * d22 - d21 * d21;
      const double y0 = S.ys[r][c], y1 = S.ys[r + 1][c];
      Z[static_cast<int64_t>(S.rixA[r + 1]) * nrhs + cc] = (d22 * y0 - d21 * y1) / det;
      Z[static_cast<int64_t>(S.rixA[r + 2]) * nrhs + cc] = (-d21 * y0 + d11 * y1) / det;
    } else if (kind == 0 && r == 0 && p > 0) {
      // second row of a 2x2 block that starts in the previous block
      const double d11 = S.dA[0], d21 = S.dB[0], d22 = S.dA[1];
      const double det = d11 * d22 - d21 * d21;
      const double y0 = readRow<kShared>(Ds + static_cast<int64_t>(p - 1) * nrhs + cc), y1 = S.ys[0][c];
      Z[static_cast<int64_t>(S.rixA[0]) * nrhs + cc] = (d22 * y0 - d21 * y1) / det;
      Z[static_cast<int64_t>(S.rixA[1]) * nrhs + cc] = (-d21 * y0 + d11 * y1) / det;
    }
  }
  __syncthreads();
}

// Backward step of own block w of front s (solve.hpp:364-383): x_w =
// Ltop(w, w)^{-T} (P_w - sum_{j>w} L(j, w)^T x_j), P_w (z less the
// ancestors' part) in the block's work rows, x_w written over it and into X.
template <typename T, bool kShared>
__device__ void bwdOwnBlock(BlockLds<typename SweepMfma<T>::E>& S, const DevFronts& fd, int s, int w, int rg,
                            int gw, int nrhs, double* __restrict__ X, double* Ds, const double* __restrict__ Dl,
                            const AncPart pa, const double* __restrict__ Mi, unsigned* fl, unsigned gen, int* err,
                            bool publish) {
  using M = SweepMfma<T>;
  using E = typename M::E;
  using MD = SweepMfma<double>;
  constexpr int NT = 256;
  const int m = fd.m[s];
  const int nf = fd.nF[s];
  const int nOwn = (nf + SB - 1) / SB;
  const int64_t io = fd.idxOff[s];
  const int* pk = fd.pkind + io;
  const T* F = static_cast<const T*>(fd.F[s]);
  const int r0 = w * SB, nr = min(SB, nf - r0);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // wave wv owns own columns tr .. tr + 15 and right-hand sides tc .. tc + 31
  const int tr = (wv & 1) * 16, tc = (wv >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  for (int e = tid; e < SB * SB; e += NT) S.Ld[e / SB][e % SB] = Mi[e];
  for (int q = tid; q < SB; q += NT) S.rixA[q] = (q < nr) ? fd.rowIdx[io + r0 + q] : 0;
  // P_w: chunk 0's value in the own rows less the other chunks' parts, in chunk order
  double p0[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int qq = tr + M::row(lane, rr), c = tc + 16 * t + lr;
      const bool ok = qq < nr && c < gw;
      double v = ldOr0(static_cast<const double*>(Ds), static_cast<int64_t>(r0 + qq) * nrhs + rg + c, ok);
      for (int k = 1; k < pa.split; ++k)
        v -= ldOr0(static_cast<const double*>(Dl),
                   (pa.off + static_cast<int64_t>(k - 1) * pa.stride + r0 + qq) * nrhs + rg + c, ok);
      p0[t][rr] = v;
    }
  E opre[4];
  auto loadO = [&](int j) {
    const int i0 = j * SB, in = min(SB, nf - i0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + u * NT;
      const int i = e % SB, qq = e / SB;
      opre[u] = static_cast<E>(lvalOr0(F, m, pk, i0 + i, r0 + qq, i < in && qq < nr));
    }
  };
  double acc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  if (nOwn - 1 > w) loadO(nOwn - 1);
  for (int j = nOwn - 1; j > w; --j) {
    if (kShared && rg == 0) waitBlock(fl + j, gen, err);
    const int i0 = j * SB, in = min(SB, nf - i0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + u * NT;
      S.Lt[e % SB][e / SB] = opre[u];
    }
    for (int e = tid; e < SB * 64; e += NT) {
      const int c = e & 63, i = e >> 6;
      const bool ok = i < in && c < gw;
      const double v = readRow<kShared>(Ds + (ok ? static_cast<int64_t>(i0 + i) * nrhs + rg + c : 0));
      S.yt[i][c] = ok ? static_cast<E>(v) : E(0);
    }
    if (j - 1 > w) loadO(j - 1);
    __syncthreads();
    typename M::v4 part[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int kk = 0; kk < SB; kk += 4) {
      const E av = S.Lt[kk + lk][tr + lr];
      part[0] = M::mma(av, S.yt[kk + lk][tc + lr], part[0]);
      part[1] = M::mma(av, S.yt[kk + lk][tc + 16 + lr], part[1]);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) acc[t][rr] += static_cast<double>(part[t][rr]);
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int qq = tr + M::row(lane, rr), c = tc + 16 * t + lr;
      S.ys[qq][c] = (qq < nr && c < gw) ? p0[t][rr] - acc[t][rr] : 0.0;
    }
  __syncthreads();
  {
    // x_w = L_ww^{-T} (...): the transposed inverse as an FP64 tile product
    typename MD::v4 xv[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int kk = 0; kk < SB; kk += 4) {
      const double av = S.Ld[kk + lk][tr + lr];
      xv[0] = MD::mma(av, S.ys[kk + lk][tc + lr], xv[0]);
      xv[1] = MD::mma(av, S.ys[kk + lk][tc + 16 + lr], xv[1]);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int qq = tr + MD::row(lane, rr), c = tc + 16 * t + lr;
        if (qq < nr && c < gw) {
          S.ys[qq][c] = xv[t][rr];
          writeRow<kShared>(Ds + static_cast<int64_t>(r0 + qq) * nrhs + rg + c, xv[t][rr]);
        }
      }
  }
  __syncthreads();
  // the earlier blocks need x only through the work block: release it, then write X
  if (kShared && publish) publishBlock(fl + w, gen);
  for (int e = tid; e < nr * gw; e += NT) {
    const int c = e % gw, r = e / gw;
    X[static_cast<int64_t>(S.rixA[r]) * nrhs + rg + c] = S.ys[r][c];
  }
  __syncthreads();
}

}  // namespace

// One increment per solve: the generation that the chain flags are compared
// with, kept on the device so that a recorded solve stays valid on replay.
__global__ void k_solve_begin(unsigned* gen) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *gen = *gen + 1u;
}

// Forward step 1 (solve.hpp:327-347) through a gather map, for every front
// of a level: front position p (pre-permutation position t = perm[p]) gets
// g = the sum of the children's delta entries that land on t, in the
// children's order (mapPtr over idxOff positions, mapSrc = (child, child
// position)) -- the sums of adding the children one after another into a
// zeroed block.  An own row p < nF gets y0 = Bp - g, the others g, the start
// of the delta handed to the parent.  info[k] describes the level's position
// k; 2^lpShift lanes take one, each a strided subset of the right-hand sides.
__global__ __launch_bounds__(256) void k_fwd_gather(const PosInfo* __restrict__ info, int64_t positions, int lpShift,
                                                     int nrhs, const double* __restrict__ Bp, double* __restrict__ Dl,
                                                     const int64_t* __restrict__ mapRow) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t item = g >> lpShift;
  const int lp = 1 << lpShift;
  const int l = static_cast<int>(g & (lp - 1));
  if (item >= positions) return;
  const PosInfo pi = info[item];
  double* dst = Dl + pi.dst * nrhs;
  const double* bp = pi.rix >= 0 ? Bp + static_cast<int64_t>(pi.rix) * nrhs : nullptr;
  // four right-hand sides per lane at a time: the children's rows are read
  // with four independent loads per entry
  for (int c0 = l; c0 < nrhs; c0 += 4 * lp) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int e = 0; e < pi.cnt; ++e) {
      const double* src = Dl + mapRow[pi.e0 + e] * nrhs;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (c0 + u * lp < nrhs) acc[u] += src[c0 + u * lp];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = c0 + u * lp;
      if (c < nrhs) dst[c] = bp ? bp[c] - acc[u] : acc[u];
    }
  }
}

// The gather's per-position records for the current factorization (the
// pivoting permutation and the row identities change with the values):
// position k = (front s, position p) of the solve's position list.
__global__ __launch_bounds__(256) void k_pos_info(const int2* __restrict__ pos, int64_t positions, DevFronts fd,
                                                   const int64_t* __restrict__ gOff,
                                                   const int64_t* __restrict__ mapPtr, PosInfo* __restrict__ info) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= positions) return;
  const int2 sp = pos[k];
  const int s = sp.x, p = sp.y;
  const int64_t io = fd.idxOff[s];
  PosInfo pi;
  pi.dst = gOff[s] + p;
  if (p < fd.m[s]) {
    const int t = fd.perm[io + p];
    pi.e0 = mapPtr[io + t];
    pi.cnt = static_cast<int>(mapPtr[io + t + 1] - pi.e0);
    pi.rix = p < fd.nF[s] ? fd.rowIdx[io + p] : -1;
  } else {
    pi.e0 = 0;
    pi.cnt = 0;
    pi.rix = -1;
  }
  info[k] = pi;
}

// Inverses of the unit lower diagonal blocks of L that the solve applies:
// block b of front s (own rows [32b, 32b + 32) of [0, nF)) gets
// M = L_bb^{-1}, row-major 32 x 32 at inv + invOff[s] + 1024 b (FP64; rows
// and columns past nF hold the identity).  One workgroup per block; thread c
// forms column c by forward substitution.  Applying M is the substitution
// reassociated (the inverted-diagonal-block form of the triangular solve).
template <typename T>
__global__ __launch_bounds__(64) void k_diag_inverse(const int* list, int count, const int* blockStart, DevFronts fd,
                                                     double* __restrict__ inv, const int64_t* __restrict__ invOff) {
  __shared__ double Ls[SB][SB + 1];
  __shared__ double Ms[SB][SB + 1];
  if (static_cast<int>(blockIdx.x) >= blockStart[count]) return;
  const int f = segmentOf<int>(blockStart, count, blockIdx.x);
  const int s = list[f];
  const int b = blockIdx.x - blockStart[f];
  const int m = fd.m[s];
  const int nf = fd.nF[s];
  if (b * SB >= nf) return;
  const int r0 = b * SB;
  const int nr = min(SB, nf - r0);
  const int64_t io = fd.idxOff[s];
  const int* pk = fd.pkind + io;
  const T* F = static_cast<const T*>(fd.F[s]);
  const int tid = threadIdx.x;
  for (int e = tid; e < SB * SB; e += 64) {
    const int r = e % SB, c = e / SB;
    Ls[r][c] = (r > c && r < nr) ? lval(F, m, pk, r0 + r, r0 + c) : 0.0;
  }
  __syncthreads();
  if (tid < SB) {
    const int c = tid;
    for (int r = 0; r < SB; ++r) Ms[r][c] = (r == c) ? 1.0 : 0.0;
    for (int r = c + 1; r < SB; ++r) {
      double a0 = 0.0, a1 = 0.0;
      int t = c;
      for (; t + 2 <= r; t += 2) {
        a0 = fma(Ls[r][t], Ms[t][c], a0);
        a1 = fma(Ls[r][t + 1], Ms[t + 1][c], a1);
      }
      if (t < r) a0 = fma(Ls[r][t], Ms[t][c], a0);
      Ms[r][c] = -(a0 + a1);
    }
  }
  __syncthreads();
  double* out = inv + invOff[s] + static_cast<int64_t>(b) * SB * SB;
  for (int e = tid; e < SB * SB; e += 64) out[e] = Ms[e / SB][e % SB];
}

// Forward own rows of fronts with a short own-row chain: one workgroup per
// front runs the blocks b = 0, 1, ... in order, reading the y_j it has
// written back from the work block (its own barriers order those stores and
// loads), so no block waits on another workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_own(const int* list, int count, DevFronts fd, int nrhs,
                                                 double* __restrict__ Z, double* __restrict__ Dl,
                                                 const int64_t* __restrict__ gOff, const double* __restrict__ inv,
                                                 const int64_t* __restrict__ invOff) {
  __shared__ BlockLds<typename SweepMfma<T>::E> S;
  if (static_cast<int>(blockIdx.x) >= count) return;
  const int s = list[blockIdx.x];
  const int nOwn = (fd.nF[s] + SB - 1) / SB;
  double* Ds = Dl + gOff[s] * nrhs;
  for (int rg = 0; rg < nrhs; rg += 64) {
    const int gw = min(64, nrhs - rg);
    for (int b = 0; b < nOwn; ++b)
      fwdOwnBlock<T, false>(S, fd, s, b, rg, gw, nrhs, Z, Ds, inv + invOff[s] + static_cast<int64_t>(b) * SB * SB,
                            nullptr, 0u, nullptr, false);
  }
}

// Forward own rows of fronts with a long own-row chain: one workgroup per
// own block; block b takes the products with the blocks j < b as they are
// published (flag per block, bounded wait), so the chain's products run on
// many workgroups at once.  A workgroup only waits for blocks of the same
// front with a lower index, which were dispatched before it.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_chain(const int* list, int count, const int* blockStart, DevFronts fd,
                                                   int nrhs, double* __restrict__ Z, double* __restrict__ Dl,
                                                   const int64_t* __restrict__ gOff, unsigned* flags,
                                                   const int64_t* __restrict__ flagOff, const unsigned* genp,
                                                   int* err, const double* __restrict__ inv,
                                                   const int64_t* __restrict__ invOff) {
  __shared__ BlockLds<typename SweepMfma<T>::E> S;
  if (static_cast<int>(blockIdx.x) >= blockStart[count]) return;
  const int f = segmentOf<int>(blockStart, count, blockIdx.x);
  const int s = list[f];
  const int b = blockIdx.x - blockStart[f];
  const int nOwn = (fd.nF[s] + SB - 1) / SB;
  if (b == 0 && threadIdx.x == 0 && nOwn > blockStart[f + 1] - blockStart[f]) raiseError(err, 2);
  if (b >= nOwn) return;
  const unsigned gen = *genp;
  double* Ds = Dl + gOff[s] * nrhs;
  for (int rg = 0; rg < nrhs; rg += 64) {
    const int gw = min(64, nrhs - rg);
    fwdOwnBlock<T, true>(S, fd, s, b, rg, gw, nrhs, Z, Ds, inv + invOff[s] + static_cast<int64_t>(b) * SB * SB,
                         flags + flagOff[s], gen, err, rg + 64 >= nrhs);
  }
}

// Rows [nF, m) of a front once its own rows are solved (solve.hpp:349-361):
// delta(i) = g(i) + L(i, 0:nF) y, g already in the work block.  A workgroup
// takes one 64-row chunk b (waves of 32 rows x 32 right-hand sides); the
// product runs over the own columns in 32-column tiles, the next tiles'
// operands loaded while the current ones are multiplied.
template <typename E>
struct DeltaLds {
  E Lt[DT][SB + 1];  // [row i][own column q]
  E yt[SB][YW];      // [own column q][right-hand side]
};
template <typename T>
__device__ void fwdDeltaChunk(DeltaLds<typename SweepMfma<T>::E>& S, const DevFronts& fd, int s, int b, int nrhs,
                              double* __restrict__ Ds) {
  using M = SweepMfma<T>;
  using E = typename M::E;
  constexpr int NT = 256;
  const int m = fd.m[s];
  const int nf = fd.nF[s];
  const int i0 = nf + b * DT, ni = min(DT, m - i0);
  const T* F = static_cast<const T*>(fd.F[s]);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  for (int rg = 0; rg < nrhs; rg += 64) {
    const int gw = min(64, nrhs - rg);
    E lpre[8], ypre[8];
    auto loadT = [&](int k0) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = tid + u * NT;
        const int i = e & (DT - 1), q = e / DT;
        lpre[u] = static_cast<E>(ldOr0(F, (i0 + i) + static_cast<int64_t>(k0 + q) * m, i < ni && k0 + q < nf));
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = tid + u * NT;
        const int c = e & 63, q = e >> 6;
        ypre[u] = static_cast<E>(ldOr0(static_cast<const double*>(Ds), static_cast<int64_t>(k0 + q) * nrhs + rg + c,
                                        k0 + q < nf && c < gw));
      }
    };
    TileAcc<T, 4> acc;  // block (bi, bj) at 2 bi + bj
    acc.zero();
    if (nf > 0) loadT(0);
    for (int k0 = 0; k0 < nf; k0 += SB) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = tid + u * NT;
        S.Lt[e & (DT - 1)][e / DT] = lpre[u];
        S.yt[e >> 6][e & 63] = ypre[u];
      }
      if (k0 + SB < nf) loadT(k0 + SB);
      __syncthreads();
      typename M::v4* part = acc.begin();
#pragma unroll 2
      for (int kk = 0; kk < SB; kk += 4) {
        const E a0 = S.Lt[wr + lr][kk + lk], a1 = S.Lt[wr + 16 + lr][kk + lk];
        const E b0 = S.yt[kk + lk][wc + lr], b1 = S.yt[kk + lk][wc + 16 + lr];
        part[0] = M::mma(a0, b0, part[0]);
        part[1] = M::mma(a0, b1, part[1]);
        part[2] = M::mma(a1, b0, part[2]);
        part[3] = M::mma(a1, b1, part[3]);
      }
      acc.end();
      __syncthreads();
    }
    // g of the chunk's rows, all loaded before any is used
    double g[2][2][4];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int r = wr + 16 * bi + M::row(lane, rr), c = wc + 16 * bj + lr;
          g[bi][bj][rr] = ldOr0(static_cast<const double*>(Ds), static_cast<int64_t>(i0 + r) * nrhs + rg + c, r < ni && c < gw);
        }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int r = wr + 16 * bi + M::row(lane, rr), c = wc + 16 * bj + lr;
          if (r < ni && c < gw) Ds[static_cast<int64_t>(i0 + r) * nrhs + rg + c] = g[bi][bj][rr] + acc.get(2 * bi + bj, rr);
        }
  }
}


template <typename T>
__global__ __launch_bounds__(256) void k_fwd_delta(const int* list, int count, const int* blockStart, DevFronts fd,
                                                   int nrhs, double* __restrict__ Dl, const int64_t* __restrict__ gOff,
                                                   int* err) {
  __shared__ DeltaLds<typename SweepMfma<T>::E> S;
  if (static_cast<int>(blockIdx.x) >= blockStart[count]) return;
  const int f = segmentOf<int>(blockStart, count, blockIdx.x);
  const int s = list[f];
  const int b = blockIdx.x - blockStart[f];
  const int nD = (fd.m[s] - fd.nF[s] + DT - 1) / DT;
  if (b == 0 && threadIdx.x == 0 && nD > blockStart[f + 1] - blockStart[f]) raiseError(err, 2);
  if (b >= nD) return;
  fwdDeltaChunk<T>(S, fd, s, b, nrhs, Dl + gOff[s] * nrhs);
}

// Backward step, the ancestors' part (solve.hpp:364-383): for own block w of
// a front, P_w = z_w - L(nF:m, w)^T x(nF:m), the ancestors' x final from the
// levels above and read from X through the front's row indices (one
// contiguous row each; the indices of a tile are read two tiles ahead, as
// wave-uniform values).  The ancestors' rows are split in `split` chunks of
// whole 64-row tiles: workgroup (w, k) takes chunk k of own block w, chunk 0
// writing z_w less its part into the block's work rows, chunk k >= 1 its part
// into the front's partial rows (AncPart).
template <typename E>
struct AncLds {
  E Lt[DT][SB + 1];  // [ancestor row i][own column q]
  E xt[DT][YW];      // [ancestor row i][right-hand side]
};
template <typename T>
__device__ void bwdAncChunk(AncLds<typename SweepMfma<T>::E>& S, const DevFronts& fd, int s, int w, int k,
                            const AncPart pa, int nrhs, const double* __restrict__ Z, const double* __restrict__ X,
                            double* __restrict__ Dl, double* __restrict__ Ds) {
  using M = SweepMfma<T>;
  using E = typename M::E;
  constexpr int NT = 256;
  const int m = fd.m[s];
  const int nf = fd.nF[s];
  const int r0 = w * SB, nr = min(SB, nf - r0);
  const int chunk = ((m - nf + pa.split - 1) / pa.split + DT - 1) / DT * DT;
  const int a0 = nf + k * chunk, a1 = min(m, a0 + chunk);
  const int64_t io = fd.idxOff[s];
  const T* F = static_cast<const T*>(fd.F[s]);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wu = __builtin_amdgcn_readfirstlane(wv);
  const int tr = (wv & 1) * 16, tc = (wv >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  for (int rg = 0; rg < nrhs; rg += 64) {
    const int gw = min(64, nrhs - rg);
    // staging: thread (wave wu, lane) loads L rows i0 + lane of own columns
    // wu + 4u, and x rows i0 + wu + 4u (right-hand side rg + lane)
    E lpre[8], xpre[16];
    int64_t rcur[16], rnext[16];
    auto rows = [&](int i0, int64_t* out) {
      // loaded unconditionally (a1 > i0 here) and held as loaded, so the
      // sixteen loads are in flight together
      int rr[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) rr[u] = fd.rowIdx[io + min(i0 + wu + 4 * u, a1 - 1)];
#pragma unroll
      for (int u = 0; u < 16; ++u) asm volatile("" : "+v"(rr[u]));
#pragma unroll
      for (int u = 0; u < 16; ++u) out[u] = (i0 + wu + 4 * u < a1) ? static_cast<int64_t>(rr[u]) * nrhs : -1;
    };
    auto loadT = [&](int i0, const int64_t* xr) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = wv + 4 * u;
        lpre[u] = static_cast<E>(ldOr0(F, (i0 + lane) + static_cast<int64_t>(r0 + q) * m, i0 + lane < a1 && q < nr));
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) xpre[u] = static_cast<E>(ldOr0(X, xr[u] + rg + lane, xr[u] >= 0 && lane < gw));
    };
    double acc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    if (a0 < a1) {
      rows(a0, rcur);
      loadT(a0, rcur);
      if (a0 + DT < a1) rows(a0 + DT, rnext);
    }
    for (int i0 = a0; i0 < a1; i0 += DT) {
#pragma unroll
      for (int u = 0; u < 8; ++u) S.Lt[lane][wv + 4 * u] = lpre[u];
#pragma unroll
      for (int u = 0; u < 16; ++u) S.xt[wv + 4 * u][lane] = xpre[u];
      if (i0 + DT < a1) {
        loadT(i0 + DT, rnext);
#pragma unroll
        for (int u = 0; u < 16; ++u) rcur[u] = rnext[u];
        if (i0 + 2 * DT < a1) rows(i0 + 2 * DT, rnext);
      }
      __syncthreads();
      typename M::v4 part[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll 4
      for (int kk = 0; kk < DT; kk += 4) {
        const E av = S.Lt[kk + lk][tr + lr];
        part[0] = M::mma(av, S.xt[kk + lk][tc + lr], part[0]);
        part[1] = M::mma(av, S.xt[kk + lk][tc + 16 + lr], part[1]);
      }
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) acc[t][rr] += static_cast<double>(part[t][rr]);
      __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int qq = tr + M::row(lane, rr), c = tc + 16 * t + lr;
        const bool ok = qq < nr && c < gw;
        const int zr = fd.rowIdx[io + r0 + (ok ? qq : 0)];
        const double z = (k == 0) ? ldOr0(Z, static_cast<int64_t>(zr) * nrhs + rg + c, ok) : 0.0;
        if (ok) {
          if (k == 0) {
            Ds[static_cast<int64_t>(r0 + qq) * nrhs + rg + c] = z - acc[t][rr];
          } else {
            Dl[(pa.off + static_cast<int64_t>(k - 1) * pa.stride + r0 + qq) * nrhs + rg + c] = acc[t][rr];
          }
        }
      }
  }
}


template <typename T>
__global__ __launch_bounds__(256) void k_bwd_anc(const int* list, int count, const int* blockStart, DevFronts fd,
                                                 int nrhs, const double* __restrict__ Z, const double* __restrict__ X,
                                                 double* __restrict__ Dl, const int64_t* __restrict__ gOff,
                                                 const AncPart* __restrict__ parts, int* err) {
  __shared__ AncLds<typename SweepMfma<T>::E> S;
  if (static_cast<int>(blockIdx.x) >= blockStart[count]) return;
  const int f = segmentOf<int>(blockStart, count, blockIdx.x);
  const int s = list[f];
  const int bb = blockIdx.x - blockStart[f];
  const AncPart pa = parts[s];
  const int alloc = (blockStart[f + 1] - blockStart[f]) / pa.split;  // own blocks planned
  const int w = bb / pa.split, k = bb - w * pa.split;
  const int nOwn = (fd.nF[s] + SB - 1) / SB;
  if (bb == 0 && threadIdx.x == 0 && nOwn > alloc) raiseError(err, 2);
  if (w >= nOwn) return;
  bwdAncChunk<T>(S, fd, s, w, k, pa, nrhs, Z, X, Dl, Dl + gOff[s] * nrhs);
}

// Backward own rows of fronts with a short own-row chain: one workgroup per
// front, blocks w = nOwn - 1 .. 0 in order, the x_j read back from the work
// block this workgroup wrote.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_own(const int* list, int count, DevFronts fd, int nrhs,
                                                 double* __restrict__ X, double* __restrict__ Dl,
                                                 const int64_t* __restrict__ gOff, const AncPart* __restrict__ parts,
                                                 const double* __restrict__ inv, const int64_t* __restrict__ invOff) {
  __shared__ BlockLds<typename SweepMfma<T>::E> S;
  if (static_cast<int>(blockIdx.x) >= count) return;
  const int s = list[blockIdx.x];
  const int nOwn = (fd.nF[s] + SB - 1) / SB;
  double* Ds = Dl + gOff[s] * nrhs;
  for (int rg = 0; rg < nrhs; rg += 64) {
    const int gw = min(64, nrhs - rg);
    for (int w = nOwn - 1; w >= 0; --w)
      bwdOwnBlock<T, false>(S, fd, s, w, rg, gw, nrhs, X, Ds, Dl, parts[s],
                            inv + invOff[s] + static_cast<int64_t>(w) * SB * SB, nullptr, 0u, nullptr, false);
  }
}

// Backward own rows of fronts with a long own-row chain: one workgroup per
// own block, taken from the last block to the first; block w waits for the
// blocks j > w of its front (dispatched before it).
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_chain(const int* list, int count, const int* blockStart, DevFronts fd,
                                                   int nrhs, double* __restrict__ X, double* __restrict__ Dl,
                                                   const int64_t* __restrict__ gOff,
                                                   const AncPart* __restrict__ parts, unsigned* flags,
                                                   const int64_t* __restrict__ flagOff, const unsigned* genp,
                                                   int* err, const double* __restrict__ inv,
                                                   const int64_t* __restrict__ invOff) {
  __shared__ BlockLds<typename SweepMfma<T>::E> S;
  if (static_cast<int>(blockIdx.x) >= blockStart[count]) return;
  const int f = segmentOf<int>(blockStart, count, blockIdx.x);
  const int s = list[f];
  const int b = blockIdx.x - blockStart[f];
  const int nOwn = (fd.nF[s] + SB - 1) / SB;
  if (b == 0 && threadIdx.x == 0 && nOwn > blockStart[f + 1] - blockStart[f]) raiseError(err, 2);
  if (b >= nOwn) return;
  const int w = nOwn - 1 - b;
  const unsigned gen = *genp;
  double* Ds = Dl + gOff[s] * nrhs;
  for (int rg = 0; rg < nrhs; rg += 64) {
    const int gw = min(64, nrhs - rg);
    bwdOwnBlock<T, true>(S, fd, s, w, rg, gw, nrhs, X, Ds, Dl, parts[s],
                         inv + invOff[s] + static_cast<int64_t>(w) * SB * SB, flags + flagOff[s], gen, err,
                         rg + 64 >= nrhs);
  }
}

// A double of lane l (wave-uniform l) to every lane.
__device__ inline double readLane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Fronts of at most 64 rows: one wave per front, lane i owning row i, the
// right-hand sides in groups of 8 held in registers.  The lane's row of L is
// read once into registers (the loops over the front's columns are unrolled,
// their bounds wave-uniform branches), and the solved value of row j reaches
// the other lanes by a lane read.  Forward: the unit-lower solve of the own
// rows, D^{-1} into Z, and the delta of the remaining rows
// (solve.hpp:303-362, :434-455) in one pass.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_tiny(const int* list, int count, DevFronts fd, int nrhs,
                                                   double* __restrict__ Z, double* __restrict__ Dl,
                                                   const int64_t* __restrict__ gOff) {
  const int f = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (f >= count) return;
  const int s = list[f];
  const int lane = threadIdx.x & 63;
  const int m = fd.m[s];
  const int nf = fd.nF[s];
  const int64_t io = fd.idxOff[s];
  const int* pk = fd.pkind + io;
  const T* F = static_cast<const T*>(fd.F[s]);
  double* Dr = Dl + (gOff[s] + lane) * nrhs;  // this lane's work row
  const bool row = lane < m;
  const bool own = lane < nf;
  // own rows subtract L(i, :) y, the others add it
  double lr[64];
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    const double l = lvalOr0(F, m, pk, lane, j, row && j < nf && j < lane);
    lr[j] = own ? -l : l;
  }
  // keep the row in registers: loaded once, not again inside the loops below
#pragma unroll
  for (int j = 0; j < 64; ++j) asm volatile("" : "+v"(lr[j]));
  const int kind = own ? pk[lane] : 0;
  const int64_t prow = own ? fd.rowIdx[io + lane] : 0;
  for (int cg = 0; cg < nrhs; cg += 8) {
    double v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = ldOr0(static_cast<const double*>(Dl), (gOff[s] + lane) * nrhs + cg + c, row && cg + c < nrhs);
#pragma unroll
    for (int j = 0; j < 64; ++j) {
      if (j < nf) {
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = fma(lr[j], readLane(v[c], j), v[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (cg + c >= nrhs) break;
      const int cc = cg + c;
      const double up = __shfl(v[c], min(lane + 1, 63), 64);
      const double dn = __shfl(v[c], max(lane - 1, 0), 64);
      if (row) Dr[cc] = v[c];
      if (own) {
        double z;
        if (kind == 1) {
          z = v[c] / fd.dDiag[io + lane];
        } else if (kind == 2) {
          const double d11 = fd.dDiag[io + lane], d21 = fd.dOff[io + lane], d22 = fd.dDiag[io + lane + 1];
          const double det = d11 * d22 - d21 * d21;
          z = (d22 * v[c] - d21 * up) / det;
        } else {
          const double d11 = fd.dDiag[io + lane - 1], d21 = fd.dOff[io + lane - 1], d22 = fd.dDiag[io + lane];
          const double det = d11 * d22 - d21 * d21;
          z = (-d21 * dn + d11 * v[c]) / det;
        }
        Z[prow * nrhs + cc] = z;
      }
    }
  }
}

// Backward pass of fronts of at most 64 rows, the same layout with the
// lane's column of L in registers: x_own = z - L(anc, :)^T x_anc, then the
// unit upper solve with L_top^T (solve.hpp:364-383).  Lane t also holds the
// ancestors' x of the front's row nF + t.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_tiny(const int* list, int count, DevFronts fd, int nrhs,
                                                   const double* __restrict__ Z, double* __restrict__ X) {
  const int f = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (f >= count) return;
  const int s = list[f];
  const int lane = threadIdx.x & 63;
  const int m = fd.m[s];
  const int nf = fd.nF[s];
  const int64_t io = fd.idxOff[s];
  const int* pk = fd.pkind + io;
  const T* F = static_cast<const T*>(fd.F[s]);
  const bool own = lane < nf;
  // column `lane` of L below the diagonal, rows [lane + 1, m)
  double lc[64];
#pragma unroll
  for (int a = 0; a < 64; ++a) lc[a] = lvalOr0(F, m, pk, a, lane, own && a > lane && a < m);
  // keep the column in registers: loaded once, not again inside the loops below
#pragma unroll
  for (int a = 0; a < 64; ++a) asm volatile("" : "+v"(lc[a]));
  const int64_t prow = own ? fd.rowIdx[io + lane] : 0;
  const int ncb = m - nf;
  const int64_t crow = (lane < ncb) ? fd.rowIdx[io + nf + lane] : 0;
  for (int cg = 0; cg < nrhs; cg += 8) {
    double x[8], xa[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      x[c] = ldOr0(Z, prow * nrhs + cg + c, own && cg + c < nrhs);
      xa[c] = ldOr0(static_cast<const double*>(X), crow * nrhs + cg + c, lane < ncb && cg + c < nrhs);
    }
#pragma unroll
    for (int a = 0; a < 64; ++a) {
      if (a < nf) continue;
      if (a >= m) break;
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] = fma(-lc[a], readLane(xa[c], a - nf), x[c]);
    }
#pragma unroll
    for (int j = 63; j > 0; --j) {
      if (j >= nf) continue;
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] = fma(-lc[j], readLane(x[c], j), x[c]);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (own && cg + c < nrhs) X[prow * nrhs + cg + c] = x[c];
  }
}

// ---------------------------------------------------------------------------
// Refinement support (refine.hpp:69-144).
// ---------------------------------------------------------------------------

// |v| for the max reductions of the norms, with NaN mapped to +inf: fmax
// drops a NaN operand, and a NaN residual or solution must not read as a
// small norm (refinement would then stop as converged on garbage).
__device__ inline double normAbs(double v) { return v != v ? static_cast<double>(INFINITY) : fabs(v); }

// Max of v over a 256-thread workgroup into *bits (bits of a non-negative
// double) with one atomic per workgroup; every thread of the group calls it.
__device__ inline void groupAtomicMax(double v, unsigned long long* bits) {
  __shared__ double wmax[4];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double mx = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    if (mx > 0.0) atomicMax(bits, static_cast<unsigned long long>(__double_as_longlong(mx)));
  }
}

// Symmetric CSR values in row order, aSym[t] = aVals[valSrc[t]]: the full
// matrix built from the caller's lower-triangle CSC values
// (selfadjointView<Lower> semantics).
__global__ void k_gather_vals(int64_t nnz, const int64_t* __restrict__ valSrc, const double* __restrict__ aVals,
                              double* __restrict__ aSym) {
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < nnz;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    aSym[t] = aVals[valSrc[t]];
}

// R = B - A X for column-major n x nrhs blocks (leading dimension n), A as
// symmetric CSR (both triangles).  One thread per row; the right-hand sides
// go in chunks of kResChunk so the row's indices and values are read once
// per chunk rather than once per column.  max |R| goes to *rmaxBits when
// given (the slot holds the bits of a non-negative double).
constexpr int kResChunk = 16;
__global__ __launch_bounds__(256) void k_residual(int n, int nrhs, const int64_t* __restrict__ rowPtr,
                                                  const int* __restrict__ colIdx, const double* __restrict__ aSym,
                                                  const double* __restrict__ B, const double* __restrict__ X,
                                                  double* __restrict__ R, unsigned long long* rmaxBits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double mx = 0.0;
  if (i < n) {
    const int64_t t0 = rowPtr[i], t1 = rowPtr[i + 1];
    for (int c0 = 0; c0 < nrhs; c0 += kResChunk) {
      const int rc = min(kResChunk, nrhs - c0);
      const double* x = X + static_cast<int64_t>(c0) * n;
      double acc[kResChunk];
#pragma unroll
      for (int u = 0; u < kResChunk; ++u) acc[u] = 0.0;
      if (rc == kResChunk) {
        for (int64_t t = t0; t < t1; ++t) {
          const double a = aSym[t];
          const double* xc = x + colIdx[t];
#pragma unroll
          for (int u = 0; u < kResChunk; ++u) acc[u] = fma(a, xc[static_cast<int64_t>(u) * n], acc[u]);
        }
      } else {
        for (int64_t t = t0; t < t1; ++t) {
          const double a = aSym[t];
          const double* xc = x + colIdx[t];
#pragma unroll
          for (int u = 0; u < kResChunk; ++u)
            if (u < rc) acc[u] = fma(a, xc[static_cast<int64_t>(u) * n], acc[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < kResChunk; ++u)
        if (u < rc) {
          const int64_t e = static_cast<int64_t>(c0 + u) * n + i;
          const double r = B[e] - acc[u];
          R[e] = r;
          mx = fmax(mx, normAbs(r));
        }
    }
  }
  if (rmaxBits) groupAtomicMax(mx, rmaxBits);
}

// Many right-hand sides (nrhs >= kResWide).  Column-major X makes the
// product's gather read one cache line per right-hand side for every matrix
// entry, so X is first transposed into Xt (row i's right-hand sides
// contiguous, k_transpose_rows) and the residual takes a wave per row with a
// lane per right-hand side: each entry's gather becomes one contiguous
// nrhs-word read.  A workgroup covers 64 rows x 64 right-hand sides and
// moves B in and R out through LDS so both go to memory in 512-byte column
// segments.
constexpr int kResWide = 16;
__global__ __launch_bounds__(256) void k_transpose_rows(int n, int nrhs, const double* __restrict__ X, int64_t ldx,
                                                        double* __restrict__ Xt) {
  __shared__ double tile[64][65];
  const int i0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 4 rows of 64 lanes
  for (int cc = ty; cc < 64; cc += 4) {  // read columns c0 + cc, rows i0 + tx (coalesced)
    const int c = c0 + cc, i = i0 + tx;
    tile[cc][tx] = (c < nrhs && i < n) ? X[i + static_cast<int64_t>(c) * ldx] : 0.0;
  }
  __syncthreads();
  for (int ii = ty; ii < 64; ii += 4) {  // write rows i0 + ii, right-hand sides c0 + tx (coalesced)
    const int i = i0 + ii, c = c0 + tx;
    if (i < n && c < nrhs) Xt[static_cast<int64_t>(i) * nrhs + c] = tile[tx][ii];
  }
}

__global__ __launch_bounds__(256) void k_residual_wide(int n, int nrhs, const int64_t* __restrict__ rowPtr,
                                                       const int* __restrict__ colIdx,
                                                       const double* __restrict__ aSym,
                                                       const double* __restrict__ B, const double* __restrict__ Xt,
                                                       double* __restrict__ R, unsigned long long* rmaxBits) {
  __shared__ double tile[64][65];  // [right-hand side][row]
  const int i0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int cc = w; cc < 64; cc += 4) {  // B tile in, column segments
    const int c = c0 + cc, i = i0 + lane;
    tile[cc][lane] = (c < nrhs && i < n) ? B[i + static_cast<int64_t>(c) * n] : 0.0;
  }
  __syncthreads();
  const int c = c0 + lane;
  const int cl = c < nrhs ? c : nrhs - 1;  // lanes past nrhs read a valid word and drop it
  for (int rr = w; rr < 64; rr += 4) {  // wave w: rows i0 + w, + 4, ...
    const int i = i0 + rr;
    if (i >= n) break;
    double acc = 0.0;
    const int64_t t0 = rowPtr[i], t1 = rowPtr[i + 1];
    // the row's entries are read once per 64 (a lane each) and taken lane by lane into scalar registers;
    // the gathers of X rows for 16 consecutive entries all go out before their products (entries past the
    // row's end gather the row's last column again and are dropped), the products in entry order
    for (int64_t base = t0; base < t1; base += 64) {
      const int cnt = static_cast<int>(min<int64_t>(64, t1 - base));
      const int64_t tl = base + min(lane, cnt - 1);
      const int colL = colIdx[tl];
      const double aL = aSym[tl];
      for (int j0 = 0; j0 < cnt; j0 += 16) {
        double xv[16], av[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int j = min(j0 + u, cnt - 1);
          const int cj = __builtin_amdgcn_readlane(colL, j);
          av[u] = readLane(aL, j);
          xv[u] = Xt[static_cast<int64_t>(cj) * nrhs + cl];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) acc = (j0 + u < cnt) ? fma(av[u], xv[u], acc) : acc;
      }
    }
    tile[lane][rr] -= acc;
  }
  __syncthreads();
  double mx = 0.0;
  for (int cc = w; cc < 64; cc += 4) {  // R tile out, column segments
    const int cR = c0 + cc, i = i0 + lane;
    if (cR < nrhs && i < n) {
      const double r = tile[cc][lane];
      R[i + static_cast<int64_t>(cR) * n] = r;
      mx = fmax(mx, normAbs(r));
    }
  }
  if (rmaxBits) groupAtomicMax(mx, rmaxBits);
}

// out = x + dx, and max |out| into *maxBits when given.
__global__ void k_axpy(int64_t len, const double* __restrict__ x, const double* __restrict__ dx,
                       double* __restrict__ out, unsigned long long* maxBits) {
  double mx = 0.0;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < len;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double v = x[e] + dx[e];
    out[e] = v;
    mx = fmax(mx, normAbs(v));
  }
  if (maxBits) groupAtomicMax(mx, maxBits);
}

// max |v| over a strided n x nrhs block, into *outBits (must start at 0).
__global__ void k_absmax(int n, int nrhs, const double* __restrict__ v, int64_t ld,
                         unsigned long long* outBits) {
  double mx = 0.0;
  for (int c = 0; c < nrhs; ++c) {
    const double* col = v + static_cast<int64_t>(c) * ld;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
      mx = fmax(mx, normAbs(col[i]));
  }
  groupAtomicMax(mx, outBits);
}

// Per-column max |v| of an n x nrhs block: workgroup (c, part) covers rows
// part, part + parts, ... of column c; outBits[c * stride] receives the max.
__global__ void k_absmax_cols(int n, int nrhs, const double* __restrict__ v, int64_t ld, int parts,
                              unsigned long long* outBits, int stride) {
  const int c = blockIdx.x / parts;
  const int part = blockIdx.x % parts;
  if (c >= nrhs) return;
  const double* col = v + static_cast<int64_t>(c) * ld;
  double mx = 0.0;
  for (int i = part * blockDim.x + threadIdx.x; i < n; i += parts * blockDim.x) mx = fmax(mx, normAbs(col[i]));
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0 && mx > 0.0)
    atomicMax(outBits + static_cast<int64_t>(c) * stride, static_cast<unsigned long long>(__double_as_longlong(mx)));
}

// max over rows of sum |a| over a row-grouped entry list.  Over the
// symmetric CSR this is the true infinity norm; over the caller's stored
// entries grouped by row it is the reference's sparseMaxAbsRowSum
// (refine.hpp:69-81).
__global__ void k_rowsum_max(int n, const int64_t* __restrict__ rowPtr, const int64_t* __restrict__ valSrc,
                             const double* __restrict__ aVals, unsigned long long* outBits) {
  double mx = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double sacc = 0.0;
    for (int64_t t = rowPtr[i]; t < rowPtr[i + 1]; ++t) sacc += normAbs(aVals[valSrc[t]]);
    mx = fmax(mx, sacc);
  }
  groupAtomicMax(mx, outBits);
}

// ---- per-system norms of a batch of independent systems ----
// The systems own consecutive row ranges; rowGroup[i] is the system of row
// i.  A wave whose first and last rows belong to one system reduces before
// its single atomic; a wave straddling a boundary falls back to one atomic
// per lane.  Loops step by whole workgroups so every lane reaches every
// shuffle.
__device__ inline void segmentMax(double v, int g, unsigned long long* out, int stride) {
  const int g0 = __shfl(g, 0, 64), g1 = __shfl(g, 63, 64);
  if (g0 == g1) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0 && g0 >= 0 && v > 0.0)
      atomicMax(out + static_cast<int64_t>(g0) * stride, static_cast<unsigned long long>(__double_as_longlong(v)));
  } else if (g >= 0 && v > 0.0) {
    atomicMax(out + static_cast<int64_t>(g) * stride, static_cast<unsigned long long>(__double_as_longlong(v)));
  }
}

// out[g * stride] = max |v(i, c)| over the rows i of system g and all columns
__global__ void k_group_absmax(int n, int nrhs, const double* __restrict__ v, int64_t ld,
                               const int* __restrict__ rowGroup, unsigned long long* out, int stride) {
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    double mx = 0.0;
    int g = -1;
    if (i < n) {
      g = rowGroup[i];
      for (int c = 0; c < nrhs; ++c) mx = fmax(mx, normAbs(v[i + static_cast<int64_t>(c) * ld]));
    }
    segmentMax(mx, g, out, stride);
  }
}

// out[g * stride + c * colStride] = max |v(i, c)| over the rows i of system g
__global__ void k_group_absmax_cols(int n, int nrhs, const double* __restrict__ v, int64_t ld,
                                    const int* __restrict__ rowGroup, unsigned long long* out, int stride,
                                    int colStride) {
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    const int g = i < n ? rowGroup[i] : -1;
    for (int c = 0; c < nrhs; ++c) {
      const double a = i < n ? normAbs(v[i + static_cast<int64_t>(c) * ld]) : 0.0;
      segmentMax(a, g, out + static_cast<int64_t>(c) * colStride, stride);
    }
  }
}

// out[g * stride] = max over the rows i of system g of sum |a| over row i's entries
__global__ void k_group_rowsum_max(int n, const int64_t* __restrict__ rowPtr, const int64_t* __restrict__ valSrc,
                                   const double* __restrict__ aVals, const int* __restrict__ rowGroup,
                                   unsigned long long* out, int stride) {
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    double sacc = 0.0;
    int g = -1;
    if (i < n) {
      g = rowGroup[i];
      for (int64_t t = rowPtr[i]; t < rowPtr[i + 1]; ++t) sacc += normAbs(aVals[valSrc[t]]);
    }
    segmentMax(sacc, g, out, stride);
  }
}

// dst(i, :) = src(i, :) for the rows of systems with accept[g] != 0
__global__ void k_group_select(int n, int nrhs, const int* __restrict__ rowGroup, const int* __restrict__ accept,
                               const double* __restrict__ src, double* __restrict__ dst, int64_t ld) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!accept[rowGroup[i]]) continue;
    for (int c = 0; c < nrhs; ++c) dst[i + static_cast<int64_t>(c) * ld] = src[i + static_cast<int64_t>(c) * ld];
  }
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0ull;
}

__global__ void k_zero_u32(unsigned* p, int64_t count) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static int gridFor(int64_t total, int threads) {
  int64_t b = (total + threads - 1) / threads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}
void launchPermuteIn(int n, int nrhs, const int* perm, const double* B, int64_t ldb, double* Bp, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return;
  const dim3 grid((n + 63) / 64, (nrhs + 63) / 64);
  k_permute_in<<<grid, 256, 0, st>>>(n, nrhs, perm, B, ldb, Bp);
  hipCheck(hipGetLastError(), "launchPermuteIn");
}
void launchPermuteOut(int n, int nrhs, const int* perm, const double* X, double* Xo, int64_t ldx, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return;
  const dim3 grid((n + 63) / 64, (nrhs + 63) / 64);
  k_permute_out<<<grid, 256, 0, st>>>(n, nrhs, perm, X, Xo, ldx);
  hipCheck(hipGetLastError(), "launchPermuteOut");
}
void launchSolveBegin(unsigned* gen, hipStream_t st) {
  k_solve_begin<<<1, 64, 0, st>>>(gen);
  hipCheck(hipGetLastError(), "launchSolveBegin");
}
void launchPosInfo(const int2* pos, int64_t positions, const DevFronts& fd, const int64_t* gOff,
                   const int64_t* mapPtr, PosInfo* info, hipStream_t st) {
  if (positions <= 0) return;
  k_pos_info<<<static_cast<unsigned>((positions + 255) / 256), 256, 0, st>>>(pos, positions, fd, gOff, mapPtr, info);
  hipCheck(hipGetLastError(), "launchPosInfo");
}
void launchFwdGather(const PosInfo* info, int64_t positions, int nrhs, const double* Bp, double* Dl,
                     const int64_t* mapRow, hipStream_t st) {
  if (positions <= 0 || nrhs <= 0) return;
  int lpShift = 0;  // lanes per position: the power of two at or above nrhs / 4, at most 16
  while ((4 << lpShift) < nrhs && lpShift < 4) ++lpShift;
  const int64_t blocks = ((positions << lpShift) + 255) / 256;
  k_fwd_gather<<<static_cast<unsigned>(blocks), 256, 0, st>>>(info, positions, lpShift, nrhs, Bp, Dl, mapRow);
  hipCheck(hipGetLastError(), "launchFwdGather");
}
void launchDiagInverse(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd,
                       double* inv, const int64_t* invOff, hipStream_t st) {
  if (count <= 0 || blocks <= 0) return;
  if (fd.fp32)
    k_diag_inverse<float><<<blocks, 64, 0, st>>>(list, count, blockStart, fd, inv, invOff);
  else
    k_diag_inverse<double><<<blocks, 64, 0, st>>>(list, count, blockStart, fd, inv, invOff);
  hipCheck(hipGetLastError(), "launchDiagInverse");
}
void launchFwdOwn(const int* list, int count, const DevFronts& fd, int nrhs, double* Z, double* Dl,
                  const int64_t* gOff, const double* inv, const int64_t* invOff, hipStream_t st) {
  if (count <= 0 || nrhs <= 0) return;
  if (fd.fp32)
    k_fwd_own<float><<<count, 256, 0, st>>>(list, count, fd, nrhs, Z, Dl, gOff, inv, invOff);
  else
    k_fwd_own<double><<<count, 256, 0, st>>>(list, count, fd, nrhs, Z, Dl, gOff, inv, invOff);
  hipCheck(hipGetLastError(), "launchFwdOwn");
}
void launchFwdChain(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd, int nrhs,
                    double* Z, double* Dl, const int64_t* gOff, unsigned* flags, const int64_t* flagOff,
                    const unsigned* gen, int* err, const double* inv, const int64_t* invOff, hipStream_t st) {
  if (count <= 0 || blocks <= 0 || nrhs <= 0) return;
  if (fd.fp32)
    k_fwd_chain<float><<<blocks, 256, 0, st>>>(list, count, blockStart, fd, nrhs, Z, Dl, gOff, flags, flagOff, gen,
                                               err, inv, invOff);
  else
    k_fwd_chain<double><<<blocks, 256, 0, st>>>(list, count, blockStart, fd, nrhs, Z, Dl, gOff, flags, flagOff, gen,
                                                err, inv, invOff);
  hipCheck(hipGetLastError(), "launchFwdChain");
}
void launchFwdDelta(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd, int nrhs,
                    double* Dl, const int64_t* gOff, int* err, hipStream_t st) {
  if (count <= 0 || blocks <= 0 || nrhs <= 0) return;
  if (fd.fp32)
    k_fwd_delta<float><<<blocks, 256, 0, st>>>(list, count, blockStart, fd, nrhs, Dl, gOff, err);
  else
    k_fwd_delta<double><<<blocks, 256, 0, st>>>(list, count, blockStart, fd, nrhs, Dl, gOff, err);
  hipCheck(hipGetLastError(), "launchFwdDelta");
}
void launchBwdAnc(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd, int nrhs,
                  const double* Z, const double* X, double* Dl, const int64_t* gOff, const AncPart* parts, int* err,
                  hipStream_t st) {
  if (count <= 0 || blocks <= 0 || nrhs <= 0) return;
  if (fd.fp32)
    k_bwd_anc<float><<<blocks, 256, 0, st>>>(list, count, blockStart, fd, nrhs, Z, X, Dl, gOff, parts, err);
  else
    k_bwd_anc<double><<<blocks, 256, 0, st>>>(list, count, blockStart, fd, nrhs, Z, X, Dl, gOff, parts, err);
  hipCheck(hipGetLastError(), "launchBwdAnc");
}
void launchBwdOwn(const int* list, int count, const DevFronts& fd, int nrhs, double* X, double* Dl,
                  const int64_t* gOff, const AncPart* parts, const double* inv, const int64_t* invOff,
                  hipStream_t st) {
  if (count <= 0 || nrhs <= 0) return;
  if (fd.fp32)
    k_bwd_own<float><<<count, 256, 0, st>>>(list, count, fd, nrhs, X, Dl, gOff, parts, inv, invOff);
  else
    k_bwd_own<double><<<count, 256, 0, st>>>(list, count, fd, nrhs, X, Dl, gOff, parts, inv, invOff);
  hipCheck(hipGetLastError(), "launchBwdOwn");
}
void launchBwdChain(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd, int nrhs,
                    double* X, double* Dl, const int64_t* gOff, const AncPart* parts, unsigned* flags,
                    const int64_t* flagOff, const unsigned* gen, int* err, const double* inv, const int64_t* invOff,
                    hipStream_t st) {
  if (count <= 0 || blocks <= 0 || nrhs <= 0) return;
  if (fd.fp32)
    k_bwd_chain<float><<<blocks, 256, 0, st>>>(list, count, blockStart, fd, nrhs, X, Dl, gOff, parts, flags, flagOff,
                                               gen, err, inv, invOff);
  else
    k_bwd_chain<double><<<blocks, 256, 0, st>>>(list, count, blockStart, fd, nrhs, X, Dl, gOff, parts, flags,
                                                flagOff, gen, err, inv, invOff);
  hipCheck(hipGetLastError(), "launchBwdChain");
}
void launchFwdTiny(const int* list, int count, const DevFronts& fd, int nrhs, double* Z, double* Dl,
                   const int64_t* gOff, hipStream_t st) {
  if (count <= 0 || nrhs <= 0) return;
  const int blocks = (count + 3) / 4;
  if (fd.fp32)
    k_fwd_tiny<float><<<blocks, 256, 0, st>>>(list, count, fd, nrhs, Z, Dl, gOff);
  else
    k_fwd_tiny<double><<<blocks, 256, 0, st>>>(list, count, fd, nrhs, Z, Dl, gOff);
  hipCheck(hipGetLastError(), "launchFwdTiny");
}
void launchBwdTiny(const int* list, int count, const DevFronts& fd, int nrhs, const double* Z, double* X,
                   hipStream_t st) {
  if (count <= 0 || nrhs <= 0) return;
  const int blocks = (count + 3) / 4;
  if (fd.fp32)
    k_bwd_tiny<float><<<blocks, 256, 0, st>>>(list, count, fd, nrhs, Z, X);
  else
    k_bwd_tiny<double><<<blocks, 256, 0, st>>>(list, count, fd, nrhs, Z, X);
  hipCheck(hipGetLastError(), "launchBwdTiny");
}
void launchResidual(int n, int nrhs, const int64_t* rowPtr, const int* colIdx, const double* aSym, const double* B,
                    const double* X, double* R, unsigned long long* rmaxBits, hipStream_t st, double* Xt) {
  if (n <= 0 || nrhs <= 0) return;
  if (Xt && nrhs >= kResWide) {
    const dim3 grid((n + 63) / 64, (nrhs + 63) / 64);
    k_transpose_rows<<<grid, 256, 0, st>>>(n, nrhs, X, n, Xt);
    k_residual_wide<<<grid, 256, 0, st>>>(n, nrhs, rowPtr, colIdx, aSym, B, Xt, R, rmaxBits);
  } else {
    k_residual<<<(n + 255) / 256, 256, 0, st>>>(n, nrhs, rowPtr, colIdx, aSym, B, X, R, rmaxBits);
  }
  hipCheck(hipGetLastError(), "launchResidual");
}
void launchGatherVals(int64_t nnz, const int64_t* valSrc, const double* aVals, double* aSym, hipStream_t st) {
  if (nnz <= 0) return;
  k_gather_vals<<<gridFor(nnz, 256), 256, 0, st>>>(nnz, valSrc, aVals, aSym);
  hipCheck(hipGetLastError(), "launchGatherVals");
}
void launchAxpy(int64_t len, const double* x, const double* dx, double* out, unsigned long long* maxBits,
                hipStream_t st) {
  k_axpy<<<std::min(gridFor(len, 256), 2048), 256, 0, st>>>(len, x, dx, out, maxBits);
  hipCheck(hipGetLastError(), "launchAxpy");
}
void launchAbsMax(int n, int nrhs, const double* v, int64_t ld, unsigned long long* outBits, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return;
  k_absmax<<<gridFor(std::min<int64_t>(n, 256 * 1024), 256), 256, 0, st>>>(n, nrhs, v, ld, outBits);
  hipCheck(hipGetLastError(), "launchAbsMax");
}
void launchAbsMaxCols(int n, int nrhs, const double* v, int64_t ld, unsigned long long* outBits, int stride,
                      hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return;
  int parts = (n + 256 * 16 - 1) / (256 * 16);
  if (parts > 64) parts = 64;
  if (parts < 1) parts = 1;
  k_absmax_cols<<<nrhs * parts, 256, 0, st>>>(n, nrhs, v, ld, parts, outBits, stride);
  hipCheck(hipGetLastError(), "launchAbsMaxCols");
}
void launchRowSumMax(int n, const int64_t* rowPtr, const int64_t* valSrc, const double* aVals,
                     unsigned long long* outBits, hipStream_t st) {
  k_rowsum_max<<<std::min(gridFor(n, 256), 2048), 256, 0, st>>>(n, rowPtr, valSrc, aVals, outBits);
  hipCheck(hipGetLastError(), "launchRowSumMax");
}
void launchGroupAbsMax(int n, int nrhs, const double* v, int64_t ld, const int* rowGroup, unsigned long long* out,
                       int stride, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return;
  k_group_absmax<<<std::min(gridFor(n, 256), 4096), 256, 0, st>>>(n, nrhs, v, ld, rowGroup, out, stride);
  hipCheck(hipGetLastError(), "launchGroupAbsMax");
}
void launchGroupAbsMaxCols(int n, int nrhs, const double* v, int64_t ld, const int* rowGroup,
                           unsigned long long* out, int stride, int colStride, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return;
  k_group_absmax_cols<<<std::min(gridFor(n, 256), 4096), 256, 0, st>>>(n, nrhs, v, ld, rowGroup, out, stride,
                                                                       colStride);
  hipCheck(hipGetLastError(), "launchGroupAbsMaxCols");
}
void launchGroupRowSumMax(int n, const int64_t* rowPtr, const int64_t* valSrc, const double* aVals,
                          const int* rowGroup, unsigned long long* out, int stride, hipStream_t st) {
  if (n <= 0) return;
  k_group_rowsum_max<<<std::min(gridFor(n, 256), 4096), 256, 0, st>>>(n, rowPtr, valSrc, aVals, rowGroup, out,
                                                                      stride);
  hipCheck(hipGetLastError(), "launchGroupRowSumMax");
}
void launchGroupSelect(int n, int nrhs, const int* rowGroup, const int* accept, const double* src, double* dst,
                       int64_t ld, hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return;
  k_group_select<<<std::min(gridFor(n, 256), 4096), 256, 0, st>>>(n, nrhs, rowGroup, accept, src, dst, ld);
  hipCheck(hipGetLastError(), "launchGroupSelect");
}
void launchZeroU64(unsigned long long* p, int count, hipStream_t st) {
  if (count <= 0) return;
  k_zero_u64<<<(count + 255) / 256, 256, 0, st>>>(p, count);
  hipCheck(hipGetLastError(), "launchZeroU64");
}
void launchZeroU32(unsigned* p, int64_t count, hipStream_t st) {
  if (count <= 0) return;
  k_zero_u32<<<gridFor(count, 256), 256, 0, st>>>(p, count);
  hipCheck(hipGetLastError(), "launchZeroU32");
}

}  // namespace symla_hip
