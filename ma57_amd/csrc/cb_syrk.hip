// cb_syrk.hip -- the deferred contribution-block update of a large front on
// the matrix cores: C -= W B^T over the lower triangle of an N x N block,
// with W = L D and B = L over the block's rows and a chunk of K pivot
// columns (include/symla/dense_kernel.hpp:806-821 summed over a front's
// panels).  N is a few thousand to ~14,000 rows and K up to 2,048, so the
// update is a large-rank product bound by the matrix cores, unlike the
// panels' rank-64 updates, which re-read C once per 64 pivots.
//
// Layout.  128 x 128 output tiles of the lower triangle, one workgroup of
// four waves each (a 64 x 64 quarter per wave, 4 x 4 MFMA blocks of 16 x 16).
// K runs in steps of KT columns: both operand tiles (128 rows x KT) are
// staged in LDS as [row][k] with k contiguous, double-buffered, and the next
// step's global loads are issued before the current step's products, so one
// barrier per step separates the LDS writes from the reads.  A lane of the
// MFMA sums 4 k values per instruction (lane l takes k index l / 16); the
// KT columns of a step are assigned so that a lane's k values over the step
// are contiguous (k = (l / 16) * KT/4 + s for instruction s), which lets it
// read its operands for the whole step as 32-byte vectors.  Both operands
// use the same assignment, so every k term is summed exactly once.
//
// The product is formed transposed (B rows as the A operand), so each
// result register holds 16 consecutive rows of one C column: C is read and
// written in 64- / 128-byte column segments.  Tiles are enumerated in 8 x 8
// super-tiles of the lower triangle and handed out in contiguous runs per
// XCD (workgroup b runs on XCD b mod 8), so the tiles in flight on one XCD
// share operand row panels in that XCD's L2.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace symla_hip {

namespace {

template <typename T>
struct SyrkMfma;
template <>
struct SyrkMfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static constexpr int KT = 16;  // k columns per step
  static constexpr int kWaves = 8;  // 32 x 64 per wave: half the accumulators, two workgroups per CU
  static __device__ inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // column within the 16 x 16 block of register r (the row is lane % 16)
  static __device__ inline int col(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <>
struct SyrkMfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static constexpr int KT = 32;
  static constexpr int kWaves = 4;  // 64 x 64 per wave
  static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ inline int col(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int kSyTile = 128;
constexpr int kSySuper = 8;  // tiles per super-tile side
constexpr int kXcds = 8;

// (ti, tj) of lower-triangle tile `t` in super-tile order: super-tiles
// column by column (SJ, then SI >= SJ), tiles row-major inside one (only
// ti >= tj on a diagonal super-tile).
__device__ inline void syrkTile(int t, int nt, int& ti, int& tj) {
  const int ns = (nt + kSySuper - 1) / kSySuper;
  for (int SJ = 0; SJ < ns; ++SJ) {
    const int c0 = SJ * kSySuper, c1 = min(c0 + kSySuper, nt);
    for (int SI = SJ; SI < ns; ++SI) {
      const int r0 = SI * kSySuper, r1 = min(r0 + kSySuper, nt);
      int cnt;
      if (SI == SJ) {
        const int w = c1 - c0;
        cnt = w * (w + 1) / 2;
      } else {
        cnt = (r1 - r0) * (c1 - c0);
      }
      if (t < cnt) {
        if (SI == SJ) {  // row r of the diagonal super-tile holds r + 1 tiles
          int r = 0;
          while (t > r) {
            t -= r + 1;
            ++r;
          }
          ti = r0 + r;
          tj = c0 + t;
        } else {
          ti = r0 + t / (c1 - c0);
          tj = c0 + t % (c1 - c0);
        }
        return;
      }
      t -= cnt;
    }
  }
  ti = tj = -1;
}

template <typename T>
__global__ __launch_bounds__(64 * SyrkMfma<T>::kWaves) void k_cb_syrk(int N, int K, const T* __restrict__ W, int ldw,
                                                                      const T* __restrict__ B, int ldb,
                                                                      T* __restrict__ C, int ldc, int nt, int per) {
  using M = SyrkMfma<T>;
  using v4 = typename M::v4;
  constexpr int TS = kSyTile;
  constexpr int KT = M::KT;
  constexpr int NW = M::kWaves;           // 4: waves of 64 x 64; 8: waves of 32 x 64
  constexpr int BI = NW == 4 ? 4 : 2;     // 16-row MFMA blocks per wave
  constexpr int SPLIT = NW * 64 / TS;     // threads per staged row
  constexpr int KH = KT / SPLIT;          // staged k columns per thread
  constexpr int KQ = KT / 4;              // k values per lane per step
  constexpr int LDK = KT + 16 / sizeof(T);  // 16-byte row pad: conflict-free 16-row vector reads
  __shared__ __attribute__((aligned(16))) T Ws[2][TS][LDK];
  __shared__ __attribute__((aligned(16))) T Bs[2][TS][LDK];

  const int total = nt * (nt + 1) / 2;
  const int t = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (t >= total) return;
  int ti, tj;
  syrkTile(t, nt, ti, tj);
  const int I0 = ti * TS, J0 = tj * TS;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * (16 * BI), wc = (w & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  const int sr = tid & (TS - 1), sh = tid / TS;  // staging: row sr, k columns sh * KH .. + KH
  const int gi = I0 + sr, gj = J0 + sr;
  const bool wIn = gi < N, bIn = gj < N;
  const T* Wr = W + gi;
  const T* Br = B + gj;

  T wreg[KH], breg[KH];
  auto gload = [&](int k0) {
#pragma unroll
    for (int u = 0; u < KH; ++u) {
      const int k = k0 + sh * KH + u;
      const bool kin = k < K;
      wreg[u] = (wIn && kin) ? Wr[static_cast<size_t>(k) * ldw] : T(0);
      breg[u] = (bIn && kin) ? Br[static_cast<size_t>(k) * ldb] : T(0);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int u = 0; u < KH; ++u) {
      Ws[buf][sr][sh * KH + u] = wreg[u];
      Bs[buf][sr][sh * KH + u] = breg[u];
    }
  };

  v4 acc[BI][4];
#pragma unroll
  for (int bi = 0; bi < BI; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = v4{0, 0, 0, 0};

  const int nk = (K + KT - 1) / KT;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * KT);
    T a[4][KQ], b[BI][KQ];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int s = 0; s < KQ; ++s) a[bj][s] = Bs[buf][wc + bj * 16 + lr][lk * KQ + s];
#pragma unroll
    for (int bi = 0; bi < BI; ++bi)
#pragma unroll
      for (int s = 0; s < KQ; ++s) b[bi][s] = Ws[buf][wr + bi * 16 + lr][lk * KQ + s];
#pragma unroll
    for (int s = 0; s < KQ; ++s)
#pragma unroll
      for (int bi = 0; bi < BI; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = M::mma(a[bj][s], b[bi][s], acc[bi][bj]);
    if (kt + 1 < nk) lstore(buf ^ 1);
    __syncthreads();
  }

  // C(row, col) -= acc: register r of block (bi, bj) holds row lr of the
  // block and column M::col(lane, r).  Every C value is loaded before any is
  // used (entries outside the triangle read C's first entry and are not
  // stored), so the loads overlap instead of waiting one by one.
  T cv[BI][4][4];
#pragma unroll
  for (int bi = 0; bi < BI; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = I0 + wr + bi * 16 + lr, col = J0 + wc + bj * 16 + M::col(lane, r);
        const bool ok = row < N && col < N && (ti != tj || row >= col);
        cv[bi][bj][r] = C[ok ? row + static_cast<size_t>(col) * ldc : 0];
      }
#pragma unroll
  for (int bi = 0; bi < BI; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = I0 + wr + bi * 16 + lr, col = J0 + wc + bj * 16 + M::col(lane, r);
        if (row < N && col < N && (ti != tj || row >= col))
          C[row + static_cast<size_t>(col) * ldc] = cv[bi][bj][r] - acc[bi][bj][r];
      }
}

// BF16x3 split products (Henry, Tang and Heinecke, "Leveraging the bfloat16
// artificial intelligence datatype for higher-precision computations", ARITH
// 2019): each FP32 operand x is split as hi = bf16(x), lo = bf16(x - hi), and
// a b is formed as hi_a hi_b + hi_a lo_b + lo_a hi_b (the lo_a lo_b term,
// below 2^-16 of |a b|, is dropped) on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation: three bf16 products of 32 k each against eight FP32
// 16x16x4 products.  Same tiling, super-tile order and transposed product as
// k_cb_syrk; the operands are split while they are staged into LDS (hi and
// lo planes, [row][k], 16-byte row pad).  Layout of the instruction on gfx950
// (measured): A[i][k] from lane i + 16 (k / 8) element k % 8, B[k][j] from
// lane j + 16 (k / 8) element k % 8, D[i][j] in lane j + 16 (i / 4) register
// i % 4 -- the D layout of v_mfma_f32_16x16x4f32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_cb_syrk_bf16x3(int N, int K, const float* __restrict__ W, int ldw,
                                                        const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                        int ldc, int nt, int per) {
  constexpr int TS = kSyTile;
  constexpr int KT = 32;       // k per step = one bf16 instruction
  constexpr int KH = KT / 2;   // staged k columns per thread
  constexpr int LDK = KT + 8;  // 16-byte row pad
  __shared__ __attribute__((aligned(16))) __bf16 Wh[2][TS][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Wl[2][TS][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[2][TS][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[2][TS][LDK];

  const int total = nt * (nt + 1) / 2;
  const int t = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (t >= total) return;
  int ti, tj;
  syrkTile(t, nt, ti, tj);
  const int I0 = ti * TS, J0 = tj * TS;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  const int sr = tid & (TS - 1), sh = tid >> 7;
  const int gi = I0 + sr, gj = J0 + sr;
  const bool wIn = gi < N, bIn = gj < N;
  const float* Wr = W + gi;
  const float* Br = B + gj;

  float wreg[KH], breg[KH];
  auto gload = [&](int k0) {
#pragma unroll
    for (int u = 0; u < KH; ++u) {
      const int k = k0 + sh * KH + u;
      const bool kin = k < K;
      wreg[u] = (wIn && kin) ? Wr[static_cast<size_t>(k) * ldw] : 0.0f;
      breg[u] = (bIn && kin) ? Br[static_cast<size_t>(k) * ldb] : 0.0f;
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int u = 0; u < KH; ++u) {
      const __bf16 wh = static_cast<__bf16>(wreg[u]);
      const __bf16 bh = static_cast<__bf16>(breg[u]);
      Wh[buf][sr][sh * KH + u] = wh;
      Wl[buf][sr][sh * KH + u] = static_cast<__bf16>(wreg[u] - static_cast<float>(wh));
      Bh[buf][sr][sh * KH + u] = bh;
      Bl[buf][sr][sh * KH + u] = static_cast<__bf16>(breg[u] - static_cast<float>(bh));
    }
  };

  f32x4 acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = f32x4{0, 0, 0, 0};

  const int nk = (K + KT - 1) / KT;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * KT);
    bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) {
      ah[bj] = *reinterpret_cast<const bf16x8*>(&Bh[buf][wc + bj * 16 + lr][lk * 8]);
      al[bj] = *reinterpret_cast<const bf16x8*>(&Bl[buf][wc + bj * 16 + lr][lk * 8]);
    }
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      bh[bi] = *reinterpret_cast<const bf16x8*>(&Wh[buf][wr + bi * 16 + lr][lk * 8]);
      bl[bi] = *reinterpret_cast<const bf16x8*>(&Wl[buf][wr + bi * 16 + lr][lk * 8]);
    }
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bj], bh[bi], acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bj], bl[bi], acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bj], bh[bi], acc[bi][bj], 0, 0, 0);
      }
    if (kt + 1 < nk) lstore(buf ^ 1);
    __syncthreads();
  }
  // C -= acc, every C value loaded before any is used (as in k_cb_syrk)
  float cv[4][4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = I0 + wr + bi * 16 + lr, col = J0 + wc + bj * 16 + 4 * lk + r;
        const bool ok = row < N && col < N && (ti != tj || row >= col);
        cv[bi][bj][r] = C[ok ? row + static_cast<size_t>(col) * ldc : 0];
      }
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = I0 + wr + bi * 16 + lr, col = J0 + wc + bj * 16 + 4 * lk + r;
        if (row < N && col < N && (ti != tj || row >= col))
          C[row + static_cast<size_t>(col) * ldc] = cv[bi][bj][r] - acc[bi][bj][r];
      }
}

}  // namespace

void launchCbSyrkBf16x3(int N, int K, const float* W, int ldw, const float* B, int ldb, float* C, int ldc,
                        hipStream_t st) {
  if (N <= 0 || K <= 0) return;
  const int nt = (N + kSyTile - 1) / kSyTile;
  const int total = nt * (nt + 1) / 2;
  const int per = (total + kXcds - 1) / kXcds;
  k_cb_syrk_bf16x3<<<dim3(per * kXcds), 256, 0, st>>>(N, K, W, ldw, B, ldb, C, ldc, nt, per);
  hipCheck(hipGetLastError(), "launchCbSyrkBf16x3");
}

void launchCbSyrk(bool fp32, int N, int K, const void* W, int ldw, const void* B, int ldb, void* C, int ldc,
                  hipStream_t st) {
  if (N <= 0 || K <= 0) return;
  const int nt = (N + kSyTile - 1) / kSyTile;
  const int total = nt * (nt + 1) / 2;
  const int per = (total + kXcds - 1) / kXcds;
  const dim3 grid(per * kXcds);
  if (fp32)
    k_cb_syrk<float><<<grid, 64 * SyrkMfma<float>::kWaves, 0, st>>>(N, K, static_cast<const float*>(W), ldw,
                                                                     static_cast<const float*>(B), ldb,
                                                                     static_cast<float*>(C), ldc, nt, per);
  else
    k_cb_syrk<double><<<grid, 64 * SyrkMfma<double>::kWaves, 0, st>>>(N, K, static_cast<const double*>(W), ldw,
                                                                       static_cast<const double*>(B), ldb,
                                                                       static_cast<double*>(C), ldc, nt, per);
  hipCheck(hipGetLastError(), "launchCbSyrk");
}

}  // namespace symla_hip
