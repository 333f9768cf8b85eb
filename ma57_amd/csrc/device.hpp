// device.hpp -- shared host/device definitions for the numeric-phase kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace symla_hip {

// Front dimensions above this use the panel path (global memory); at or below
// it the whole packed lower triangle lives in LDS.
constexpr int kSmallFrontMax = 192;
constexpr int kSmallFrontDefault = 64;  // the plan's default bound (tunable up to kSmallFrontMax)
// Panel width of the large-front path.
constexpr int kPanel = 64;
// Threads of the small-front and panel kernels.
constexpr int kSmallThreads = 256;
constexpr int kPanelThreads = 512;

// Per-front metadata and outputs, all device arrays indexed by supernode id.
// Per-position arrays (frontIdx, perm, rowIdx, dDiag, dOff, pkind, relmap)
// are indexed by idxOff[s] + local position, local position in [0, m[s]).
// One front position of the solve's forward gather: its work row, the
// children's delta rows summed into it (mapRow[e0 .. e0 + cnt)), and the
// final-order row of B it starts from (own rows) or -1.
struct PosInfo {
  int64_t dst;
  int64_t e0;
  int cnt;
  int rix;
};
// The backward pass's product with a front's ancestors' rows, L(nF:m, own)^T
// x_anc, split over those rows in `split` chunks: chunk 0 leaves z less its
// part in the own work rows, chunk k >= 1 its part in the work rows
// off + (k - 1) * stride + (own row), Dl rows past the fronts' blocks.
struct AncPart {
  int64_t off;
  int stride;
  int split;
};

struct DevFronts {
  const int* m;          // front order
  const int* nE;         // eligible (fully summed) columns
  const int* ncols;      // own columns
  void* const* F;        // front storage, column-major, ld = m (double, or float when fp32)
  int fp32;              // 1: fronts hold float (FP32 factor storage and Schur updates)
  const int64_t* idxOff; // offset into per-position arrays
  const int* frontIdx;   // final-order global index of pre-permutation local position
  const int* relmap;     // child position -> parent pre-permutation position (-1: none)
  const int* childPtr;   // children CSR (ordered as the reference: increasing id)
  const int* childList;
  const int* firstCol;   // first own column (final order)
  const int64_t* acolPtr;  // A entries per final column (CSR over columns)
  const int* aRow;       // local pre-permutation row in the owning front
  const int64_t* aSrc;   // index into the caller's value array
  const int* schurCut;   // the panels' trailing updates cover columns [0, schurCut); when it is nE < m the
                         // contribution block [nE, m)^2 gets one rank-nE update after the last panel
  // outputs / state
  int* perm;             // local permutation: position -> pre-permutation position
  int* rowIdx;           // frontIdx[perm[i]]
  double* dDiag;
  double* dOff;
  int* pkind;            // 1: 1x1 at position, 2: first of 2x2, 0: second of 2x2 / none
  int* nF;               // pivots factored
  int* inertia;          // 3 per front: pos, neg, zero
  unsigned long long* origMaxBits;  // max|F| at front start (bits of a non-negative double)
  unsigned long long* runMaxBits;   // running max (bits)
  int* kcur;             // next pivot position (panel path)
  int* pcount;           // pivots in the last panel (panel path)
  int* stopped;          // 1 when a front stopped on a delay
  int* cbDone;           // 1 when the speculative panel also wrote the last panel's rows [nE, m)
  double* pert;          // static mode: sum |perturbation|
  int* npert;            // static mode: perturbed pivots
  unsigned long long* schurFlops;  // algorithmic flops of the MFMA trailing updates (p*r*(r+1) per update)
};

// Cooperative panel of one front whose eligible rows exceed what one
// workgroup holds: W workgroups of kPanelThreads threads, kMwRows rows each,
// exchange per pivot through one MwSync record in device memory.
constexpr int kMwRows = 2 * kPanelThreads;
constexpr int kMwMaxW = 32;
constexpr int kMwSlots = 16;  // giant fronts per launch
struct MwSync {
  unsigned long long flag[kMwMaxW];  // tag of the last step each workgroup posted
  double lv[2][kMwMaxW];             // per-step local max |a_ik| (double-buffered by step parity)
  int li[2][kMwMaxW];                // and its row
  double extra[2][12];               // workgroup 0: a_kk and the pivot column at rows k + 1 .. k + 8
  int state[4];                      // workgroup 0 after a general step: k, p, stop
  unsigned epoch;                    // launches so far on this record (tags stay unique)
  int pad;
};

struct FactorParams {
  double zeroTol;        // DenseLDLTOptions::zero_tolerance
  double relFloor;       // DenseLDLTOptions::relative_pivot_floor
  int staticMode;        // 1: factorStatic semantics
  double staticDelta;    // > 0: delta; <= 0: sqrt(eps) * orig_max per front
  double staticRelFloor;
  double staticAbsFloor;
  const int* expectedSign;  // final-order expected signs (static mode) or null
  int specPanel;         // 1: panels of fronts with more than specMinRows eligible rows
  int specMinRows;       //    start with the speculative blocked step
};
__host__ __device__ inline bool specFront(const FactorParams& p, int nE) { return p.specPanel && nE > p.specMinRows; }

// Speculative panel step: a panel's pivots are first taken in order as 1x1
// without interchanges (diagonal block factored in LDS, the rows below by
// forward substitution), then the threshold tests are checked for every
// column at once; the prefix that passes is kept and the exact panel kernel
// continues from the first column that fails.  Its per-front record lives in
// the front's workspace after the two column buffers (2 * nE doubles):
//   colmax[2][kPanel]  max |c_iq| over the eligible rows below the block,
//                      rounded up to float (bits of non-negative doubles;
//                      double-buffered by panel iteration parity)
//   emax[kPanel]       the same maxima, exact
//   nmax[kPanel]       max |c_iq| over the rows [nE, m), exact
//   cdiag[kPanel]      a_qq at pivot q (before any static perturbation)
//   dval[kPanel]       the pivot d_q (1 past the block)
//   bmax[kPanel]       max |c_iq| over the block rows below q
//   rdiv[kPanel]       1 / d_q (1 past the block)
//   lblk[kPanel * kPanel]  multipliers of the block, l_iq at [i * kPanel + q]
//   wmat[kPanel * kPanel]  d_j l_qj at [q * kPanel + j] (j < q): the row
//                      substitution's coefficients
// followed, for fronts that take the step, by the backup of the raw panel
// values of the rows below the block (m x kPanel front elements, row i of
// panel column q at i + q m), from which a failed test restores them.
constexpr int kSpecRec = 8 * kPanel + 2 * kPanel * kPanel;
constexpr int kSpecRows = 256;  // rows below the block per workgroup (one per thread)
__host__ __device__ inline int64_t panelWorkspace(int m, int nE, bool spec) {
  return 2 * static_cast<int64_t>(nE) + kSpecRec + (spec ? static_cast<int64_t>(m) * kPanel : 0);
}

}  // namespace symla_hip
