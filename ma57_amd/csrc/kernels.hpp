// kernels.hpp -- host-side launchers of the numeric-phase kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "device.hpp"

namespace symla_hip {

// Throws DeviceError for a failed HIP call (defined with the plan).
void hipCheck(hipError_t e, const char* what);

// factor_kernels.hip
size_t smallFrontSmemBytes(int m);
void launchResetFronts(int numSN, const DevFronts& fd, hipStream_t st);
void launchResetPerm(const int* list, int count, const DevFronts& fd, hipStream_t st);
void launchAssemble(const int4* tasks, int ntasks, const DevFronts& fd, const double* aVals, int monotone,
                    hipStream_t st);
void launchFactorSmall(const int* list, int count, int maxM, const DevFronts& fd, const FactorParams& prm,
                       hipStream_t st);
void launchFrontMaxAbs(const int* list, int count, int maxM, const DevFronts& fd, hipStream_t st);
void launchPanel(const int* list, int count, int maxRows, const DevFronts& fd, const FactorParams& prm, double* ws,
                 const int64_t* wsOff, hipStream_t st);
// cooperative panels: list[f] factored by workgroups [wgStart[f], wgStart[f+1]) with sync[f]
void launchPanelMw(const int* list, int count, const int* wgStart, int blocks, const DevFronts& fd,
                   const FactorParams& prm, double* ws, const int64_t* wsOff, MwSync* sync, int* err,
                   hipStream_t st);
// Speculative panel step (both passes) for the fronts of one panel iteration;
// wgStart[f] .. wgStart[f+1] are front f's workgroups (kSpecRows rows each,
// at least one), parity = iteration index & 1.
void launchSpecPanel(const int* list, int count, const int* wgStart, int blocks, int parity, const DevFronts& fd,
                     const FactorParams& prm, double* ws, const int64_t* wsOff, hipStream_t st);
void launchSpecReset(const int* list, int count, const DevFronts& fd, double* ws, const int64_t* wsOff,
                     hipStream_t st);
void launchPanelRows(const int* list, int count, const int* blockStart, int totalBlocks, const DevFronts& fd,
                     hipStream_t st);
// tileTp: per listed front, the planned 64-row tiles per side of its trailing block
void launchSyrk(const int* list, int count, const int* tileStart, const int* tileTp, int totalTiles,
                const DevFronts& fd, hipStream_t st);
// W = L(nE:m, q0:q0+kc) D of front s (rows = m - nE, column-major, ld = rows;
// float for FP32 fronts), zero past the pivots taken
void launchFormW(const DevFronts& fd, int s, int rows, int q0, int kc, void* W, hipStream_t st);
void launchFrontFinish(const int* list, int count, const DevFronts& fd, hipStream_t st);
void launchMakeRowIdx(const int* list, int count, const DevFronts& fd, hipStream_t st);
void launchReduceStats(int numSN, const DevFronts& fd, long long* outI, double* outD, hipStream_t st);

// solve_kernels.hip
void launchPermuteIn(int n, int nrhs, const int* perm, const double* B, int64_t ldb, double* Bp, hipStream_t st);
void launchPermuteOut(int n, int nrhs, const int* perm, const double* X, double* Xo, int64_t ldx, hipStream_t st);
void launchSolveBegin(unsigned* gen, hipStream_t st);
// Inside a solve, Bp, Z and X are n x nrhs row-major (ld nrhs) and a front's
// work block Dl is m x nrhs row-major at gOff[s] * nrhs.
// per-position records of the forward gather (pos[k] = (front, position)),
// formed at the start of each solve
void launchPosInfo(const int2* pos, int64_t positions, const DevFronts& fd, const int64_t* gOff,
                   const int64_t* mapPtr, PosInfo* info, hipStream_t st);
// forward gather over the positions info[0 .. positions); mapRow: the
// children's delta rows (work-row numbers) summed into each position
void launchFwdGather(const PosInfo* info, int64_t positions, int nrhs, const double* Bp, double* Dl,
                     const int64_t* mapRow, hipStream_t st);
// own rows per solve block: the solve applies the inverse of each block's
// unit lower diagonal block (kSolveBlock^2 doubles, row-major, at
// invOff[s] + kSolveBlock^2 * block), formed once per factorization
constexpr int kSolveBlock = 32;
void launchDiagInverse(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd,
                       double* inv, const int64_t* invOff, hipStream_t st);
// own rows, one workgroup per front (a short chain of own blocks)
void launchFwdOwn(const int* list, int count, const DevFronts& fd, int nrhs, double* Z, double* Dl,
                  const int64_t* gOff, const double* inv, const int64_t* invOff, hipStream_t st);
// own rows, one workgroup per own block (blockStart: ceil(nF / kSolveBlock) per front), dependency flags
void launchFwdChain(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd, int nrhs,
                    double* Z, double* Dl, const int64_t* gOff, unsigned* flags, const int64_t* flagOff,
                    const unsigned* gen, int* err, const double* inv, const int64_t* invOff, hipStream_t st);
// rows [nF, m) after the own rows (blockStart: ceil((m - nF) / 64) per front)
void launchFwdDelta(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd, int nrhs,
                    double* Dl, const int64_t* gOff, int* err, hipStream_t st);
// blockStart: split * ceil(nF / kSolveBlock) workgroups per front
void launchBwdAnc(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd, int nrhs,
                  const double* Z, const double* X, double* Dl, const int64_t* gOff, const AncPart* parts, int* err,
                  hipStream_t st);
void launchBwdOwn(const int* list, int count, const DevFronts& fd, int nrhs, double* X, double* Dl,
                  const int64_t* gOff, const AncPart* parts, const double* inv, const int64_t* invOff,
                  hipStream_t st);
void launchBwdChain(const int* list, int count, const int* blockStart, int blocks, const DevFronts& fd, int nrhs,
                    double* X, double* Dl, const int64_t* gOff, const AncPart* parts, unsigned* flags,
                    const int64_t* flagOff, const unsigned* gen, int* err, const double* inv, const int64_t* invOff,
                    hipStream_t st);
// fronts of at most kTinyFront rows: one wave per front, no dependency flags
constexpr int kTinyFront = 64;
void launchFwdTiny(const int* list, int count, const DevFronts& fd, int nrhs, double* Z, double* Dl,
                   const int64_t* gOff, hipStream_t st);
void launchBwdTiny(const int* list, int count, const DevFronts& fd, int nrhs, const double* Z, double* X,
                   hipStream_t st);
// R = B - A X over symmetric CSR values aSym (X, B, R: n x nrhs column-major, ld n); max |R| into
// *rmaxBits if non-null.  Xt (n x nrhs doubles, may be null): workspace for a row-major copy of X, used
// from 16 right-hand sides on
void launchResidual(int n, int nrhs, const int64_t* rowPtr, const int* colIdx, const double* aSym, const double* B,
                    const double* X, double* R, unsigned long long* rmaxBits, hipStream_t st, double* Xt = nullptr);
void launchGatherVals(int64_t nnz, const int64_t* valSrc, const double* aVals, double* aSym, hipStream_t st);
// out = x + dx; max |out| into *maxBits if non-null
void launchAxpy(int64_t len, const double* x, const double* dx, double* out, unsigned long long* maxBits,
                hipStream_t st);
void launchAbsMax(int n, int nrhs, const double* v, int64_t ld, unsigned long long* outBits, hipStream_t st);
// per-column max |v| into outBits[c * stride] (bits of non-negative doubles)
void launchAbsMaxCols(int n, int nrhs, const double* v, int64_t ld, unsigned long long* outBits, int stride,
                      hipStream_t st);
void launchRowSumMax(int n, const int64_t* rowPtr, const int64_t* valSrc, const double* aVals,
                     unsigned long long* outBits, hipStream_t st);
// per-system norms of a batch (systems own consecutive rows; rowGroup[i] =
// system of row i): outputs at out[g * stride] (+ c * colStride per column),
// bits of non-negative doubles folded with atomic max (start them at 0)
void launchGroupAbsMax(int n, int nrhs, const double* v, int64_t ld, const int* rowGroup, unsigned long long* out,
                       int stride, hipStream_t st);
void launchGroupAbsMaxCols(int n, int nrhs, const double* v, int64_t ld, const int* rowGroup,
                           unsigned long long* out, int stride, int colStride, hipStream_t st);
void launchGroupRowSumMax(int n, const int64_t* rowPtr, const int64_t* valSrc, const double* aVals,
                          const int* rowGroup, unsigned long long* out, int stride, hipStream_t st);
// rows of systems with accept[g] != 0 copied from src to dst (n x nrhs, leading dimension ld)
void launchGroupSelect(int n, int nrhs, const int* rowGroup, const int* accept, const double* src, double* dst,
                       int64_t ld, hipStream_t st);
void launchZeroU64(unsigned long long* p, int count, hipStream_t st);
void launchZeroU32(unsigned* p, int64_t count, hipStream_t st);

}  // namespace symla_hip
