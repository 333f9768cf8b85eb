// cb_update.cpp -- the library SYRK of the deferred contribution-block
// update: CB -= W B^T on the lower triangle (rocBLAS syrkx), W = L D and
// B = L over the contribution-block rows of a front's pivot columns.
#include <rocblas/rocblas.h>

#include <string>

#include "kernels.hpp"
#include "plan.hpp"

namespace symla_hip {

namespace {
void blasCheck(rocblas_status s, const char* what) {
  if (s != rocblas_status_success) throw DeviceError(std::string("symla: rocBLAS ") + what + " failed (status " +
                                                     std::to_string(static_cast<int>(s)) + ")");
}
}  // namespace

void* blasCreate(hipStream_t st) {
  rocblas_handle h = nullptr;
  blasCheck(rocblas_create_handle(&h), "create_handle");
  blasCheck(rocblas_set_stream(h, st), "set_stream");
  // no split-K atomics: the same inputs give the same factors on every run
  blasCheck(rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed), "set_atomics_mode");
  return h;
}

void blasDestroy(void* h) {
  if (h) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(h));
}

void blasSyrkxLower(void* h, bool fp32, int n, int k, const void* W, int ldw, const void* B, int ldb, void* C,
                    int ldc) {
  auto hd = static_cast<rocblas_handle>(h);
  if (fp32) {
    const float alpha = -1.0f, beta = 1.0f;
    blasCheck(rocblas_ssyrkx(hd, rocblas_fill_lower, rocblas_operation_none, n, k, &alpha,
                             static_cast<const float*>(W), ldw, static_cast<const float*>(B), ldb, &beta,
                             static_cast<float*>(C), ldc),
              "ssyrkx");
  } else {
    const double alpha = -1.0, beta = 1.0;
    blasCheck(rocblas_dsyrkx(hd, rocblas_fill_lower, rocblas_operation_none, n, k, &alpha,
                             static_cast<const double*>(W), ldw, static_cast<const double*>(B), ldb, &beta,
                             static_cast<double*>(C), ldc),
              "dsyrkx");
  }
}

}  // namespace symla_hip
