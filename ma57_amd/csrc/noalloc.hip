// noalloc.hip -- the static device scratch of the allocation-free solve
// entry point (symla_solve_device_noalloc in include/symla_hip.h).
//
// The solve's per-call workspace (permuted right-hand sides, forward and
// backward sweep vectors, per-front work blocks) normally comes from plan
// buffers grown with hipMalloc.  Callers that cannot allocate -- code under
// stream capture, tight serving loops -- take it from this file-scope array
// instead.  Each GPU the library runs on holds its own copy of the array
// until the process ends.  In host code the array's name is only a host-side
// stand-in, so its device address is looked up once per device, with that
// device current, and only that address is handed to kernels.  Every call on
// one GPU shares that GPU's array: calls that overlap on one GPU (two plans
// on two streams, or two host threads) race and must be serialized by the
// caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"

namespace symla_hip {

namespace {
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

std::mutex g_scratchMutex;
std::vector<double*> g_scratchPtr;  // per device ordinal; null until looked up
}  // namespace

double* noAllocScratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (device < 0) throw std::invalid_argument("symla: invalid device ordinal");
  if (static_cast<size_t>(device) >= g_scratchPtr.size()) g_scratchPtr.resize(device + 1, nullptr);
  if (!g_scratchPtr[device]) {
    int cur = -1;
    hipCheck(hipGetDevice(&cur), "hipGetDevice");
    if (cur != device) hipCheck(hipSetDevice(device), "hipSetDevice");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    hipCheck(e, "hipGetSymbolAddress(scratch)");
    if (!p) throw DeviceError("symla: the static solve scratch has no device address");
    g_scratchPtr[device] = static_cast<double*>(p);
  }
  return g_scratchPtr[device];
}

}  // namespace symla_hip
