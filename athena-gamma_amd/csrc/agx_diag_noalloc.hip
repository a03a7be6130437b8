// agx_diag_noalloc.hip -- the history sums without any device allocation.
//
// agx_history_sum_noalloc is agx_history_sum with its per-workgroup partial sums in a
// file-scope device array (`scratch`, HIST_PARTIAL_DOUBLES doubles = 192 KiB) instead of
// memory allocated per call, for callers that cannot allocate (code under stream
// capture, tight loops, harnesses that forbid hipMalloc).  Each GPU the library runs on
// holds its own copy of the array until the process ends.  The work split caps the
// workgroup count at HIST_NW_MAX, so every mesh size fits the array; inputs beyond the
// 32-bit row index (nblocks * nx3 * nx2 > 2^31 - 1) are rejected.  Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU (other streams, other host
// threads) race on it, and a call leaves its partials behind (every partial row the final
// sum reads is first written by the same call, so nothing needs clearing).
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "agx_internal.hpp"

namespace agx {

static __device__ __align__(256) double scratch[HIST_PARTIAL_DOUBLES];

// Device address of `scratch` on the current device.  In host code the array's name is
// a stand-in, not a device address, so it is never passed to a kernel or a copy: each
// device's address comes from hipGetSymbolAddress, taken once per device with that
// device current, and kept.
static std::mutex g_scratch_mu;
static std::vector<double *> g_scratch_ptr;   // per device ordinal

static int scratch_ptr(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr.empty()) {
    int n = 0;
    e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    g_scratch_ptr.assign(n > 0 ? n : 1, nullptr);
  }
  if (dev < 0 || dev >= (int)g_scratch_ptr.size()) {
    set_error("agx_history_sum_noalloc: device ordinal %d out of range", dev);
    return -1;
  }
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace agx

using namespace agx;

extern "C" int agx_history_sum_noalloc(const double *u, const double *bcc, const double *phi, int32_t nblocks,
                                       int32_t nhydro, int32_t nx1, int32_t nx2, int32_t nx3, int32_t ng,
                                       const double *vol, double *hst, void *stream) {
  int rc = history_check_dense(u, bcc, nblocks, nhydro, nx1, nx2, nx3, ng, vol, hst, "agx_history_sum_noalloc");
  if (rc) return rc;
  double *part = nullptr;
  rc = scratch_ptr(&part);
  if (rc) return rc;
  return history_dense(u, bcc, phi, nblocks, nhydro, nx1, nx2, nx3, ng, vol, part, hst,
                       static_cast<hipStream_t>(stream));
}
