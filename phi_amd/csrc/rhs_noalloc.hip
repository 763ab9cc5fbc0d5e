// phi_hip_rhs_noalloc: the batched right-hand side of phi_hip_rhs without
// any device allocation, for callers that cannot allocate (code under HIP
// graph capture, serving loops).  Its input and output are staged in a
// fixed device array, `scratch`, instead of buffers allocated per call.
//
// Each device holds its own copy of `scratch` (kScratchBytes) from the
// first call on that device until the process ends.  Its address on a
// device comes from hipGetSymbolAddress, looked up once per device with that
// device current, and only that address reaches kernels and copies.  Every
// call on a device shares that device's array: calls that overlap on one
// device race (the caller serialises them).  The kernels write every output
// element, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <complex>
#include <mutex>

#include "capi_internal.h"

namespace {

// Input and output halves of up to kScratchComplex complex doubles each.
constexpr size_t kScratchComplex = (size_t)1 << 20;
constexpr size_t kScratchBytes = 2 * kScratchComplex * sizeof(double2);  // 32 MiB

__device__ __align__(256) double2 scratch[2 * kScratchComplex];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double2* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (looked up on the first
// call there, with that device current).
int ScratchOf(int dev, double2** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return phi::capi::Err(PHI_HIP_EINVAL, "device id beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess)
      return phi::capi::Err(PHI_HIP_EHIP, "hipGetDevice failed");
    if (cur != dev && hipSetDevice(dev) != hipSuccess)
      return phi::capi::Err(PHI_HIP_EHIP, "hipSetDevice failed");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return phi::capi::Err(PHI_HIP_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                                 hipGetErrorString(e));
    g_ptr[dev] = static_cast<double2*>(p);
  }
  *out = g_ptr[dev];
  return PHI_HIP_OK;
}

}  // namespace

extern "C" int phi_hip_rhs_noalloc(phi_ctx* c, const double* in, double* out, size_t n) {
  return phi::capi::Guarded(__func__, [&]() -> int {
    if (!c || !in || !out) return phi::capi::Err(PHI_HIP_EINVAL, "null argument");
    if (n != c->e->state_elems()) return phi::capi::Err(PHI_HIP_EINVAL, "state size mismatch");
    if (n > kScratchComplex)
      return phi::capi::Err(PHI_HIP_EINVAL,
                            "state of " + std::to_string(n) +
                                " complex elements exceeds the allocation-free staging of " +
                                std::to_string(kScratchComplex) + " (use phi_hip_rhs)");
    double2* s = nullptr;
    int rc = ScratchOf(c->e->device(), &s);
    if (rc) return rc;
    rc = c->e->RhsStaged(reinterpret_cast<const std::complex<double>*>(in),
                         reinterpret_cast<std::complex<double>*>(out),
                         reinterpret_cast<double*>(s),
                         reinterpret_cast<double*>(s + kScratchComplex));
    return rc ? phi::capi::Err(rc, c->e->last_error()) : phi::capi::Ok();
  });
}
