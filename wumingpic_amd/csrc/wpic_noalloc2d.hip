// wpic_noalloc2d.hip -- wpic_gauss2d_noalloc: the Gauss residual of
// wpic_gauss2d computed without any device allocation, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// wpic_gauss2d allocates, on its first call, a device buffer for the charge
// density plane and the residual it copies back.  Here both live in the file
// scope array `scratch` below, of a fixed maximum size that every GPU the
// library runs on holds until the process ends.  Host code never uses the
// array's name as a device address (in host code it is a stand-in symbol):
// the first call on a device looks up that device's copy with
// hipGetSymbolAddress, with the device current, and keeps the address per
// device; kernels receive that address as an argument.
#include <mutex>

#include "wpic_internal.h"

namespace wpic {
namespace {

constexpr size_t SCRATCH_DOUBLES = size_t(5) << 20;  // 40 MiB per device
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAXDEV = 64;
std::mutex g_mu;
double* g_ptr[MAXDEV];  // this device's copy of `scratch`, or null before the first call there

// the device address of `scratch` on `dev` (looked up once per device)
hipError_t scratch_on(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (dev < 0 || dev >= MAXDEV) return hipErrorInvalidDevice;
  if (!g_ptr[dev]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) {
      const hipError_t e2 = hipSetDevice(cur);
      if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

}  // namespace
}  // namespace wpic

using namespace wpic;

extern "C" int wpic_gauss2d_noalloc(wpic_ctx* c, double* G) {
  if (!c || !G) return set_error(WPIC_EINVAL, "null argument");
  const size_t pl = (size_t)c->g.fw * c->g.fh, nc = (size_t)c->g.nx * c->g.ny;
  if (pl + nc > SCRATCH_DOUBLES)
    return set_error(WPIC_EINVAL, "wpic_gauss2d_noalloc: (nx+4)*(ny+4) + nx*ny = " + std::to_string(pl + nc) +
                                      " doubles exceeds the fixed scratch of " + std::to_string(SCRATCH_DOUBLES) +
                                      " (use wpic_gauss2d)");
  double* s = nullptr;
  hipError_t e = scratch_on(c->device, &s);
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  // rho (cleared by a fill kernel inside launch_gauss) then G, as wpic_gauss2d's buffer
  e = launch_gauss(c, s + pl, s);
  if (e != hipSuccess) return hip_fail(e, "launch_gauss");
  e = hipMemcpyAsync(G, s + pl, sizeof(double) * nc, hipMemcpyDeviceToHost, c->st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
  e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  return 0;
}
