// lp_noalloc.hip -- lp_tridag_array_noalloc: the batched Thomas solve of tridag_array
// for callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It takes device arrays only, needs
// no context, and keeps its scratch -- the gam factors of every mode and the zero-pivot
// flag -- in one file-scope device array of fixed size, resident on every device the
// library runs on until the process ends.
//
// The array's host-side name is not a device address: its address on the current device
// comes from hipGetSymbolAddress, taken once per device and kept; only that address is
// passed to kernels and fills.  Calls on one device share the array, so calls that
// overlap on one device race (each call clears the flag before its solve).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "../../include/lesgo_amd.h"
#include "lp_internal.hpp"
#include "lp_kernels.hpp"

namespace {

// 256 bytes for the flag, then gam[(nz+1) lh ny]: 256^3 (lh 129, ny 256, nz+1 258) fits
constexpr size_t kScratchDoubles = (size_t)9 << 20;  // 72 MiB
constexpr size_t kFlagDoubles = 32;                  // 256 bytes ahead of gam
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
std::map<int, double *> g_ptr;  // device -> address of `scratch` on that device

// its address on the current device (the first call on a device looks it up)
hipError_t scratch_ptr(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_ptr.find(dev);
  if (it != g_ptr.end()) {
    *out = it->second;
    return hipSuccess;
  }
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[dev] = (double *)p;
  *out = (double *)p;
  return hipSuccess;
}

}  // namespace

extern "C" int lp_tridag_array_noalloc(int32_t lh, int32_t ny, int32_t nz, const double *a, const double *b,
                                       const double *c, const double *r, double *u, void *stream) {
  if (lh < 2 || ny < 2 || nz < 1 || !a || !b || !c || !r || !u)
    return lp::set_error(LP_ERR_ARG, "lp_tridag_array_noalloc: bad argument");
  const size_t need = (size_t)lh * ny * (size_t)(nz + 1);
  if (need > kScratchDoubles - kFlagDoubles)
    return lp::set_error(LP_ERR_ARG,
                         "lp_tridag_array_noalloc: lh*ny*(nz+1) = %zu exceeds its fixed scratch of %zu values",
                         need, kScratchDoubles - kFlagDoubles);
  double *base = nullptr;
  hipError_t e = scratch_ptr(&base);
  if (e != hipSuccess)
    return lp::set_error(LP_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  int *flag = reinterpret_cast<int *>(base);
  double *gam = base + kFlagDoubles;
  const long nh = (long)lh * ny;
  hipLaunchKernelGGL(lp::k_fill<int>, dim3(1), dim3(64), 0, s, flag, 1L, 0);  // a fill kernel, no memset
  hipLaunchKernelGGL(lp::k_tridag_arrays, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, a, b, c,
                     (const lp::cx<double> *)r, (lp::cx<double> *)u, gam, lh, ny, nz + 1, flag);
  e = hipGetLastError();
  int h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return lp::set_error(LP_ERR_HIP, "lp_tridag_array_noalloc: %s", hipGetErrorString(e));
  if (h) return lp::set_error(LP_ERR_PIVOT, "tridag_array: zero pivot");
  return LP_OK;
}
