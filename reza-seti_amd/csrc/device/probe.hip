// Plain device copy: the achievable-bandwidth reference for the roofline and
// the calibration kernel for HBM traffic counters (a launch of known bytes).
#include <hip/hip_runtime.h>

#include <string>

#include "../capi_internal.h"

namespace lbm {
namespace {

constexpr int PB = 256;

// One 16-byte element per thread: on gfx950 this out-copies a grid-stride
// loop (6.2 vs 4.5 TB/s for 1 GiB), whose threads hold one load in flight.
__global__ __launch_bounds__(PB) void copy_probe_kernel(const float4* __restrict__ in, float4* __restrict__ out,
                                                        int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * PB + threadIdx.x;
  if (i < n) out[i] = in[i];
}

}  // namespace
}  // namespace lbm

extern "C" int lbm_copy_probe(int32_t device, int64_t bytes, int32_t iters, double* gbps) {
  using namespace lbm;
  if (!gbps || bytes < 16 || iters < 1) return fail(LBM_ERR_ARG, "invalid argument");
  const int64_t n = bytes / 16;
  if ((n + PB - 1) / PB >= (int64_t(1) << 31)) return fail(LBM_ERR_ARG, "copy probe size too large");
  float4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LBM_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == LBM_OK) rc = fail(LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return e == hipSuccess;
  };
  if (chk(hipSetDevice(device), "hipSetDevice") && chk(hipMalloc(&a, n * 16), "hipMalloc") &&
      chk(hipMalloc(&b, n * 16), "hipMalloc") && chk(hipMemset(a, 0, n * 16), "hipMemset") &&
      chk(hipEventCreate(&e0), "hipEventCreate") && chk(hipEventCreate(&e1), "hipEventCreate")) {
    const unsigned grid = static_cast<unsigned>((n + PB - 1) / PB);
    hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(PB), 0, 0, a, b, n);  // warm-up
    chk(hipEventRecord(e0, 0), "hipEventRecord");
    for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(PB), 0, 0, a, b, n);
    chk(hipGetLastError(), "copy launch");
    chk(hipEventRecord(e1, 0), "hipEventRecord");
    if (chk(hipEventSynchronize(e1), "hipEventSynchronize")) {
      float ms = 0.0f;
      chk(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      *gbps = 2.0 * double(n * 16) * iters / (double(ms) * 1e-3) / 1e9;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}
