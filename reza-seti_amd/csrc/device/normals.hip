// Near-wall normals on the GPU: SubDomain::calculateBoundaryNormals
// (Subdomain.cpp:499-572) with one thread per near-wall point.
//
// The per-triangle terms (unit normal, area, centroid) depend only on the
// triangle; they are computed once on the host with the reference's
// expressions and passed in. Every thread then visits the triangles in file
// order and accumulates exactly the reference's sum, so the result is bit
// identical to the host loop (FP64, no contraction, IEEE division and sqrt).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../host/normals.h"

namespace lbm {

namespace {

constexpr int NB = 256;
constexpr int TILE = 256;

__global__ __launch_bounds__(NB) void normals_kernel(const double* __restrict__ pts, int64_t n,
                                                     const TriTerms* __restrict__ tt, int64_t nt,
                                                     double* __restrict__ out) {
  __shared__ TriTerms tile[TILE];
  const int64_t i = int64_t(blockIdx.x) * NB + threadIdx.x;
  const bool active = i < n;
  const double x = active ? pts[3 * i] : 0.0, y = active ? pts[3 * i + 1] : 0.0, z = active ? pts[3 * i + 2] : 0.0;
  double n0 = 0.0, n1 = 0.0, n2 = 0.0;
  for (int64_t base = 0; base < nt; base += TILE) {
    const int64_t cnt = (nt - base) < TILE ? (nt - base) : TILE;
    __syncthreads();
    if (threadIdx.x < cnt) tile[threadIdx.x] = tt[base + threadIdx.x];
    __syncthreads();
    if (active) {
      for (int t = 0; t < cnt; ++t) {
        const TriTerms& T = tile[t];
        const double dx = x - T.cx, dy = y - T.cy, dz = z - T.cz;
        const double distance = sqrt(dx * dx + dy * dy + dz * dz);
        const double weight = T.area / (distance * distance);
        n0 += T.nx * weight;
        n1 += T.ny * weight;
        n2 += T.nz * weight;
      }
    }
  }
  if (!active) return;
  const double norm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  if (norm > 0) {
    n0 /= norm;
    n1 /= norm;
    n2 /= norm;
  }
  out[3 * i] = n0;
  out[3 * i + 1] = n1;
  out[3 * i + 2] = n2;
}

}  // namespace

bool device_boundary_normals(const std::vector<double>& pts, const std::vector<TriTerms>& tt,
                             std::vector<double>& out, int device, std::string* err) {
  const int64_t n = static_cast<int64_t>(pts.size() / 3), nt = static_cast<int64_t>(tt.size());
  out.assign(pts.size(), 0.0);
  if (n == 0) return true;
  double *dp = nullptr, *dout = nullptr;
  TriTerms* dt = nullptr;
  auto fail = [&](const char* what, hipError_t e) {
    if (err) *err = std::string(what) + ": " + hipGetErrorString(e);
    if (dp) (void)hipFree(dp);
    if (dout) (void)hipFree(dout);
    if (dt) (void)hipFree(dt);
    return false;
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return fail("hipSetDevice", e);
  if ((e = hipMalloc(&dp, pts.size() * sizeof(double))) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&dout, pts.size() * sizeof(double))) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&dt, std::max<size_t>(1, tt.size()) * sizeof(TriTerms))) != hipSuccess)
    return fail("hipMalloc", e);
  if ((e = hipMemcpy(dp, pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return fail("hipMemcpy", e);
  if (nt && (e = hipMemcpy(dt, tt.data(), tt.size() * sizeof(TriTerms), hipMemcpyHostToDevice)) != hipSuccess)
    return fail("hipMemcpy", e);
  const unsigned grid = static_cast<unsigned>((n + NB - 1) / NB);
  hipLaunchKernelGGL(normals_kernel, dim3(grid), dim3(NB), 0, 0, dp, n, dt, nt, dout);
  if ((e = hipGetLastError()) != hipSuccess) return fail("normals launch", e);
  if ((e = hipMemcpy(out.data(), dout, pts.size() * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
    return fail("hipMemcpy", e);
  (void)hipFree(dp);
  (void)hipFree(dout);
  (void)hipFree(dt);
  return true;
}

}  // namespace lbm
