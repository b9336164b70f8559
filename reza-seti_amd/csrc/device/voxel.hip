// Ray-parity voxeliser on the GPU: Geometry::labelPoints with
// countIntersectinsTriangle (ComplexGeometry.cpp:265-310, 440-500), the same
// +z ray test as the host voxeliser (host/raytest.h), one thread per grid
// point over its column's candidate triangles (the host's candidate lists,
// in the same order). Labels are identical to the host's.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../host/raytest.h"
#include "../host/voxel.h"

namespace lbm {

namespace {

constexpr int VB = 256;

struct ColumnGrid {
  double min_z, max_z, spacing, ray_length;
  int64_t NY, NZ, npoints;
};

__global__ __launch_bounds__(VB) void label_kernel(const Triangle* __restrict__ tris, const int32_t* __restrict__ cand,
                                                  const int64_t* __restrict__ start, const double* __restrict__ rx,
                                                  const double* __restrict__ ry, ColumnGrid g,
                                                  uint8_t* __restrict__ fluid) {
  const int64_t stride = int64_t(gridDim.x) * VB;
  for (int64_t gi = int64_t(blockIdx.x) * VB + threadIdx.x; gi < g.npoints; gi += stride) {
    const int64_t c = gi / g.NZ, k = gi - c * g.NZ;
    const int64_t i = c / g.NY, j = c - i * g.NY;
    const double x = rx[i], y = ry[j];
    // GridSpec::pz and the ray ends of the host loop, term by term.
    double sz = g.min_z + double(k) * g.spacing;
    double ez = g.min_z + double(k) * g.spacing + g.ray_length + 0.1;
    if (sz == g.min_z) sz += 0.01;
    if (ez == g.max_z) ez -= 0.01;
    int count = 0;
    for (int64_t e = start[c]; e < start[c + 1]; ++e)
      count += segment_counts(x, y, sz, x, y, ez, tris[cand[e]]) ? 1 : 0;
    fluid[gi] = (count % 2 == 1) ? 1 : 0;
  }
}

}  // namespace

bool device_label_points(const std::vector<Triangle>& tris, const std::vector<int32_t>& cand,
                         const std::vector<int64_t>& start, const std::vector<double>& rx,
                         const std::vector<double>& ry, double min_z, double max_z, double spacing, int NZ,
                         std::vector<uint8_t>& fluid, int device, std::string* err) {
  const int64_t NX = static_cast<int64_t>(rx.size()), NY = static_cast<int64_t>(ry.size());
  const int64_t P = NX * NY * int64_t(NZ);
  fluid.assign(static_cast<size_t>(P), 0);
  if (P == 0) return true;
  if (static_cast<int64_t>(start.size()) != NX * NY + 1) {
    if (err) *err = "device voxeliser: candidate offsets do not match the grid";
    return false;
  }
  Triangle* dt = nullptr;
  int32_t* dc = nullptr;
  int64_t* ds = nullptr;
  double *drx = nullptr, *dry = nullptr;
  uint8_t* df = nullptr;
  auto release = [&]() {
    void* ptrs[] = {dt, dc, ds, drx, dry, df};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  };
  auto fail = [&](const char* what, hipError_t e) {
    if (err) *err = std::string("device voxeliser: ") + what + ": " + hipGetErrorString(e);
    release();
    return false;
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return fail("hipSetDevice", e);
  if ((e = hipMalloc(&dt, std::max<size_t>(tris.size(), 1) * sizeof(Triangle))) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&dc, std::max<size_t>(cand.size(), 1) * 4)) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&ds, start.size() * 8)) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&drx, NX * 8)) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&dry, NY * 8)) != hipSuccess) return fail("hipMalloc", e);
  if ((e = hipMalloc(&df, static_cast<size_t>(P))) != hipSuccess) return fail("hipMalloc", e);
  if ((!tris.empty() && (e = hipMemcpy(dt, tris.data(), tris.size() * sizeof(Triangle), hipMemcpyHostToDevice))) ||
      (!cand.empty() && (e = hipMemcpy(dc, cand.data(), cand.size() * 4, hipMemcpyHostToDevice))) ||
      (e = hipMemcpy(ds, start.data(), start.size() * 8, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(drx, rx.data(), NX * 8, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(dry, ry.data(), NY * 8, hipMemcpyHostToDevice)))
    return fail("hipMemcpy", e);
  const ColumnGrid g{min_z, max_z, spacing, max_z - min_z, NY, int64_t(NZ), P};
  const int64_t want = (P + VB - 1) / VB;
  const unsigned grid = static_cast<unsigned>(want < (int64_t(1) << 20) ? want : (int64_t(1) << 20));
  hipLaunchKernelGGL(label_kernel, dim3(grid), dim3(VB), 0, 0, dt, dc, ds, drx, dry, g, df);
  if ((e = hipGetLastError()) != hipSuccess) return fail("label launch", e);
  if ((e = hipMemcpy(fluid.data(), df, static_cast<size_t>(P), hipMemcpyDeviceToHost)) != hipSuccess)
    return fail("hipMemcpy", e);
  release();
  return true;
}

}  // namespace lbm
