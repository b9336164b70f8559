// Node classification and connectivity on the GPU, from the grid labels:
//   boundary points    Geometry::findBoundaryPoints (ComplexGeometry.cpp:517-564):
//                      a fluid point with a non-fluid point among its 26
//                      neighbours, or on a face of the box
//   node kinds         SubDomain::CreateNodes (Subdomain.cpp:364-442) with the
//                      plane tests of Geometry::getInOutBP
//                      (ComplexGeometry.cpp:567-662): inlet, then the first
//                      matching outlet block, then wall
//   node order         Subdomain.cpp:598-611: grid order, i outer, k inner
//   push neighbours    SubDomain::Connectivity (Subdomain.cpp:447-471)
// The host chain (geometry.cpp build_from_labels) computes the same tables;
// the plane tests evaluate the same binary64 expressions (-ffp-contract=off
// on both sides). The host keeps the creation-order lists, the nodes behind
// inlets/outlets and the near-wall set, which it derives from these tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../host/classify.h"
#include "../lattice.h"

namespace lbm {

namespace {

constexpr int CB = 256;
constexpr int MAX_PLANES = 32;

struct DevPlane {
  int32_t axis;  // 0 x, 1 y, 2 z, -1 none
  double coordinate;
  double rx[2], ry[2], rz[2];
};

struct ClassArgs {
  const uint8_t* fluid;
  int32_t NX, NY, NZ;
  int32_t periodic[3];
  double min_x, min_y, min_z, spacing;
  int32_t n_in, n_out;
  const DevPlane* planes;  // inlets, then outlets
  const int32_t* col_base; // first node of each (i, j) column
  int32_t* node_of;        // per grid point, -1 for non-fluid
  int32_t* ijk;            // 3 per node
  uint8_t* kind;           // per node
  int32_t* outlet_id;      // per node
  int32_t* bp_count;       // boundary points per column
};

__device__ inline int wrap(int v, int n, int per, bool& ok) {
  if (v >= 0 && v < n) return v;
  if (per) return (v + n) % n;
  ok = false;
  return v;
}

// Geometry::getInOutBP's plane/range test (ComplexGeometry.cpp:583-597).
__device__ inline bool in_plane(const DevPlane& s, double x, double y, double z) {
  const double tolerance = 2.5e-2;
  if (s.axis == 0 && fabs(x - s.coordinate) <= tolerance)
    return y >= s.ry[0] && y <= s.ry[1] && z >= s.rz[0] && z <= s.rz[1];
  if (s.axis == 1 && fabs(y - s.coordinate) <= tolerance)
    return x >= s.rx[0] && x <= s.rx[1] && z >= s.rz[0] && z <= s.rz[1];
  if (s.axis == 2 && fabs(z - s.coordinate) <= tolerance)
    return x >= s.rx[0] && x <= s.rx[1] && y >= s.ry[0] && y <= s.ry[1];
  return false;
}

// Fluid points per (i, j) column: one block per column.
__global__ __launch_bounds__(CB) void column_count_kernel(const uint8_t* fluid, int32_t NZ, int32_t* count) {
  const int64_t c = blockIdx.x;
  const uint8_t* col = fluid + c * NZ;
  int32_t s = 0;
  for (int32_t k = threadIdx.x; k < NZ; k += CB) s += col[k] ? 1 : 0;
  __shared__ int32_t red[CB];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = CB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) count[c] = red[0];
}

// Numbers the fluid points of one column in k order (block scan over
// chunks of CB points) and classifies them.
__global__ __launch_bounds__(CB) void column_nodes_kernel(ClassArgs a) {
  const int64_t c = blockIdx.x;
  const int32_t i = static_cast<int32_t>(c / a.NY), j = static_cast<int32_t>(c - int64_t(i) * a.NY);
  const int64_t base_g = c * a.NZ;
  __shared__ int32_t scan[CB];
  int32_t running = a.col_base[c];
  int32_t nbp = 0;
  for (int32_t k0 = 0; k0 < a.NZ; k0 += CB) {
    const int32_t k = k0 + threadIdx.x;
    const bool f = k < a.NZ && a.fluid[base_g + k];
    scan[threadIdx.x] = f ? 1 : 0;
    __syncthreads();
    // inclusive scan (Hillis-Steele)
    for (int off = 1; off < CB; off <<= 1) {
      const int32_t v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
      __syncthreads();
      scan[threadIdx.x] += v;
      __syncthreads();
    }
    const int32_t incl = scan[threadIdx.x], total = scan[CB - 1];
    if (k < a.NZ) a.node_of[base_g + k] = f ? running + incl - 1 : -1;
    if (f) {
      const int32_t n = running + incl - 1;
      a.ijk[3 * int64_t(n)] = i;
      a.ijk[3 * int64_t(n) + 1] = j;
      a.ijk[3 * int64_t(n) + 2] = k;
      bool bp = (!a.periodic[0] && (i == 0 || i == a.NX - 1)) || (!a.periodic[1] && (j == 0 || j == a.NY - 1)) ||
                (!a.periodic[2] && (k == 0 || k == a.NZ - 1));
      for (int dx = -1; dx <= 1 && !bp; ++dx)
        for (int dy = -1; dy <= 1 && !bp; ++dy)
          for (int dz = -1; dz <= 1 && !bp; ++dz) {
            bool ok = true;
            const int x = wrap(i + dx, a.NX, a.periodic[0], ok), y = wrap(j + dy, a.NY, a.periodic[1], ok),
                      z = wrap(k + dz, a.NZ, a.periodic[2], ok);
            if (ok && !a.fluid[(int64_t(x) * a.NY + y) * a.NZ + z]) bp = true;
          }
      uint8_t kd = 0;
      int32_t oid = -1;
      if (bp) {
        ++nbp;
        // GridSpec::px/py/pz, term by term.
        const double px = a.min_x + i * a.spacing, py = a.min_y + j * a.spacing, pz = a.min_z + k * a.spacing;
        bool inl = false;
        for (int s = 0; s < a.n_in; ++s) inl = inl || in_plane(a.planes[s], px, py, pz);
        if (inl) {
          kd = 2;
        } else {
          for (int s = 0; s < a.n_out && oid < 0; ++s)
            if (in_plane(a.planes[a.n_in + s], px, py, pz)) oid = s;
          kd = oid >= 0 ? 3 : 1;
        }
      }
      a.kind[n] = kd;
      a.outlet_id[n] = oid;
    }
    running += total;
    __syncthreads();
  }
  __shared__ int32_t red[CB];
  red[threadIdx.x] = nbp;
  __syncthreads();
  for (int w = CB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.bp_count[c] = red[0];
}

struct NbArgs {
  const int32_t* ijk;
  const int32_t* node_of;
  int64_t N;
  int32_t NX, NY, NZ;
  int32_t periodic[3];
  int32_t Q;
  int32_t cx[27], cy[27], cz[27];
  int32_t* nb;  // Q * N
};

// SubDomain::Connectivity: node at x + c_q, or -1.
__global__ __launch_bounds__(CB) void neighbours_kernel(NbArgs a) {
  const int64_t stride = int64_t(gridDim.x) * CB;
  for (int64_t n = int64_t(blockIdx.x) * CB + threadIdx.x; n < a.N; n += stride) {
    const int i = a.ijk[3 * n], j = a.ijk[3 * n + 1], k = a.ijk[3 * n + 2];
    for (int q = 0; q < a.Q; ++q) {
      bool ok = true;
      const int x = wrap(i + a.cx[q], a.NX, a.periodic[0], ok), y = wrap(j + a.cy[q], a.NY, a.periodic[1], ok),
                z = wrap(k + a.cz[q], a.NZ, a.periodic[2], ok);
      a.nb[int64_t(q) * a.N + n] = ok ? a.node_of[(int64_t(x) * a.NY + y) * a.NZ + z] : -1;
    }
  }
}

int32_t axis_of(const std::string& plane) { return plane == "x" ? 0 : plane == "y" ? 1 : plane == "z" ? 2 : -1; }

}  // namespace

bool device_classify(const GridSpec& g, const std::vector<uint8_t>& fluid, const std::vector<PlaneSpec>& inlets,
                     const std::vector<PlaneSpec>& outlets, int Q, const int periodic[3], int device,
                     DeviceTables& out, std::string* err) {
  const int64_t ncol = int64_t(g.NX) * g.NY, P = g.npoints();
  if (static_cast<int64_t>(fluid.size()) != P) {
    if (err) *err = "device classifier: label count does not match the grid";
    return false;
  }
  if (inlets.size() + outlets.size() > MAX_PLANES) {
    if (err) *err = "device classifier: more than 32 inlet and outlet blocks";
    return false;
  }
  if (ncol >= (int64_t(1) << 31)) {
    if (err) *err = "device classifier: more than 2^31 grid columns";
    return false;
  }
  std::vector<DevPlane> planes;
  for (const auto* lst : {&inlets, &outlets})
    for (const PlaneSpec& s : *lst) {
      DevPlane d{};
      d.axis = axis_of(s.plane);
      d.coordinate = s.coordinate;
      for (int e = 0; e < 2; ++e) {
        d.rx[e] = s.range_x[e];
        d.ry[e] = s.range_y[e];
        d.rz[e] = s.range_z[e];
      }
      planes.push_back(d);
    }
  uint8_t* df = nullptr;
  int32_t *dcount = nullptr, *dbase = nullptr, *dnode = nullptr, *dijk = nullptr, *doid = nullptr, *dnb = nullptr,
          *dbp = nullptr;
  uint8_t* dkind = nullptr;
  DevPlane* dpl = nullptr;
  auto release = [&]() {
    void* ptrs[] = {df, dcount, dbase, dnode, dijk, doid, dnb, dbp, dkind, dpl};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  };
  auto fail = [&](const char* what, hipError_t e) {
    if (err) *err = std::string("device classifier: ") + what + ": " + hipGetErrorString(e);
    release();
    return false;
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return fail("hipSetDevice", e);
  if ((e = hipMalloc(&df, std::max<int64_t>(P, 1))) != hipSuccess ||
      (e = hipMalloc(&dcount, ncol * 4)) != hipSuccess || (e = hipMalloc(&dbase, ncol * 4)) != hipSuccess ||
      (e = hipMalloc(&dbp, ncol * 4)) != hipSuccess ||
      (e = hipMalloc(&dpl, std::max<size_t>(planes.size(), 1) * sizeof(DevPlane))) != hipSuccess)
    return fail("hipMalloc", e);
  if ((e = hipMemcpy(df, fluid.data(), P, hipMemcpyHostToDevice)) != hipSuccess ||
      (!planes.empty() &&
       (e = hipMemcpy(dpl, planes.data(), planes.size() * sizeof(DevPlane), hipMemcpyHostToDevice)) != hipSuccess))
    return fail("hipMemcpy", e);
  hipLaunchKernelGGL(column_count_kernel, dim3(static_cast<unsigned>(ncol)), dim3(CB), 0, 0, df, g.NZ, dcount);
  if ((e = hipGetLastError()) != hipSuccess) return fail("count launch", e);
  std::vector<int32_t> count(ncol), base(ncol);
  if ((e = hipMemcpy(count.data(), dcount, ncol * 4, hipMemcpyDeviceToHost)) != hipSuccess) return fail("hipMemcpy", e);
  int64_t N = 0;
  for (int64_t c = 0; c < ncol; ++c) {
    base[c] = static_cast<int32_t>(N);
    N += count[c];
    if (N >= (int64_t(1) << 31) - 1) {
      if (err) *err = "device classifier: more than 2^31-1 nodes";
      release();
      return false;
    }
  }
  out.N = N;
  if ((e = hipMemcpy(dbase, base.data(), ncol * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail("hipMemcpy", e);
  const int64_t N1 = std::max<int64_t>(N, 1);
  if ((e = hipMalloc(&dnode, std::max<int64_t>(P, 1) * 4)) != hipSuccess ||
      (e = hipMalloc(&dijk, 3 * N1 * 4)) != hipSuccess || (e = hipMalloc(&dkind, N1)) != hipSuccess ||
      (e = hipMalloc(&doid, N1 * 4)) != hipSuccess)
    return fail("hipMalloc", e);
  ClassArgs a{};
  a.fluid = df;
  a.NX = g.NX;
  a.NY = g.NY;
  a.NZ = g.NZ;
  for (int d = 0; d < 3; ++d) a.periodic[d] = periodic[d];
  a.min_x = g.bbox.min_x;
  a.min_y = g.bbox.min_y;
  a.min_z = g.bbox.min_z;
  a.spacing = g.spacing;
  a.n_in = static_cast<int32_t>(inlets.size());
  a.n_out = static_cast<int32_t>(outlets.size());
  a.planes = dpl;
  a.col_base = dbase;
  a.node_of = dnode;
  a.ijk = dijk;
  a.kind = dkind;
  a.outlet_id = doid;
  a.bp_count = dbp;
  hipLaunchKernelGGL(column_nodes_kernel, dim3(static_cast<unsigned>(ncol)), dim3(CB), 0, 0, a);
  if ((e = hipGetLastError()) != hipSuccess) return fail("classify launch", e);
  // The labels are not needed any more: make room for the push table.
  (void)hipFree(df);
  df = nullptr;
  if ((e = hipMalloc(&dnb, int64_t(Q) * N1 * 4)) != hipSuccess) return fail("hipMalloc", e);
  NbArgs b{};
  b.ijk = dijk;
  b.node_of = dnode;
  b.N = N;
  b.NX = g.NX;
  b.NY = g.NY;
  b.NZ = g.NZ;
  for (int d = 0; d < 3; ++d) b.periodic[d] = periodic[d];
  const LatticeView L = lattice_view(Q);
  b.Q = L.Q;
  for (int q = 0; q < L.Q; ++q) {
    b.cx[q] = L.cx[q];
    b.cy[q] = L.cy[q];
    b.cz[q] = L.cz[q];
  }
  b.nb = dnb;
  if (N > 0) {
    const int64_t want = (N + CB - 1) / CB;
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(want, int64_t(1) << 20));
    hipLaunchKernelGGL(neighbours_kernel, dim3(grid), dim3(CB), 0, 0, b);
    if ((e = hipGetLastError()) != hipSuccess) return fail("neighbour launch", e);
  }
  out.ijk.resize(3 * N);
  out.kind.resize(N);
  out.outlet_id.resize(N);
  out.nb.resize(static_cast<size_t>(Q) * N);
  std::vector<int32_t> bp(ncol);
  if ((N > 0 && ((e = hipMemcpy(out.ijk.data(), dijk, 3 * N * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
                 (e = hipMemcpy(out.kind.data(), dkind, N, hipMemcpyDeviceToHost)) != hipSuccess ||
                 (e = hipMemcpy(out.outlet_id.data(), doid, N * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
                 (e = hipMemcpy(out.nb.data(), dnb, int64_t(Q) * N * 4, hipMemcpyDeviceToHost)) != hipSuccess)) ||
      (e = hipMemcpy(bp.data(), dbp, ncol * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return fail("hipMemcpy", e);
  out.n_boundary_points = 0;
  for (int32_t v : bp) out.n_boundary_points += v;
  release();
  return true;
}

}  // namespace lbm
