// Voxelisation, node classification and the sparse node tables.
//
// Host-side restatement of the reference's setup chain with identical
// floating-point predicates (so classification is bit-exact) but without its
// O(points x triangles) cost:
//   bounding box / grid       ComplexGeometry.cpp:83-151 (pitch doubled, :157)
//   +z ray parity labelling   ComplexGeometry.cpp:208-310, 314-336, 440-514
//   boundary points           ComplexGeometry.cpp:517-564
//   inlet/outlet/wall sets    ComplexGeometry.cpp:567-662
//   node kinds & precedence   Subdomain.cpp:364-442 (+ Utility.cpp:59-74)
//   connectivity              Subdomain.cpp:447-471
//   near-wall set & normals   Subdomain.cpp:475-572
//   node order                Subdomain.cpp:598-611 (i outer, k inner)
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "config.h"
#include "stl.h"

namespace lbm {

enum NodeKind : uint8_t { KIND_FLUID = 0, KIND_WALL = 1, KIND_INLET = 2, KIND_OUTLET = 3 };

struct BBox {
  double min_x, max_x, min_y, max_y, min_z, max_z;
};

struct GridSpec {
  BBox bbox{};
  double spacing = 0.0;  // lattice pitch actually used (2 x grid_size)
  int NX = 0, NY = 0, NZ = 0;  // number of points per axis
  double px(int i) const { return bbox.min_x + i * spacing; }
  double py(int j) const { return bbox.min_y + j * spacing; }
  double pz(int k) const { return bbox.min_z + k * spacing; }
  int64_t npoints() const { return int64_t(NX) * NY * NZ; }
  int64_t gidx(int i, int j, int k) const { return (int64_t(i) * NY + j) * NZ + k; }
};

BBox bounding_box(const std::vector<Triangle>& tris);
// grid_size is config.txt's value; the pitch used is 2 x grid_size.
GridSpec make_grid(const BBox& bb, double grid_size);

// Ray-parity label per grid point (1 = Fluid), index GridSpec::gidx.
// `verify_stride` > 0 re-evaluates every verify_stride-th column against all
// triangles (no candidate filter) and returns false on any disagreement.
// label_device = 1 evaluates the same ray test on HIP device `device`
// (device/voxel.hip) over the same candidate lists: the same labels.
bool label_points(const GridSpec& g, const std::vector<Triangle>& tris,
                  std::vector<uint8_t>& fluid, int nthreads, int verify_stride,
                  std::string* err, int label_device = 0, int device = 0);

// One (point, triangle) evaluation of the reference's ray test; exposed for
// tests. (x, y, z) is the grid point before the bounding-box offsets.
bool ray_counts_triangle(const GridSpec& g, double x, double y, double z, const Triangle& t);

struct BuildOptions {
  int Q = 19;
  int periodic[3] = {0, 0, 0};  // extension: wrap neighbours along an axis
  int nthreads = 0;              // 0 = hardware concurrency
  int verify_stride = 0;         // see label_points
  int compute_normals = 1;       // 0 none, 1 host threads, 2 GPU (same sum)
  int device = 0;
  int label_device = 0;          // ray labels: 0 host threads, 1 GPU (same labels)
  int classify_device = 0;       // boundary points, kinds, order, push table: 0 host, 1 GPU (same tables)
};

struct Tables {
  GridSpec grid;
  int Q = 19;
  int periodic[3] = {0, 0, 0};
  int64_t N = 0;                   // nodes (fluid + wall + inlet + outlet)
  std::vector<int32_t> ijk;        // 3N, reference node order
  std::vector<uint8_t> kind;       // N, NodeKind
  std::vector<int32_t> outlet_id;  // N, outlet block index or -1
  std::vector<int32_t> inlet_nodes, wall_nodes;  // creation order
  std::vector<std::vector<int32_t>> outlet_nodes;
  std::array<double, 3> inlet_normal{0, 0, 0};
  bool has_inlet_normal = false;
  std::vector<std::array<double, 3>> outlet_normals;
  std::vector<int32_t> nb;         // Q*N push table: node at x + c_q, or -1
  std::vector<int32_t> bc_nbr;     // N: node at x - n_hat for inlet/outlet nodes, else -1
  std::vector<int32_t> nearwall;   // ascending node indices
  std::vector<double> normals;     // 3 * nearwall.size()
  std::vector<uint8_t> fluid;      // per grid point label (kept for tests)
  int64_t n_boundary_points = 0;
  int64_t upstream_out_of_box = 0; // outlet upstream lookups outside the grid
};

// Full chain from a parsed config (stl_file relative to base_dir).
bool build_from_config(const SimulationParameters& p, const std::string& base_dir,
                       const BuildOptions& o, Tables& t, std::vector<Triangle>* tris_out,
                       std::string* err);

// Chain from an explicit fluid label array (skips STL + ray casting).
bool build_from_labels(const GridSpec& g, const std::vector<uint8_t>& fluid,
                       const std::vector<PlaneSpec>& inlets, const std::vector<PlaneSpec>& outlets,
                       const std::vector<Triangle>& tris, const BuildOptions& o, Tables& t,
                       std::string* err);

// Node tables given directly (a reference SubDomain's), checked; derives the
// node lists and the inlet/outlet neighbour behind each boundary node.
struct TableInput {
  int Q = 19;
  int64_t N = 0;
  int dims[3] = {0, 0, 0};
  double origin[3] = {0, 0, 0};
  double spacing = 0.0;
  const int32_t* ijk = nullptr;
  const uint8_t* kind = nullptr;
  const int32_t* outlet_id = nullptr;
  const int32_t* nb = nullptr;
  bool has_inlet_normal = false;
  double inlet_normal[3] = {0, 0, 0};
  int n_outlets = 0;
  const double* outlet_normals = nullptr;
  int64_t n_nearwall = 0;
  const int32_t* nearwall = nullptr;
  const double* normals = nullptr;
};
bool build_from_tables(const TableInput& in, Tables& t, std::string* err);

// Near-wall normal of one point (Subdomain.cpp:519-570), exposed for tests.
std::array<double, 3> boundary_normal(double x, double y, double z, const std::vector<Triangle>& tris);

}  // namespace lbm
