// Voxeliser, classifier and table builder (see geometry.h for the reference
// map). All floating-point predicates are written with the reference's
// operation order; this file must be compiled with -ffp-contract=off and
// without fast-math so that the host evaluates them exactly like a plain
// IEEE-754 binary64 build of the reference.
#include "geometry.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <thread>
#include <unordered_set>

#include "../lattice.h"
#include "classify.h"
#include "normals.h"
#include "parallel.h"
#include "raytest.h"
#include "voxel.h"

namespace lbm {

namespace {

// Ray end points of labelPoints case 0 plus adjustLineCoordinates case 0
// (ComplexGeometry.cpp:314-336, 463-468).
struct Ray {
  double sx, sy, sz, ex, ey, ez;
};

inline double adjust_xy(double v, double lo, double hi) {
  const double off = 0.01;
  if (v == lo) v += off;
  if (v == hi) v -= off;
  return v;
}

inline Ray make_ray(const GridSpec& g, double x, double y, double z) {
  const BBox& b = g.bbox;
  const double rayLength = b.max_z - b.min_z;
  Ray r;
  r.sx = adjust_xy(x, b.min_x, b.max_x);
  r.sy = adjust_xy(y, b.min_y, b.max_y);
  r.ex = r.sx;
  r.ey = r.sy;
  r.sz = z;
  r.ez = z + rayLength + 0.1;
  if (r.sz == b.min_z) r.sz += 0.01;
  if (r.ez == b.max_z) r.ez -= 0.01;
  return r;
}

// A triangle that can never count: zero file-normal z (denominator is exactly
// zero for a +z ray) or a degenerate vertex set (rejected first thing in the
// inside test).
bool never_counts(const Triangle& T) {
  if (T.nz == 0.0) return true;
  const double uX = T.x2 - T.x1, uY = T.y2 - T.y1, uZ = T.z2 - T.z1;
  const double vX = T.x3 - T.x1, vY = T.y3 - T.y1, vZ = T.z3 - T.z1;
  const double nX = uY * vZ - uZ * vY, nY = uZ * vX - uX * vZ, nZ = uX * vY - uY * vX;
  const double l2 = nX * nX + nY * nY + nZ * nZ;
  if (l2 < 1e-16 * 1e-16) return true;
  if (std::fabs(nZ) < 1e-16) return true;
  return false;
}

}  // namespace

bool ray_counts_triangle(const GridSpec& g, double x, double y, double z, const Triangle& t) {
  const Ray r = make_ray(g, x, y, z);
  return segment_counts(r.sx, r.sy, r.sz, r.ex, r.ey, r.ez, t);
}

BBox bounding_box(const std::vector<Triangle>& tris) {
  // Geometry::GetBoundingbox (ComplexGeometry.cpp:90-104), same min/max nesting.
  const double inf = std::numeric_limits<double>::infinity();
  double mnx = inf, mny = inf, mnz = inf, mxx = -inf, mxy = -inf, mxz = -inf;
  for (const Triangle& t : tris) {
    mnx = std::min(mnx, std::min(t.x1, std::min(t.x2, t.x3)));
    mny = std::min(mny, std::min(t.y1, std::min(t.y2, t.y3)));
    mnz = std::min(mnz, std::min(t.z1, std::min(t.z2, t.z3)));
    mxx = std::max(mxx, std::max(t.x1, std::max(t.x2, t.x3)));
    mxy = std::max(mxy, std::max(t.y1, std::max(t.y2, t.y3)));
    mxz = std::max(mxz, std::max(t.z1, std::max(t.z2, t.z3)));
  }
  return {mnx, mxx, mny, mxy, mnz, mxz};
}

GridSpec make_grid(const BBox& bb, double grid_size) {
  // DistributeTriangles doubles the spacing (ComplexGeometry.cpp:157) before
  // generatePointsInBoundingBox derives nx = ceil(range / spacing) (:124-128).
  GridSpec g;
  g.bbox = bb;
  double spacing = grid_size;
  spacing *= 2;
  g.spacing = spacing;
  const int nx = static_cast<int>(std::ceil((bb.max_x - bb.min_x) / spacing));
  const int ny = static_cast<int>(std::ceil((bb.max_y - bb.min_y) / spacing));
  const int nz = static_cast<int>(std::ceil((bb.max_z - bb.min_z) / spacing));
  g.NX = nx + 1;
  g.NY = ny + 1;
  g.NZ = nz + 1;
  return g;
}

bool label_points(const GridSpec& g, const std::vector<Triangle>& tris, std::vector<uint8_t>& fluid,
                  int nthreads, int verify_stride, std::string* err, int label_device, int device) {
  const int NX = g.NX, NY = g.NY, NZ = g.NZ;
  if (NX <= 0 || NY <= 0 || NZ <= 0) {
    if (err) *err = "empty grid";
    return false;
  }
  fluid.assign(static_cast<size_t>(g.npoints()), 0);

  // Ray x (resp. y) of every column, after the bounding-box offsets.
  std::vector<double> rx(NX), ry(NY);
  for (int i = 0; i < NX; ++i) rx[i] = adjust_xy(g.px(i), g.bbox.min_x, g.bbox.max_x);
  for (int j = 0; j < NY; ++j) ry[j] = adjust_xy(g.py(j), g.bbox.min_y, g.bbox.max_y);
  std::vector<std::pair<double, int>> sx(NX), sy(NY);
  for (int i = 0; i < NX; ++i) sx[i] = {rx[i], i};
  for (int j = 0; j < NY; ++j) sy[j] = {ry[j], j};
  std::sort(sx.begin(), sx.end());
  std::sort(sy.begin(), sy.end());

  // Candidate filter: a triangle can only count for a ray whose (x, y) lies in
  // the triangle's xy bounding box. The box is widened by a margin far above
  // the rounding error of the barycentric test, so no counting pair is lost.
  auto cols_of = [&](const Triangle& T, std::vector<int>& is, std::vector<int>& js) {
    is.clear();
    js.clear();
    const double bx0 = std::min(T.x1, std::min(T.x2, T.x3)), bx1 = std::max(T.x1, std::max(T.x2, T.x3));
    const double by0 = std::min(T.y1, std::min(T.y2, T.y3)), by1 = std::max(T.y1, std::max(T.y2, T.y3));
    const double ext = std::max(bx1 - bx0, by1 - by0);
    const double scale = std::max({std::fabs(bx0), std::fabs(bx1), std::fabs(by0), std::fabs(by1), 1.0});
    const double m = 1e-6 * ext + 1e-9 * scale;
    auto lo_x = std::lower_bound(sx.begin(), sx.end(), std::make_pair(bx0 - m, -1));
    for (auto it = lo_x; it != sx.end() && it->first <= bx1 + m; ++it) is.push_back(it->second);
    auto lo_y = std::lower_bound(sy.begin(), sy.end(), std::make_pair(by0 - m, -1));
    for (auto it = lo_y; it != sy.end() && it->first <= by1 + m; ++it) js.push_back(it->second);
  };

  const int64_t ncol = int64_t(NX) * NY;
  std::vector<int64_t> start(ncol + 1, 0);
  {
    std::vector<int> is, js;
    for (const Triangle& T : tris) {
      if (never_counts(T)) continue;
      cols_of(T, is, js);
      for (int i : is)
        for (int j : js) start[int64_t(i) * NY + j + 1]++;
    }
  }
  for (int64_t c = 0; c < ncol; ++c) start[c + 1] += start[c];
  std::vector<int32_t> cand(static_cast<size_t>(start[ncol]));
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    std::vector<int> is, js;
    for (size_t ti = 0; ti < tris.size(); ++ti) {
      const Triangle& T = tris[ti];
      if (never_counts(T)) continue;
      cols_of(T, is, js);
      for (int i : is)
        for (int j : js) cand[fill[int64_t(i) * NY + j]++] = static_cast<int32_t>(ti);
    }
  }

  const double rayLength = g.bbox.max_z - g.bbox.min_z;
  if (label_device) {
    if (!device_label_points(tris, cand, start, rx, ry, g.bbox.min_z, g.bbox.max_z, g.spacing, NZ, fluid, device,
                             err))
      return false;
  } else {
  parallel_for(ncol, nthreads, [&](int64_t c) {
    const int i = static_cast<int>(c / NY), j = static_cast<int>(c % NY);
    const double x = rx[i], y = ry[j];
    const int64_t a = start[c], b = start[c + 1];
    uint8_t* col = fluid.data() + c * NZ;
    for (int k = 0; k < NZ; ++k) {
      double sz = g.pz(k);
      double ez = g.pz(k) + rayLength + 0.1;
      if (sz == g.bbox.min_z) sz += 0.01;
      if (ez == g.bbox.max_z) ez -= 0.01;
      int count = 0;
      for (int64_t e = a; e < b; ++e)
        count += segment_counts(x, y, sz, x, y, ez, tris[cand[e]]) ? 1 : 0;
      col[k] = (count % 2 == 1) ? 1 : 0;
    }
  });
  }

  if (verify_stride > 0) {
    std::atomic<int64_t> bad{0};
    const int64_t nver = (ncol + verify_stride - 1) / verify_stride;
    parallel_for(nver, nthreads, [&](int64_t v) {
      const int64_t c = v * verify_stride;
      const int i = static_cast<int>(c / NY), j = static_cast<int>(c % NY);
      for (int k = 0; k < NZ; ++k) {
        int count = 0;
        for (const Triangle& T : tris)
          count += ray_counts_triangle(g, g.px(i), g.py(j), g.pz(k), T) ? 1 : 0;
        if (((count % 2 == 1) ? 1 : 0) != fluid[c * NZ + k]) bad++;
      }
    });
    if (bad.load() != 0) {
      if (err) *err = "voxeliser candidate filter disagreed with the full scan on " +
                      std::to_string(bad.load()) + " points";
      return false;
    }
  }
  return true;
}

// Per-triangle part of SubDomain::calculateBoundaryNormals (Subdomain.cpp:530-550).
static TriTerms tri_terms(const Triangle& T) {
  const double ux = T.x2 - T.x1, uy = T.y2 - T.y1, uz = T.z2 - T.z1;
  const double vx = T.x3 - T.x1, vy = T.y3 - T.y1, vz = T.z3 - T.z1;
  double nx = uy * vz - uz * vy;
  double ny = uz * vx - ux * vz;
  double nz = ux * vy - uy * vx;
  const double norm = std::sqrt(nx * nx + ny * ny + nz * nz);
  if (norm > 0) {
    nx /= norm;
    ny /= norm;
    nz /= norm;
  }
  TriTerms t;
  t.nx = nx;
  t.ny = ny;
  t.nz = nz;
  t.area = 0.5 * norm;
  t.cx = (T.x1 + T.x2 + T.x3) / 3.0;
  t.cy = (T.y1 + T.y2 + T.y3) / 3.0;
  t.cz = (T.z1 + T.z2 + T.z3) / 3.0;
  t.pad = 0.0;
  return t;
}

// Point part (Subdomain.cpp:551-570). The reference's std::pow(d, 2) is the
// exact square (compilers fold it to d*d).
static std::array<double, 3> normal_from_terms(double x, double y, double z, const std::vector<TriTerms>& tt) {
  double n0 = 0.0, n1 = 0.0, n2 = 0.0;
  for (const TriTerms& T : tt) {
    const double ddx = x - T.cx, ddy = y - T.cy, ddz = z - T.cz;
    const double distance = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    const double weight = T.area / (distance * distance);
    n0 += T.nx * weight;
    n1 += T.ny * weight;
    n2 += T.nz * weight;
  }
  const double norm = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  if (norm > 0) {
    n0 /= norm;
    n1 /= norm;
    n2 /= norm;
  }
  return {n0, n1, n2};
}

std::array<double, 3> boundary_normal(double x, double y, double z, const std::vector<Triangle>& tris) {
  std::vector<TriTerms> tt;
  tt.reserve(tris.size());
  for (const Triangle& T : tris) tt.push_back(tri_terms(T));
  return normal_from_terms(x, y, z, tt);
}

namespace {

// Plane/range test of Geometry::getInOutBP (ComplexGeometry.cpp:583-597).
bool in_plane_block(const PlaneSpec& s, double x, double y, double z) {
  const long double tolerance = 2.5e-2;
  if (s.plane == "x" && std::abs(x - s.coordinate) <= tolerance)
    return y >= s.range_y[0] && y <= s.range_y[1] && z >= s.range_z[0] && z <= s.range_z[1];
  if (s.plane == "y" && std::abs(y - s.coordinate) <= tolerance)
    return x >= s.range_x[0] && x <= s.range_x[1] && z >= s.range_z[0] && z <= s.range_z[1];
  if (s.plane == "z" && std::abs(z - s.coordinate) <= tolerance)
    return x >= s.range_x[0] && x <= s.range_x[1] && y >= s.range_y[0] && y <= s.range_y[1];
  return false;
}

bool check_block(const PlaneSpec& s, const char* what, std::string* err) {
  if (s.range_x.size() < 2 || s.range_y.size() < 2 || s.range_z.size() < 2) {
    if (err) *err = std::string(what) + " block has a range with fewer than 2 values";
    return false;
  }
  return true;
}

}  // namespace

namespace {

// Near-wall set and its normals from finished node tables (the last steps of
// build_from_labels, shared by the host and the device classification).
bool nearwall_and_normals(const GridSpec& g, const std::vector<Triangle>& tris, const BuildOptions& o, Tables& t,
                          std::string* err) {
  const int64_t N = t.N;
  const int Q = t.Q;
  // Near-wall set (SubDomain::IdentifyNextToWallLattices, Subdomain.cpp:475-493):
  // plain-fluid neighbours of wall nodes. The reference keeps them in an
  // unordered_set (output-only order); here they are sorted by node index.
  {
    std::vector<uint8_t> mark(N, 0);
    for (int32_t w : t.wall_nodes)
      for (int q = 0; q < Q; ++q) {
        const int32_t nbr = t.nb[static_cast<size_t>(q) * N + w];
        if (nbr >= 0 && t.kind[nbr] == KIND_FLUID) mark[nbr] = 1;
      }
    t.nearwall.clear();
    for (int64_t n = 0; n < N; ++n)
      if (mark[n]) t.nearwall.push_back(static_cast<int32_t>(n));
  }
  t.normals.assign(3 * t.nearwall.size(), 0.0);
  if (o.compute_normals) {
    std::vector<TriTerms> tt;
    tt.reserve(tris.size());
    for (const Triangle& T : tris) tt.push_back(tri_terms(T));
    std::vector<double> pts(3 * t.nearwall.size());
    for (size_t a = 0; a < t.nearwall.size(); ++a) {
      const int32_t n = t.nearwall[a];
      pts[3 * a] = g.px(t.ijk[3 * n]);
      pts[3 * a + 1] = g.py(t.ijk[3 * n + 1]);
      pts[3 * a + 2] = g.pz(t.ijk[3 * n + 2]);
    }
    if (o.compute_normals == 2) {
      // Same sum evaluated on the GPU (bit identical; see normals.hip).
      std::vector<double> out;
      if (!device_boundary_normals(pts, tt, out, o.device, err)) return false;
      t.normals.swap(out);
    } else {
      parallel_for(static_cast<int64_t>(t.nearwall.size()), o.nthreads, [&](int64_t a) {
        const auto nv = normal_from_terms(pts[3 * a], pts[3 * a + 1], pts[3 * a + 2], tt);
        t.normals[3 * a] = nv[0];
        t.normals[3 * a + 1] = nv[1];
        t.normals[3 * a + 2] = nv[2];
      });
    }
  }
  return true;
}

}  // namespace

bool classify_on_device(const GridSpec& g, const std::vector<uint8_t>& fluid,
                               const std::vector<PlaneSpec>& inlets, const std::vector<PlaneSpec>& outlets,
                               const std::vector<Triangle>& tris, const BuildOptions& o, Tables& t,
                               std::string* err);

bool build_from_labels(const GridSpec& g, const std::vector<uint8_t>& fluid,
                       const std::vector<PlaneSpec>& inlets, const std::vector<PlaneSpec>& outlets,
                       const std::vector<Triangle>& tris, const BuildOptions& o, Tables& t,
                       std::string* err) {
  if (o.Q != 19 && o.Q != 27) {
    if (err) *err = "Q must be 19 or 27";
    return false;
  }
  if (!(g.spacing > 2e-6)) {
    // Index-set membership stands in for the reference's 1e-6 coordinate match
    // (Utility.cpp:59-74); both agree only while grid points are > 2e-6 apart.
    if (err) *err = "lattice pitch must exceed 2e-6";
    return false;
  }
  for (const auto& s : inlets)
    if (!check_block(s, "inlet", err)) return false;
  for (const auto& s : outlets)
    if (!check_block(s, "outlet", err)) return false;
  if (o.classify_device) return classify_on_device(g, fluid, inlets, outlets, tris, o, t, err);

  const LatticeView L = lattice_view(o.Q);
  const int NX = g.NX, NY = g.NY, NZ = g.NZ;
  t = Tables();
  t.grid = g;
  t.Q = o.Q;
  for (int a = 0; a < 3; ++a) t.periodic[a] = o.periodic[a];
  t.fluid = fluid;
  const int dims[3] = {NX, NY, NZ};

  auto wrap = [&](int v, int axis, bool& ok) -> int {
    if (v >= 0 && v < dims[axis]) return v;
    if (o.periodic[axis]) return (v + dims[axis]) % dims[axis];
    ok = false;
    return v;
  };

  // Boundary points (Geometry::findBoundaryPoints, ComplexGeometry.cpp:517-564).
  const int64_t P = g.npoints();
  std::vector<uint8_t> is_bp(static_cast<size_t>(P), 0);
  parallel_for(NX, o.nthreads, [&](int64_t ii) {
    const int i = static_cast<int>(ii);
    for (int j = 0; j < NY; ++j)
      for (int k = 0; k < NZ; ++k) {
        const int64_t gi = g.gidx(i, j, k);
        if (!fluid[gi]) continue;
        bool bp = false;
        for (int dx = -1; dx <= 1 && !bp; ++dx)
          for (int dy = -1; dy <= 1 && !bp; ++dy)
            for (int dz = -1; dz <= 1 && !bp; ++dz) {
              bool ok = true;
              const int a = wrap(i + dx, 0, ok), b = wrap(j + dy, 1, ok), c = wrap(k + dz, 2, ok);
              if (ok && !fluid[g.gidx(a, b, c)]) bp = true;
            }
        const bool at_face = (!o.periodic[0] && (i == 0 || i == NX - 1)) ||
                             (!o.periodic[1] && (j == 0 || j == NY - 1)) ||
                             (!o.periodic[2] && (k == 0 || k == NZ - 1));
        if (at_face) bp = true;
        is_bp[gi] = bp ? 1 : 0;
      }
  });

  // Node kinds with CreateNodes precedence: inlet > first outlet > wall > fluid.
  std::vector<int32_t> node_of(static_cast<size_t>(P), -1);
  int64_t N = 0;
  for (int64_t gi = 0; gi < P; ++gi)
    if (fluid[gi]) node_of[gi] = static_cast<int32_t>(N++);
  if (N >= (int64_t(1) << 31) - 1) {
    if (err) *err = "more than 2^31-1 nodes";
    return false;
  }
  t.N = N;
  t.ijk.resize(3 * N);
  t.kind.assign(N, KIND_FLUID);
  t.outlet_id.assign(N, -1);
  t.outlet_nodes.assign(outlets.size(), {});
  t.outlet_normals.clear();
  for (const auto& s : outlets) {
    std::array<double, 3> n{0, 0, 0};
    for (size_t a = 0; a < 3 && a < s.normal.size(); ++a) n[a] = s.normal[a];
    t.outlet_normals.push_back(n);
  }
  if (!inlets.empty()) {
    // Only the last inlet block's normal survives (ComplexGeometry.cpp:605).
    const auto& nrm = inlets.back().normal;
    if (nrm.size() < 3) {
      if (err) *err = "inlet normal must have 3 components";
      return false;
    }
    t.inlet_normal = {nrm[0], nrm[1], nrm[2]};
    t.has_inlet_normal = true;
  }
  int64_t nbp = 0;
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < NY; ++j)
      for (int k = 0; k < NZ; ++k) {
        const int64_t gi = g.gidx(i, j, k);
        const int32_t n = node_of[gi];
        if (n < 0) continue;
        t.ijk[3 * n] = i;
        t.ijk[3 * n + 1] = j;
        t.ijk[3 * n + 2] = k;
        if (!is_bp[gi]) continue;
        ++nbp;
        const double x = g.px(i), y = g.py(j), z = g.pz(k);
        bool inl = false;
        for (const auto& s : inlets) inl = inl || in_plane_block(s, x, y, z);
        if (inl) {
          t.kind[n] = KIND_INLET;
          t.inlet_nodes.push_back(n);
          continue;
        }
        int oid = -1;
        for (size_t s = 0; s < outlets.size(); ++s)
          if (in_plane_block(outlets[s], x, y, z)) {
            oid = static_cast<int>(s);
            break;
          }
        if (oid >= 0) {
          t.kind[n] = KIND_OUTLET;
          t.outlet_id[n] = oid;
          t.outlet_nodes[oid].push_back(n);
          continue;
        }
        t.kind[n] = KIND_WALL;
        t.wall_nodes.push_back(n);
      }
  t.n_boundary_points = nbp;

  // Push connectivity (SubDomain::Connectivity, Subdomain.cpp:447-471).
  const int Q = L.Q;
  t.nb.assign(static_cast<size_t>(Q) * N, -1);
  parallel_for(N, o.nthreads, [&](int64_t n) {
    const int i = t.ijk[3 * n], j = t.ijk[3 * n + 1], k = t.ijk[3 * n + 2];
    for (int q = 0; q < Q; ++q) {
      bool ok = true;
      const int a = wrap(i + L.cx[q], 0, ok), b = wrap(j + L.cy[q], 1, ok), c = wrap(k + L.cz[q], 2, ok);
      t.nb[static_cast<size_t>(q) * N + n] = ok ? node_of[g.gidx(a, b, c)] : -1;
    }
  });

  // Node behind an inlet/outlet along its block normal: the outlet copy source
  // of Simulation::ApplyOutletBc (Simulation.cpp:650-741, else-if order
  // +x,-x,+y,-y,+z,-z) and the interior neighbour used by extension BCs.
  t.bc_nbr.assign(N, -1);
  auto behind = [&](int32_t n, const std::array<double, 3>& nv) -> int32_t {
    int di = 0, dj = 0, dk = 0;
    if (nv[0] > 0) di = -1;
    else if (nv[0] < 0) di = 1;
    else if (nv[1] > 0) dj = -1;
    else if (nv[1] < 0) dj = 1;
    else if (nv[2] > 0) dk = -1;
    else if (nv[2] < 0) dk = 1;
    else return -1;
    bool ok = true;
    const int a = wrap(t.ijk[3 * n] + di, 0, ok), b = wrap(t.ijk[3 * n + 1] + dj, 1, ok),
              c = wrap(t.ijk[3 * n + 2] + dk, 2, ok);
    if (!ok) {
      t.upstream_out_of_box++;
      return -1;
    }
    return node_of[g.gidx(a, b, c)];
  };
  for (size_t s = 0; s < t.outlet_nodes.size(); ++s)
    for (int32_t n : t.outlet_nodes[s]) t.bc_nbr[n] = behind(n, t.outlet_normals[s]);
  if (t.has_inlet_normal)
    for (int32_t n : t.inlet_nodes) t.bc_nbr[n] = behind(n, t.inlet_normal);

  return nearwall_and_normals(g, tris, o, t, err);
}

// build_from_labels with the boundary points, node kinds, node order and
// push table computed on the GPU (device_classify); the creation-order lists,
// the nodes behind inlets/outlets and the near-wall set follow from those
// tables here, with the same rules.
bool classify_on_device(const GridSpec& g, const std::vector<uint8_t>& fluid,
                        const std::vector<PlaneSpec>& inlets, const std::vector<PlaneSpec>& outlets,
                        const std::vector<Triangle>& tris, const BuildOptions& o, Tables& t, std::string* err) {
  DeviceTables d;
  if (!device_classify(g, fluid, inlets, outlets, o.Q, o.periodic, o.device, d, err)) return false;
  const LatticeView L = lattice_view(o.Q);
  const int Q = L.Q;
  const int64_t N = d.N;
  t = Tables();
  t.grid = g;
  t.Q = o.Q;
  for (int a = 0; a < 3; ++a) t.periodic[a] = o.periodic[a];
  t.fluid = fluid;
  t.N = N;
  t.ijk = std::move(d.ijk);
  t.kind = std::move(d.kind);
  t.outlet_id = std::move(d.outlet_id);
  t.nb = std::move(d.nb);
  t.n_boundary_points = d.n_boundary_points;
  t.outlet_nodes.assign(outlets.size(), {});
  for (const auto& s : outlets) {
    std::array<double, 3> n{0, 0, 0};
    for (size_t a = 0; a < 3 && a < s.normal.size(); ++a) n[a] = s.normal[a];
    t.outlet_normals.push_back(n);
  }
  if (!inlets.empty()) {
    const auto& nrm = inlets.back().normal;  // only the last block's normal survives (ComplexGeometry.cpp:605)
    if (nrm.size() < 3) {
      if (err) *err = "inlet normal must have 3 components";
      return false;
    }
    t.inlet_normal = {nrm[0], nrm[1], nrm[2]};
    t.has_inlet_normal = true;
  }
  for (int64_t n = 0; n < N; ++n) {
    if (t.kind[n] == KIND_INLET) t.inlet_nodes.push_back(static_cast<int32_t>(n));
    else if (t.kind[n] == KIND_WALL) t.wall_nodes.push_back(static_cast<int32_t>(n));
    else if (t.kind[n] == KIND_OUTLET) t.outlet_nodes[t.outlet_id[n]].push_back(static_cast<int32_t>(n));
  }
  // The node behind an inlet/outlet node along its block normal (else-if
  // order +x,-x,+y,-y,+z,-z), from the push table: the neighbour in that
  // axis direction, or none outside a non-periodic box.
  t.bc_nbr.assign(N, -1);
  const int dims[3] = {g.NX, g.NY, g.NZ};
  auto behind = [&](int32_t n, const std::array<double, 3>& nv) -> int32_t {
    int dv[3] = {0, 0, 0};
    if (nv[0] > 0) dv[0] = -1;
    else if (nv[0] < 0) dv[0] = 1;
    else if (nv[1] > 0) dv[1] = -1;
    else if (nv[1] < 0) dv[1] = 1;
    else if (nv[2] > 0) dv[2] = -1;
    else if (nv[2] < 0) dv[2] = 1;
    else return -1;
    for (int a = 0; a < 3; ++a) {
      const int v = t.ijk[3 * int64_t(n) + a] + dv[a];
      if ((v < 0 || v >= dims[a]) && !o.periodic[a]) {
        t.upstream_out_of_box++;
        return -1;
      }
    }
    for (int q = 0; q < Q; ++q)
      if (L.cx[q] == dv[0] && L.cy[q] == dv[1] && L.cz[q] == dv[2]) return t.nb[static_cast<size_t>(q) * N + n];
    return -1;
  };
  for (size_t s = 0; s < t.outlet_nodes.size(); ++s)
    for (int32_t n : t.outlet_nodes[s]) t.bc_nbr[n] = behind(n, t.outlet_normals[s]);
  if (t.has_inlet_normal)
    for (int32_t n : t.inlet_nodes) t.bc_nbr[n] = behind(n, t.inlet_normal);
  return nearwall_and_normals(g, tris, o, t, err);
}

bool build_from_tables(const TableInput& in, Tables& t, std::string* err) {
  auto bad = [&](const std::string& m) {
    if (err) *err = "lbm_tables: " + m;
    return false;
  };
  if (in.Q != 19 && in.Q != 27) return bad("q must be 19 or 27");
  if (in.N <= 0 || in.N >= (int64_t(1) << 31) - 1) return bad("n_nodes must be in [1, 2^31 - 1)");
  if (!in.ijk || !in.kind || !in.nb) return bad("ijk, kind and nb are required");
  if (in.dims[0] <= 0 || in.dims[1] <= 0 || in.dims[2] <= 0) return bad("grid extents must be positive");
  if (!(in.spacing > 2e-6)) return bad("spacing must exceed 2e-6");
  if (in.n_outlets < 0 || (in.n_outlets > 0 && !in.outlet_normals)) return bad("outlet normals missing");
  if (in.n_nearwall < 0 || (in.n_nearwall > 0 && (!in.nearwall || !in.normals))) return bad("near-wall arrays missing");
  const LatticeView L = lattice_view(in.Q);
  const int Q = L.Q;
  const int64_t N = in.N;
  t = Tables();
  t.Q = Q;
  t.N = N;
  t.grid.NX = in.dims[0];
  t.grid.NY = in.dims[1];
  t.grid.NZ = in.dims[2];
  t.grid.spacing = in.spacing;
  t.grid.bbox = {in.origin[0], in.origin[0] + (in.dims[0] - 1) * in.spacing,
                 in.origin[1], in.origin[1] + (in.dims[1] - 1) * in.spacing,
                 in.origin[2], in.origin[2] + (in.dims[2] - 1) * in.spacing};
  t.ijk.assign(in.ijk, in.ijk + 3 * N);
  t.kind.assign(in.kind, in.kind + N);
  t.outlet_id.assign(N, -1);
  t.outlet_nodes.assign(in.n_outlets, {});
  for (int k = 0; k < in.n_outlets; ++k)
    t.outlet_normals.push_back({in.outlet_normals[3 * k], in.outlet_normals[3 * k + 1], in.outlet_normals[3 * k + 2]});
  t.has_inlet_normal = in.has_inlet_normal;
  if (in.has_inlet_normal) t.inlet_normal = {in.inlet_normal[0], in.inlet_normal[1], in.inlet_normal[2]};

  // Positions inside the grid and unique (one node per grid point).
  std::vector<int64_t> key(N);
  for (int64_t n = 0; n < N; ++n) {
    const int32_t* c = &t.ijk[3 * n];
    if (c[0] < 0 || c[0] >= in.dims[0] || c[1] < 0 || c[1] >= in.dims[1] || c[2] < 0 || c[2] >= in.dims[2])
      return bad("node " + std::to_string(n) + " lies outside the grid");
    key[n] = t.grid.gidx(c[0], c[1], c[2]);
  }
  {
    std::vector<int64_t> sorted(key);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) return bad("two nodes share a grid point");
  }
  // Kinds and outlet blocks; creation-order lists (node order is creation order).
  for (int64_t n = 0; n < N; ++n) {
    const uint8_t k = t.kind[n];
    const int32_t oid = in.outlet_id ? in.outlet_id[n] : -1;
    if (k > KIND_OUTLET) return bad("node " + std::to_string(n) + " has an unknown kind");
    if ((k == KIND_OUTLET) != (oid >= 0)) return bad("node " + std::to_string(n) + ": outlet block and kind disagree");
    if (oid >= in.n_outlets) return bad("node " + std::to_string(n) + ": outlet block out of range");
    if (k == KIND_INLET) t.inlet_nodes.push_back(static_cast<int32_t>(n));
    if (k == KIND_WALL) t.wall_nodes.push_back(static_cast<int32_t>(n));
    if (k == KIND_OUTLET) {
      t.outlet_id[n] = oid;
      t.outlet_nodes[oid].push_back(static_cast<int32_t>(n));
    }
    if (k != KIND_FLUID) t.n_boundary_points++;
  }
  // Neighbour entries: in range and at x + c_q (no wrap: the reference has
  // no periodic axes).
  t.nb.assign(in.nb, in.nb + static_cast<size_t>(Q) * N);
  for (int q = 0; q < Q; ++q)
    for (int64_t n = 0; n < N; ++n) {
      const int32_t m = t.nb[static_cast<size_t>(q) * N + n];
      if (m < 0) continue;
      if (m >= N) return bad("neighbour entry out of range");
      const int32_t* a = &t.ijk[3 * n];
      const int32_t* b = &t.ijk[3 * int64_t(m)];
      if (b[0] != a[0] + L.cx[q] || b[1] != a[1] + L.cy[q] || b[2] != a[2] + L.cz[q])
        return bad("neighbour " + std::to_string(q) + " of node " + std::to_string(n) + " is not at x + c_q");
    }
  // The node behind an inlet/outlet node: its neighbour along -n_hat, taken
  // in the reference's else-if order of the normal's components.
  t.bc_nbr.assign(N, -1);
  auto behind = [&](int32_t n, const std::array<double, 3>& nv) -> int32_t {
    int d[3] = {0, 0, 0};
    if (nv[0] > 0) d[0] = -1;
    else if (nv[0] < 0) d[0] = 1;
    else if (nv[1] > 0) d[1] = -1;
    else if (nv[1] < 0) d[1] = 1;
    else if (nv[2] > 0) d[2] = -1;
    else if (nv[2] < 0) d[2] = 1;
    else return -1;
    const int32_t* c = &t.ijk[3 * int64_t(n)];
    for (int a = 0; a < 3; ++a)
      if (c[a] + d[a] < 0 || c[a] + d[a] >= in.dims[a]) {
        t.upstream_out_of_box++;
        return -1;
      }
    for (int q = 0; q < Q; ++q)
      if (L.cx[q] == d[0] && L.cy[q] == d[1] && L.cz[q] == d[2]) return t.nb[static_cast<size_t>(q) * N + n];
    return -1;
  };
  for (size_t s = 0; s < t.outlet_nodes.size(); ++s)
    for (int32_t n : t.outlet_nodes[s]) t.bc_nbr[n] = behind(n, t.outlet_normals[s]);
  if (t.has_inlet_normal)
    for (int32_t n : t.inlet_nodes) t.bc_nbr[n] = behind(n, t.inlet_normal);
  // Near-wall nodes as given.
  std::vector<uint8_t> seen(N, 0);
  for (int64_t i = 0; i < in.n_nearwall; ++i) {
    const int32_t n = in.nearwall[i];
    if (n < 0 || n >= N) return bad("near-wall node out of range");
    if (t.kind[n] != KIND_FLUID) return bad("near-wall node " + std::to_string(n) + " is not a plain fluid node");
    if (seen[n]++) return bad("near-wall node " + std::to_string(n) + " listed twice");
    t.nearwall.push_back(n);
  }
  t.normals.assign(in.normals, in.normals + 3 * in.n_nearwall);
  // Grid labels from the node positions (Point::Type Fluid), where the grid
  // is small enough to hold them.
  if (t.grid.npoints() <= (int64_t(1) << 32)) {
    t.fluid.assign(static_cast<size_t>(t.grid.npoints()), 0);
    for (int64_t n = 0; n < N; ++n) t.fluid[key[n]] = 1;
  }
  return true;
}

bool build_from_config(const SimulationParameters& p, const std::string& base_dir, const BuildOptions& o,
                       Tables& t, std::vector<Triangle>* tris_out, std::string* err) {
  std::string path = p.stlFile;
  if (!path.empty() && path[0] != '/' && !base_dir.empty()) path = base_dir + "/" + path;
  std::vector<Triangle> tris;
  if (!readSTLASCII(path, tris)) {
    if (err) *err = "cannot open STL file " + path;
    return false;
  }
  if (tris.empty()) {
    if (err) *err = "STL file has no facets: " + path;
    return false;
  }
  if (!(p.dx > 0)) {
    if (err) *err = "grid_size must be positive";
    return false;
  }
  const GridSpec g = make_grid(bounding_box(tris), p.dx);
  std::vector<uint8_t> fluid;
  if (!label_points(g, tris, fluid, o.nthreads, o.verify_stride, err, o.label_device, o.device)) return false;
  if (!build_from_labels(g, fluid, p.inlets, p.outlets, tris, o, t, err)) return false;
  if (tris_out) tris_out->swap(tris);
  return true;
}

}  // namespace lbm
