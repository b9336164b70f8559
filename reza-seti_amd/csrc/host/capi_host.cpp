// Host half of the C ABI: config parsing and geometry/node tables.
#include <cmath>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <limits>
#include <string>

#include "../capi_internal.h"
#include "output.h"
#include "../lattice.h"

namespace lbm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

static void copy_block(const PlaneSpec& s, lbm_plane_block& b) {
  std::memset(&b, 0, sizeof(b));
  std::strncpy(b.plane, s.plane.c_str(), sizeof(b.plane) - 1);
  b.coordinate = s.coordinate;
  for (int i = 0; i < 2; ++i) {
    b.range_x[i] = i < (int)s.range_x.size() ? s.range_x[i] : 0.0;
    b.range_y[i] = i < (int)s.range_y.size() ? s.range_y[i] : 0.0;
    b.range_z[i] = i < (int)s.range_z.size() ? s.range_z[i] : 0.0;
  }
  b.normal_len = static_cast<int32_t>(s.normal.size());
  for (int i = 0; i < 3 && i < (int)s.normal.size(); ++i) b.normal[i] = s.normal[i];
}

static PlaneSpec to_spec(const lbm_plane_block& b) {
  PlaneSpec s;
  s.plane = std::string(b.plane, strnlen(b.plane, sizeof(b.plane)));
  s.coordinate = b.coordinate;
  s.range_x = {b.range_x[0], b.range_x[1]};
  s.range_y = {b.range_y[0], b.range_y[1]};
  s.range_z = {b.range_z[0], b.range_z[1]};
  const int n = b.normal_len < 0 ? 0 : (b.normal_len > 3 ? 3 : b.normal_len);
  for (int i = 0; i < n; ++i) s.normal.push_back(b.normal[i]);
  return s;
}

static BuildOptions to_opts(const lbm_geometry_options* o) {
  BuildOptions b;
  if (!o) return b;
  b.Q = o->q;
  for (int a = 0; a < 3; ++a) b.periodic[a] = o->periodic[a];
  b.nthreads = o->nthreads;
  b.verify_stride = o->verify_stride;
  b.compute_normals = o->compute_normals;
  b.device = o->device;
  b.label_device = o->label_device;
  b.classify_device = o->classify_device;
  return b;
}

}  // namespace lbm

using namespace lbm;

extern "C" {

const char* lbm_last_error(void) { return g_last_error.c_str(); }

// Simulation::getInletVelocity (Simulation.cpp:519-534) over the three
// polynomial pieces of Pulsatile::InletVelequation1..3 (pulsatile.h:55-69),
// with the reference's terms, powers and branch bounds.
double lbm_inlet_velocity(double t, double period) {
  const double x = std::fmod(t, period);
  if (x >= 0 && x < (0.1 * period))
    return -8.579e+08 * std::pow(x, 8) + 4.059e+08 * std::pow(x, 7) - 7.817e+07 * std::pow(x, 6) +
           7.994e+06 * std::pow(x, 5) - 4.835e+05 * std::pow(x, 4) + 1.782e+04 * std::pow(x, 3) -
           321.6 * std::pow(x, 2) + 3.215 * x + 0.1423;
  if (x >= (0.1 * period) && x < (0.3 * period))
    return -4.32e+06 * std::pow(x, 8) + 7.94e+06 * std::pow(x, 7) - 6.298e+06 * std::pow(x, 6) +
           2.82e+06 * std::pow(x, 5) - 7.809e+05 * std::pow(x, 4) + 1.37e+05 * std::pow(x, 3) -
           1.484e+04 * std::pow(x, 2) + 901.9 * x - 22.87;
  if (x >= (0.3 * period) && x < (0.9 * period))
    return -5.123 * std::pow(x, 4) + 14.15 * std::pow(x, 3) - 14.21 * std::pow(x, 2) + 5.865 * x - 0.5696;
  return 0.0;
}

int lbm_lattice(int32_t q, int32_t* c, int32_t* opp, double* w) {
  if (q != 19 && q != 27) return fail(LBM_ERR_ARG, "q must be 19 or 27");
  const LatticeView L = lattice_view(q);
  for (int i = 0; i < q; ++i) {
    if (c) {
      c[3 * i] = L.cx[i];
      c[3 * i + 1] = L.cy[i];
      c[3 * i + 2] = L.cz[i];
    }
    if (opp) opp[i] = L.opp[i];
    if (w) w[i] = L.w[i];
  }
  return LBM_OK;
}

int lbm_config_parse(const char* path, lbm_config* out) {
  if (!path || !out) return fail(LBM_ERR_ARG, "null argument");
  SimulationParameters p;
  try {
    parseInputFile(path, p);
  } catch (const std::runtime_error& e) {
    return fail(LBM_ERR_IO, e.what());
  } catch (const std::exception& e) {
    return fail(LBM_ERR_ARG, std::string("config value: ") + e.what());
  }
  std::memset(out, 0, sizeof(*out));
  std::strncpy(out->stl_file, p.stlFile.c_str(), sizeof(out->stl_file) - 1);
  std::strncpy(out->flow_type, p.FlowType.c_str(), sizeof(out->flow_type) - 1);
  out->dt = p.dt;
  out->rho = p.rho;
  out->nu_blood = p.nu_blood;
  out->tau = p.tau;
  out->dx = p.dx;
  out->pressure = p.pressure;
  out->reynolds_number = p.reynoldsNumber;
  out->cte_vel = p.CteVel;
  out->systole = p.Systole;
  out->diastole = p.Diastole;
  out->num_out = p.num_Out;
  out->num_cycle = p.NumCycle;
  out->max_iter = p.MaxIter;
  out->iter_to_write = p.IterToWrite;
  out->n_equations = static_cast<int32_t>(p.equations.size());
  if (p.inlets.size() > LBM_MAX_BLOCKS || p.outlets.size() > LBM_MAX_BLOCKS)
    return fail(LBM_ERR_ARG, "more than 16 inlet or outlet blocks");
  out->n_inlets = static_cast<int32_t>(p.inlets.size());
  out->n_outlets = static_cast<int32_t>(p.outlets.size());
  for (size_t i = 0; i < p.inlets.size(); ++i) copy_block(p.inlets[i], out->inlets[i]);
  for (size_t i = 0; i < p.outlets.size(); ++i) copy_block(p.outlets[i], out->outlets[i]);
  return LBM_OK;
}

void lbm_geometry_options_default(lbm_geometry_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->q = 19;
  o->compute_normals = 1;
}

int lbm_geometry_from_config(const char* config_path, const lbm_geometry_options* opts, lbm_geometry** out) {
  if (!config_path || !out) return fail(LBM_ERR_ARG, "null argument");
  *out = nullptr;
  SimulationParameters p;
  try {
    parseInputFile(config_path, p);
  } catch (const std::runtime_error& e) {
    return fail(LBM_ERR_IO, e.what());
  } catch (const std::exception& e) {
    return fail(LBM_ERR_ARG, std::string("config value: ") + e.what());
  }
  std::string dir(config_path);
  const size_t slash = dir.find_last_of('/');
  dir = (slash == std::string::npos) ? std::string(".") : dir.substr(0, slash);
  auto* g = new lbm_geometry();
  std::string err;
  bool ok = false;
  try {
    ok = build_from_config(p, dir, to_opts(opts), g->t, &g->tris, &err);
  } catch (const std::exception& e) {
    err = e.what();
  }
  if (!ok) {
    delete g;
    return fail(err.rfind("cannot open", 0) == 0 ? LBM_ERR_IO : LBM_ERR_ARG, err);
  }
  *out = g;
  return LBM_OK;
}

int lbm_geometry_from_labels(const uint8_t* fluid, int32_t nx, int32_t ny, int32_t nz, const double origin[3],
                             double spacing, const lbm_plane_block* inlets, int32_t n_inlets,
                             const lbm_plane_block* outlets, int32_t n_outlets, const double* tris, int64_t ntris,
                             const lbm_geometry_options* opts, lbm_geometry** out) {
  if (!fluid || !origin || !out || nx <= 0 || ny <= 0 || nz <= 0 || n_inlets < 0 || n_outlets < 0 || ntris < 0)
    return fail(LBM_ERR_ARG, "invalid argument");
  if ((n_inlets && !inlets) || (n_outlets && !outlets) || (ntris && !tris))
    return fail(LBM_ERR_ARG, "null block or triangle array");
  *out = nullptr;
  GridSpec g;
  g.spacing = spacing;
  g.NX = nx;
  g.NY = ny;
  g.NZ = nz;
  g.bbox = {origin[0], origin[0] + (nx - 1) * spacing, origin[1], origin[1] + (ny - 1) * spacing,
            origin[2], origin[2] + (nz - 1) * spacing};
  std::vector<uint8_t> lab(fluid, fluid + g.npoints());
  std::vector<PlaneSpec> in, outl;
  for (int i = 0; i < n_inlets; ++i) in.push_back(to_spec(inlets[i]));
  for (int i = 0; i < n_outlets; ++i) outl.push_back(to_spec(outlets[i]));
  auto* geo = new lbm_geometry();
  geo->tris.resize(static_cast<size_t>(ntris));
  for (int64_t i = 0; i < ntris; ++i) {
    const double* r = tris + 12 * i;
    geo->tris[i] = {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11]};
  }
  std::string err;
  bool ok = false;
  try {
    ok = build_from_labels(g, lab, in, outl, geo->tris, to_opts(opts), geo->t, &err);
  } catch (const std::exception& e) {
    err = e.what();
  }
  if (!ok) {
    delete geo;
    return fail(LBM_ERR_ARG, err);
  }
  *out = geo;
  return LBM_OK;
}

int lbm_geometry_get_info(const lbm_geometry* g, lbm_geometry_info* o) {
  if (!g || !o) return fail(LBM_ERR_ARG, "null argument");
  std::memset(o, 0, sizeof(*o));
  const Tables& t = g->t;
  o->n_nodes = t.N;
  o->n_wall = static_cast<int64_t>(t.wall_nodes.size());
  o->n_inlet = static_cast<int64_t>(t.inlet_nodes.size());
  int64_t nout = 0;
  for (const auto& v : t.outlet_nodes) nout += static_cast<int64_t>(v.size());
  o->n_outlet = nout;
  o->n_fluid = t.N - o->n_wall - o->n_inlet - nout;
  o->n_nearwall = static_cast<int64_t>(t.nearwall.size());
  o->n_boundary_points = t.n_boundary_points;
  o->n_triangles = static_cast<int64_t>(g->tris.size());
  o->upstream_out_of_box = t.upstream_out_of_box;
  o->grid[0] = t.grid.NX;
  o->grid[1] = t.grid.NY;
  o->grid[2] = t.grid.NZ;
  o->q = t.Q;
  o->n_outlets = static_cast<int32_t>(t.outlet_nodes.size());
  o->has_inlet_normal = t.has_inlet_normal ? 1 : 0;
  o->spacing = t.grid.spacing;
  const BBox& b = t.grid.bbox;
  o->bbox[0] = b.min_x;
  o->bbox[1] = b.max_x;
  o->bbox[2] = b.min_y;
  o->bbox[3] = b.max_y;
  o->bbox[4] = b.min_z;
  o->bbox[5] = b.max_z;
  for (int a = 0; a < 3; ++a) o->inlet_normal[a] = t.inlet_normal[a];
  return LBM_OK;
}

int lbm_geometry_get_labels(const lbm_geometry* g, uint8_t* fluid) {
  if (!g || !fluid) return fail(LBM_ERR_ARG, "null argument");
  if (static_cast<int64_t>(g->t.fluid.size()) != g->t.grid.npoints())
    return fail(LBM_ERR_UNSUPPORTED, "no grid labels kept for this geometry (tables over more than 2^32 grid points)");
  std::memcpy(fluid, g->t.fluid.data(), g->t.fluid.size());
  return LBM_OK;
}

int lbm_geometry_from_tables(const lbm_tables* in, lbm_geometry** out) {
  if (!in || !out) return fail(LBM_ERR_ARG, "null argument");
  *out = nullptr;
  TableInput ti;
  ti.Q = in->q;
  ti.N = in->n_nodes;
  for (int a = 0; a < 3; ++a) {
    ti.dims[a] = in->grid[a];
    ti.origin[a] = in->origin[a];
    ti.inlet_normal[a] = in->inlet_normal[a];
  }
  ti.spacing = in->spacing;
  ti.ijk = in->ijk;
  ti.kind = in->kind;
  ti.outlet_id = in->outlet_id;
  ti.nb = in->nb;
  ti.has_inlet_normal = in->has_inlet_normal != 0;
  ti.n_outlets = in->n_outlets;
  ti.outlet_normals = in->outlet_normals;
  ti.n_nearwall = in->n_nearwall;
  ti.nearwall = in->nearwall;
  ti.normals = in->normals;
  auto* g = new lbm_geometry();
  std::string err;
  bool ok = false;
  try {
    ok = build_from_tables(ti, g->t, &err);
  } catch (const std::exception& e) {
    err = e.what();
  }
  if (!ok) {
    delete g;
    return fail(LBM_ERR_ARG, err);
  }
  *out = g;
  return LBM_OK;
}

int lbm_geometry_get_nodes(const lbm_geometry* g, int32_t* ijk, uint8_t* kind, int32_t* outlet_id) {
  if (!g) return fail(LBM_ERR_ARG, "null argument");
  const Tables& t = g->t;
  if (ijk) std::memcpy(ijk, t.ijk.data(), t.ijk.size() * sizeof(int32_t));
  if (kind) std::memcpy(kind, t.kind.data(), t.kind.size());
  if (outlet_id) std::memcpy(outlet_id, t.outlet_id.data(), t.outlet_id.size() * sizeof(int32_t));
  return LBM_OK;
}

int lbm_geometry_get_neighbours(const lbm_geometry* g, int32_t* nb) {
  if (!g || !nb) return fail(LBM_ERR_ARG, "null argument");
  std::memcpy(nb, g->t.nb.data(), g->t.nb.size() * sizeof(int32_t));
  return LBM_OK;
}

int lbm_geometry_get_nearwall(const lbm_geometry* g, int32_t* nodes, double* normals) {
  if (!g) return fail(LBM_ERR_ARG, "null argument");
  if (nodes) std::memcpy(nodes, g->t.nearwall.data(), g->t.nearwall.size() * sizeof(int32_t));
  if (normals) std::memcpy(normals, g->t.normals.data(), g->t.normals.size() * sizeof(double));
  return LBM_OK;
}

int lbm_geometry_get_bc_neighbour(const lbm_geometry* g, int32_t* bc_nbr) {
  if (!g || !bc_nbr) return fail(LBM_ERR_ARG, "null argument");
  std::memcpy(bc_nbr, g->t.bc_nbr.data(), g->t.bc_nbr.size() * sizeof(int32_t));
  return LBM_OK;
}

int lbm_geometry_get_outlet_normals(const lbm_geometry* g, double* normals) {
  if (!g || !normals) return fail(LBM_ERR_ARG, "null argument");
  for (size_t i = 0; i < g->t.outlet_normals.size(); ++i)
    for (int a = 0; a < 3; ++a) normals[3 * i + a] = g->t.outlet_normals[i][a];
  return LBM_OK;
}

int lbm_geometry_get_triangles(const lbm_geometry* g, double* tris) {
  if (!g || !tris) return fail(LBM_ERR_ARG, "null argument");
  for (size_t i = 0; i < g->tris.size(); ++i) {
    const Triangle& T = g->tris[i];
    const double v[12] = {T.x1, T.y1, T.z1, T.x2, T.y2, T.z2, T.x3, T.y3, T.z3, T.nx, T.ny, T.nz};
    std::memcpy(tris + 12 * i, v, sizeof(v));
  }
  return LBM_OK;
}

void lbm_geometry_destroy(lbm_geometry* g) { delete g; }


int lbm_write_vtu(const lbm_geometry* g, const double* rho, const double* u, const char* path) {
  if (!g || !rho || !u || !path) return lbm::fail(LBM_ERR_ARG, "null argument");
  const int64_t N = g->t.N;
  std::vector<double> r(rho, rho + N), v(u, u + 3 * N);
  std::string err;
  if (!lbm::write_vtu(g->t, r, v, path, &err)) return lbm::fail(LBM_ERR_IO, err);
  return LBM_OK;
}

int lbm_smooth_surface_csv(const char* in_path, const char* out_path, double radius, int32_t pulsatile) {
  if (!in_path || !out_path) return lbm::fail(LBM_ERR_ARG, "null argument");
  std::string err;
  if (!lbm::smooth_surface_csv(in_path, out_path, radius, pulsatile != 0, &err)) return lbm::fail(LBM_ERR_IO, err);
  return LBM_OK;
}

}  // extern "C"
