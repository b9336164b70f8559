// ORACLE (test infrastructure only): a C ABI over the reference's own
// compiled set-up classes, used by the CPU tests to pin the product's
// classification, connectivity, near-wall normals, Zou-He rule, STL reader
// and config parser against the real code (single MPI rank). Built by
// oracle/ref/Makefile into oracle/_ref/libref.so; never part of the product.
#include <unistd.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <vector>

#include "Boundary.h"
#include "ComplexGeometry.h"
#include "GeometryReader.h"
#include "LBM.h"
#include "LatticeModel.h"
#include "Subdomain.h"
#include "pulsatile.h"

extern "C" {

struct ref_tables {
  int32_t grid[3];
  double spacing;
  double bbox[6];
  uint8_t* labels;      // (i * ny + j) * nz + k
  int64_t n;            // nodes in Involved_Lat_ order
  int32_t* ijk;         // 3n
  uint8_t* kind;        // 0 fluid, 1 wall, 2 inlet, 3 outlet
  int32_t* outlet_id;   // n
  int32_t* nb;          // 19 * n push table, q-major
  int32_t* fluid_bt;    // n: BoundaryType_ bytes found in plain fluid nodes, -1 for boundary nodes
  int64_t n_nw;
  int32_t* nw_ijk;      // 3 n_nw, nextToTheWall order
  double* nw_normal;    // 3 n_nw
  double inlet_normal[3];
};

struct ref_block {
  char plane[16];
  double coordinate;
  double range_x[2], range_y[2], range_z[2];
  double normal[3];
  int32_t normal_len;
};

struct ref_config {
  char stl_file[512];
  char flow_type[32];
  double dt, rho, nu_blood, tau, dx, pressure, reynolds_number, cte_vel, systole, diastole;
  int32_t num_out, num_cycle, max_iter, iter_to_write, n_inlets, n_outlets, n_equations;
  ref_block inlets[16];
  ref_block outlets[16];
};

static void ensure_mpi() {
  int init = 0;
  MPI_Initialized(&init);
  if (!init) {
    int argc = 0;
    char** argv = nullptr;
    MPI_Init(&argc, &argv);
  }
}

static void fill_block(const SimulationParameters::Inlet& s, ref_block& b) {
  std::memset(&b, 0, sizeof(b));
  std::strncpy(b.plane, s.plane.c_str(), sizeof(b.plane) - 1);
  b.coordinate = s.coordinate;
  for (int i = 0; i < 2; ++i) {
    b.range_x[i] = i < (int)s.range_x.size() ? s.range_x[i] : 0.0;
    b.range_y[i] = i < (int)s.range_y.size() ? s.range_y[i] : 0.0;
    b.range_z[i] = i < (int)s.range_z.size() ? s.range_z[i] : 0.0;
  }
  b.normal_len = (int32_t)s.normal.size();
  for (int i = 0; i < 3 && i < (int)s.normal.size(); ++i) b.normal[i] = s.normal[i];
}

int ref_parse_config(const char* path, ref_config* out) {
  SimulationParameters p;
  try {
    parseInputFile(path, p);
  } catch (...) {
    return 1;
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
  out->n_equations = (int32_t)p.equations.size();
  out->n_inlets = (int32_t)p.inlets.size();
  out->n_outlets = (int32_t)p.outlets.size();
  for (size_t i = 0; i < p.inlets.size() && i < 16; ++i) fill_block(p.inlets[i], out->inlets[i]);
  for (size_t i = 0; i < p.outlets.size() && i < 16; ++i) {
    SimulationParameters::Inlet tmp{p.outlets[i].plane, p.outlets[i].coordinate, p.outlets[i].range_x,
                                    p.outlets[i].range_y, p.outlets[i].range_z, p.outlets[i].normal};
    fill_block(tmp, out->outlets[i]);
  }
  return 0;
}

int64_t ref_read_stl(const char* path, double* out, int64_t max) {
  Reader r;
  std::vector<Triangle> t = r.readSTLASCII(path);
  const int64_t n = (int64_t)t.size();
  for (int64_t i = 0; i < n && i < max; ++i) {
    const Triangle& T = t[i];
    const double v[12] = {T.x1, T.y1, T.z1, T.x2, T.y2, T.z2, T.x3, T.y3, T.z3, T.nx, T.ny, T.nz};
    std::memcpy(out + 12 * i, v, sizeof(v));
  }
  return n;
}

int ref_lattice(int32_t* c, int32_t* opp, double* w) {
  LatticeModel m;
  for (int q = 0; q < _nLatNodes; ++q) {
    const std::vector<int> d = m.GetDirection(q);
    for (int a = 0; a < 3; ++a) c[3 * q + a] = d[a];
    opp[q] = m.OppositeDirection(q);
    w[q] = m.GetWeight(q);
  }
  return _nLatNodes;
}

// The reference's pulsatile waveform: Pulsatile::InletVelequation1..3 are
// the reference header itself (pulsatile.h:49-71, compiled here unmodified).
// Simulation::getInletVelocity (Simulation.cpp:519-534), which selects the
// piece, lives in Simulation.cpp, which needs VTK and does not build here; its
// eight lines of branch logic are restated below around the compiled pieces.
double ref_inlet_velocity(double x, double period) {
  Pulsatile p;
  double vel = 0.0;
  const double nx = fmod(x, period);
  if (nx >= 0 && nx < (0.1 * period)) vel = p.InletVelequation1(nx);
  else if (nx >= (0.1 * period) && nx < (0.3 * period)) vel = p.InletVelequation2(nx);
  else if (nx >= (0.3 * period) && nx < (0.9 * period)) vel = p.InletVelequation3(nx);
  return vel;
}

// One piece on its own (1, 2 or 3), for checks across the branch bounds.
double ref_pulsatile_piece(int piece, double x) {
  Pulsatile p;
  return piece == 1 ? p.InletVelequation1(x) : piece == 2 ? p.InletVelequation2(x) : p.InletVelequation3(x);
}

int ref_zouhe(double vel, const double* normal3, double* f) {
  auto vs = std::make_shared<LatticeModel>();
  Boundary b(0, 0, 0, vs, NodeType::Boundary);
  b.m_distributions.assign(f, f + _nLatNodes);
  std::vector<double> n(normal3, normal3 + 3);
  double v = vel;
  b.ApplyBc(v, n);
  std::memcpy(f, b.m_distributions.data(), sizeof(double) * _nLatNodes);
  return 0;
}

// Runs Geometry + SubDomain on config_name inside config_dir (the reference
// reads the STL relative to the working directory and writes debug VTK files
// there).
int ref_build(const char* config_dir, const char* config_name, ref_tables* out) {
  ensure_mpi();
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd))) return 1;
  if (chdir(config_dir) != 0) return 1;
  int rc = 0;
  try {
    SimulationParameters sp;
    parseInputFile(config_name, sp);
    auto vs = std::make_shared<LatticeModel>();
    Geometry geom(vs, sp);
    SubDomain sub(geom);
    std::memset(out, 0, sizeof(*out));
    const int nx = (int)geom.pointsBox.size(), ny = (int)geom.pointsBox[0].size(),
              nz = (int)geom.pointsBox[0][0].size();
    out->grid[0] = nx;
    out->grid[1] = ny;
    out->grid[2] = nz;
    out->spacing = geom.spacing;
    const BoundingBox& b = geom.boundingbox;
    const double bb[6] = {b.min_x, b.max_x, b.min_y, b.max_y, b.min_z, b.max_z};
    std::memcpy(out->bbox, bb, sizeof(bb));
    out->labels = (uint8_t*)std::malloc((size_t)nx * ny * nz);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k)
          out->labels[((size_t)i * ny + j) * nz + k] = geom.pointsBox[i][j][k].Type == NodeType::Fluid ? 1 : 0;
    const auto& L = sub.Involved_Lat_;
    const int64_t n = (int64_t)L.size();
    out->n = n;
    out->ijk = (int32_t*)std::malloc(sizeof(int32_t) * 3 * n);
    out->kind = (uint8_t*)std::malloc(n);
    out->outlet_id = (int32_t*)std::malloc(sizeof(int32_t) * n);
    out->nb = (int32_t*)std::malloc(sizeof(int32_t) * _nLatNodes * n);
    out->fluid_bt = (int32_t*)std::malloc(sizeof(int32_t) * n);
    std::unordered_map<const Node*, int64_t> index;
    for (int64_t i = 0; i < n; ++i) index[L[i].get()] = i;
    std::unordered_map<const Node*, int> outlet_of;
    for (size_t k = 0; k < sub.OutletBoundaryNodes_.size(); ++k)
      for (const auto& p : sub.OutletBoundaryNodes_[k].first) outlet_of[p.get()] = (int)k;
    for (int64_t i = 0; i < n; ++i) {
      const Node* nd = L[i].get();
      out->ijk[3 * i] = nd->x_position;
      out->ijk[3 * i + 1] = nd->y_position;
      out->ijk[3 * i + 2] = nd->z_position;
      uint8_t kd = 0;
      int32_t raw = -1;
      if (nd->NodeType_ == NodeType::Boundary) {
        kd = nd->BoundaryType_ == BoundaryType::Inlet ? 2 : nd->BoundaryType_ == BoundaryType::Outlet ? 3 : 1;
      } else {
        int v = 0;
        std::memcpy(&v, &nd->BoundaryType_, sizeof(v) < sizeof(nd->BoundaryType_) ? sizeof(v)
                                                                                    : sizeof(nd->BoundaryType_));
        raw = v;
      }
      out->kind[i] = kd;
      out->fluid_bt[i] = raw;
      auto it = outlet_of.find(nd);
      out->outlet_id[i] = it == outlet_of.end() ? -1 : it->second;
      for (int q = 0; q < _nLatNodes; ++q) {
        const Node* m = nd->m_neighbours[q].get();
        out->nb[(size_t)q * n + i] = m ? (int32_t)index.at(m) : -1;
      }
    }
    const auto& W = sub.nextToTheWall;
    out->n_nw = (int64_t)W.size();
    out->nw_ijk = (int32_t*)std::malloc(sizeof(int32_t) * 3 * (W.size() + 1));
    out->nw_normal = (double*)std::malloc(sizeof(double) * 3 * (W.size() + 1));
    for (size_t i = 0; i < W.size(); ++i) {
      out->nw_ijk[3 * i] = W[i]->x_position;
      out->nw_ijk[3 * i + 1] = W[i]->y_position;
      out->nw_ijk[3 * i + 2] = W[i]->z_position;
      for (int a = 0; a < 3; ++a) out->nw_normal[3 * i + a] = W[i]->normal[a];
    }
    const auto& in = sub.InletBoundaryNode_.second;
    for (int a = 0; a < 3 && a < (int)in.size(); ++a) out->inlet_normal[a] = in[a];
  } catch (...) {
    rc = 2;
  }
  if (chdir(cwd) != 0) rc = rc ? rc : 3;
  return rc;
}

void ref_free(ref_tables* t) {
  std::free(t->labels);
  std::free(t->ijk);
  std::free(t->kind);
  std::free(t->outlet_id);
  std::free(t->nb);
  std::free(t->fluid_bt);
  std::free(t->nw_ijk);
  std::free(t->nw_normal);
  std::memset(t, 0, sizeof(*t));
}

}  // extern "C"
