// C++ drop-in for the reference's class surface over the C ABI.
//
// Mirrors the construction chain of LBM.cpp:95-111
//   parseInputFile -> LatticeModel -> Geometry -> SubDomain -> Simulation
// and the public methods of Simulation (Simulation.h:74-87), VelocitySet /
// LatticeModel (VelocitySet.h:60-65, LatticeModel.h:50-54). Everything that
// computes runs behind include/lbm_c_api.h on the GPU; these classes only hold
// handles and forward calls.
//
// One reference iteration is one fused launch. Collision() issues it (the
// launch streams, applies the inlet/outlet/wall rules, accumulates the wall
// shear stress of the current iteration and performs the collision of the
// next one). Communicate, Streaming, ApplyZouHeBc, ApplyOutletBc, ApplyWallBc
// and ComputeWallShearStress are therefore folded into it and do nothing on
// their own; calling the reference sequence in order gives the reference
// iteration.
#pragma once

#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../../include/lbm_c_api.h"
#include "../config.h"

namespace reza {

using SimulationParameters = lbm::SimulationParameters;

inline void parseInputFile(const std::string& fileName, SimulationParameters& p) {
  lbm::parseInputFile(fileName, p);
}

inline void check(int rc, const char* what) {
  if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

// VelocitySet / LatticeModel getters (VelocitySet.cpp:52-82, LatticeModel.cpp:149-191).
class LatticeModel {
 public:
  explicit LatticeModel(int q = 19) : q_(q), c_(3 * q), opp_(q), w_(q) {
    check(lbm_lattice(q, c_.data(), opp_.data(), w_.data()), "lbm_lattice");
  }
  double GetWeight(int dir) const { return w_[dir]; }
  std::vector<int> GetDirection(int dir) const { return {c_[3 * dir], c_[3 * dir + 1], c_[3 * dir + 2]}; }
  double Get_c_square(int) const { return 1.0 / 3.0; }
  int Get_nDimensions() const { return 3; }
  int Get_nDirections() const { return q_; }
  int OppositeDirection(int dir) const { return opp_[dir]; }

 private:
  int q_;
  std::vector<int32_t> c_, opp_;
  std::vector<double> w_;
};

// Geometry: STL -> grid -> labels -> boundary sets (ComplexGeometry.cpp:48-73).
class Geometry {
 public:
  Geometry(std::shared_ptr<LatticeModel> velSet, SimulationParameters& simParams, const std::string& config_path)
      : velSet_(std::move(velSet)), simParams(simParams) {
    lbm_geometry_options o;
    lbm_geometry_options_default(&o);
    o.q = velSet_->Get_nDirections();
    lbm_geometry* g = nullptr;
    check(lbm_geometry_from_config(config_path.c_str(), &o, &g), "Geometry");
    handle_.reset(g, lbm_geometry_destroy);
    check(lbm_geometry_get_info(g, &info_), "Geometry info");
    spacing = info_.spacing;
    GeomSize = {info_.grid[0], info_.grid[1], info_.grid[2]};
  }
  // The same, from node tables a reference SubDomain already holds
  // (lbm_geometry_from_tables): no second voxelisation.
  Geometry(std::shared_ptr<LatticeModel> velSet, SimulationParameters& simParams, const lbm_tables& tables)
      : velSet_(std::move(velSet)), simParams(simParams) {
    lbm_geometry* g = nullptr;
    check(lbm_geometry_from_tables(&tables, &g), "Geometry");
    handle_.reset(g, lbm_geometry_destroy);
    check(lbm_geometry_get_info(g, &info_), "Geometry info");
    spacing = info_.spacing;
    GeomSize = {info_.grid[0], info_.grid[1], info_.grid[2]};
  }
  lbm_geometry* handle() const { return handle_.get(); }
  const lbm_geometry_info& info() const { return info_; }

  std::shared_ptr<LatticeModel> velSet_;
  SimulationParameters simParams;
  double spacing = 0.0;
  std::vector<int> GeomSize;

 private:
  std::shared_ptr<lbm_geometry> handle_;
  lbm_geometry_info info_{};
};

// Boundary::ApplyBc (Boundary.h:61, Boundary.cpp:111-289) on one node's
// populations: the reference's Zou-He inlet rule, evaluated by the device
// code the fused step runs for inlet nodes (lbm_zouhe).
class Boundary {
 public:
  explicit Boundary(int device = 0) : m_distributions(19, 0.0), device_(device) {}
  void ApplyBc(double& vel, std::vector<double>& normal) {
    if (normal.size() < 3 || m_distributions.size() != 19)
      throw std::invalid_argument("ApplyBc: 19 populations and a 3-component normal");
    check(lbm_zouhe(device_, 1, m_distributions.data(), &vel, normal.data()), "ApplyBc");
  }
  std::vector<double> m_distributions;

 private:
  int device_;
};

// SubDomain: node lists of one rank (Subdomain.cpp:43-62). On one GPU the
// whole lattice is one subdomain.
class SubDomain {
 public:
  explicit SubDomain(Geometry& g) : Geometry_(g) {}
  int64_t NumberOfNodes() const { return Geometry_.info().n_nodes; }
  Geometry& Geometry_;
};

class Simulation {
 public:
  // Simulation.cpp:44-58: set up and run the whole loop.
  Simulation(SubDomain& sub, SimulationParameters& simParams) : Simulation(sub, simParams, nullptr) { Iterate(); }

  // Set-up only, with solver parameters beyond the reference's (precision,
  // boundary modes); `extra` may be null.
  Simulation(SubDomain& sub, SimulationParameters& simParams, const lbm_params* extra)
      : SubDomain_(sub), simParams(simParams) {
    lbm_params p;
    if (extra) p = *extra;
    else lbm_params_default(&p);
    p.tau = simParams.tau;
    p.dt = simParams.dt;
    p.cte_vel = simParams.CteVel;
    p.flow_type = simParams.FlowType == "Pulsatile" ? LBM_FLOW_PULSATILE : LBM_FLOW_CONSTANT;
    SetRelaxation(simParams.tau);
    m_dt = simParams.dt;
    lbm_solver* s = nullptr;
    check(lbm_create(sub.Geometry_.handle(), &p, &s), "Simulation");
    solver_.reset(s, lbm_destroy);
  }

  // One reference iteration per call (see the file comment).
  void Collision() {
    check(lbm_step(solver_.get(), 1), "Collision");
    osi_current_ = false;
  }
  void Communicate() {}      // one GPU: nothing to exchange (Simulation.cpp:270-516)
  void Streaming() {}        // folded into Collision()
  void ApplyZouHeBc() {}     // folded into Collision()
  void ApplyOutletBc() {}    // folded into Collision()
  void ApplyWallBc() {}      // folded into Collision()
  void ComputeWallShearStress(int) {}  // accumulated by the fused launch

  // Simulation::ComputeOSIAndTAWSS (Simulation.cpp:804-826). OSI and TAWSS
  // are a pure function of the stress accumulators, so the per-iteration call
  // only marks them out of date; FetchOSIAndTAWSS (called by WriteOutput)
  // reads them from the device when they are used, instead of a device
  // round trip every iteration.
  void ComputeOSIAndTAWSS(int) { osi_current_ = false; }
  void FetchOSIAndTAWSS() {
    if (osi_current_) return;
    const int64_t n = SubDomain_.Geometry_.info().n_nearwall;
    osi.resize(n);
    tawss.resize(n);
    check(lbm_get_osi_tawss(solver_.get(), osi.data(), tawss.data()), "ComputeOSIAndTAWSS");
    osi_current_ = true;
  }

  // Simulation::getInletVelocity (Simulation.cpp:519-534, pulsatile.h:55-69):
  // the same host function lbm_step evaluates for a pulsatile inlet.
  double getInletVelocity(double x, double period) { return lbm_inlet_velocity(x, period); }

  void SetRelaxation(double tau) { m_relaxation = tau; }

  // Simulation::Iterate (Simulation.cpp:538-618): loop condition, output
  // cadence and progress line of the reference.
  void Iterate() {
    double total_runtime = 0.0;
    int iter = 0;
    int systole_count = 0, diastole_count = 0;
    const double tolerance = m_dt / 2;
    const double period = 0.9;
    while ((simParams.FlowType == "Constant" && iter < simParams.MaxIter) ||
           (simParams.FlowType == "Pulsatile" && total_runtime <= 2.7)) {
      Collision();
      Communicate();
      Streaming();
      ApplyZouHeBc();
      ApplyOutletBc();
      ApplyWallBc();
      ComputeWallShearStress(iter);
      if (simParams.FlowType == "Pulsatile") {
        ComputeOSIAndTAWSS(iter);
        const double time_in_cycle = std::fmod(total_runtime, period);
        if (std::fabs(time_in_cycle - simParams.Systole) < tolerance)
          WriteOutput(iter, "Systole_" + std::to_string(++systole_count));
        else if (std::fabs(time_in_cycle - simParams.Diastole) < tolerance)
          WriteOutput(iter, "Diastole_" + std::to_string(++diastole_count));
      }
      if (simParams.FlowType == "Constant" && iter > 0 && simParams.IterToWrite > 0 &&
          iter % simParams.IterToWrite == 0)
        WriteOutput(iter, "Iteration_" + std::to_string(iter));
      if (simParams.FlowType == "Pulsatile") total_runtime += m_dt;
      if (verbose) std::printf("Iteration %d has been performed successfully\n", iter);
      iter++;
    }
    iterations_done = iter;
  }

  // Simulation::WriteOutput (Simulation.cpp:835-856) in output_dir: the
  // fluid-variables VTU (Fluid_Variables/<iter>.vtu), the near-wall CSV
  // (<iter>_FlowHemodynamics.csv: X,Y,Z,WSS[,OSI,TAWSS],VelMag) and its
  // smoothed copy (FlowHemodynamics/<phase>_<iter>_FlowHemodynamics.csv).
  void WriteOutput(int iter, const std::string& phase_name) {
    if (output_dir.empty()) return;
    const bool pulsatile = simParams.FlowType == "Pulsatile";
    if (pulsatile) FetchOSIAndTAWSS();
    check(lbm_write_output(SubDomain_.Geometry_.handle(), solver_.get(), iter, phase_name.c_str(),
                           output_dir.c_str(), pulsatile ? 1 : 0),
          "WriteOutput");
  }

  lbm_solver* solver() const { return solver_.get(); }

  SubDomain& SubDomain_;
  SimulationParameters simParams;
  double m_relaxation = 1.0;
  double m_dt = 0.0;
  std::vector<double> osi, tawss;
  std::string output_dir;  // empty: WriteOutput writes nothing
  bool verbose = false;
  int iterations_done = 0;

 private:
  std::shared_ptr<lbm_solver> solver_;
  bool osi_current_ = false;
};

}  // namespace reza
