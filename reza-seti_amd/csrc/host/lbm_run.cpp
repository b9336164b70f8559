// lbm_run: the reference's driver (LBM.cpp:60-125) over the MI355X path.
//
//   lbm_run [config.txt] [--out DIR] [--verbose] [--dump FILE] [--steps N]
//
// Same construction order as the reference main: parse config.txt, build the
// lattice model, the geometry, the subdomain, then run Simulation::Iterate.
// --dump writes the final post-boundary populations (FP64, node-major, the
// reference's Involved_Lat_ order) for comparison; --steps overrides MaxIter.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "reza/Simulation.h"

int main(int argc, char** argv) {
  std::string config = "config.txt", out_dir, dump;
  bool verbose = false;
  int steps = -1;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--out" && i + 1 < argc) out_dir = argv[++i];
    else if (a == "--dump" && i + 1 < argc) dump = argv[++i];
    else if (a == "--steps" && i + 1 < argc) steps = std::atoi(argv[++i]);
    else if (a == "--verbose") verbose = true;
    else config = a;
  }
  const auto t0 = std::chrono::steady_clock::now();
  try {
    reza::SimulationParameters simParams;
    // The reference swallows a parse failure and runs with zero parameters
    // (LBM.cpp:94-99); here it is reported and the run stops.
    reza::parseInputFile(config, simParams);
    if (steps >= 0) simParams.MaxIter = steps;
    auto velSet = std::make_shared<reza::LatticeModel>(19);
    reza::Geometry geometry(velSet, simParams, config);
    reza::SubDomain subdomain(geometry);
    reza::Simulation sim(subdomain, simParams, nullptr);
    sim.verbose = verbose;
    sim.output_dir = out_dir;
    std::printf("Geometry: %lld nodes (%lld wall, %lld inlet, %lld outlet), %lld near-wall\n",
                (long long)geometry.info().n_nodes, (long long)geometry.info().n_wall,
                (long long)geometry.info().n_inlet, (long long)geometry.info().n_outlet,
                (long long)geometry.info().n_nearwall);
    const auto t1 = std::chrono::steady_clock::now();
    sim.Iterate();
    reza::check(lbm_synchronize(sim.solver()), "synchronize");
    const auto t2 = std::chrono::steady_clock::now();
    const double loop_s = std::chrono::duration<double>(t2 - t1).count();
    std::printf("Iterations: %d, loop %.3f s, %.1f MLUPS\n", sim.iterations_done, loop_s,
                double(geometry.info().n_nodes) * sim.iterations_done / loop_s / 1e6);
    if (!dump.empty()) {
      std::vector<double> f(static_cast<size_t>(geometry.info().n_nodes) * geometry.info().q);
      reza::check(lbm_get_populations(sim.solver(), f.data()), "populations");
      std::ofstream o(dump, std::ios::binary);
      o.write(reinterpret_cast<const char*>(f.data()), static_cast<std::streamsize>(f.size() * sizeof(double)));
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "lbm_run: %s\n", e.what());
    return 1;
  }
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Total time: %.3f s\n", total);
  return 0;
}
