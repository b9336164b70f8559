// config.txt model and parser with the reference's semantics.
//
// Mirrors `struct SimulationParameters` (SimulationParameters.h:49-89) and
// `parseInputFile` (LBM.h:81-188). Quirks kept on purpose:
//   * a section header is any trimmed line whose last character is ':'
//     (LBM.h:95-98);
//   * an inlet/outlet block runs until a line that is empty BEFORE trimming
//     (LBM.h:163), so whitespace-only lines do not end a block and whatever
//     follows is swallowed into it;
//   * `NumCycle` is parsed with stod and stored in an int (LBM.h:121);
//   * `period`, `pressure` semantics etc. are only stored;
//   * missing ranges default to [-inf, +inf] (LBM.h:165-167);
//   * a block without `coordinate` leaves it unset in the reference (an
//     uninitialised double); here it is NaN, so no point can match it.
#pragma once

#include <string>
#include <vector>

namespace lbm {

struct PlaneSpec {
  std::string plane;          // "x" | "y" | "z"
  double coordinate = 0.0;    // plane coordinate (physical units)
  std::vector<double> range_x, range_y, range_z;
  std::vector<double> normal;
};

struct Equation {
  std::string name;
  std::vector<double> coeffs;
};

struct SimulationParameters {
  std::string stlFile;
  std::string FlowType;
  double dt = 0.0;
  double rho = 0.0;
  double nu_blood = 0.0;
  double tau = 0.0;
  double dx = 0.0;
  double pressure = 0.0;
  double reynoldsNumber = 0.0;
  int num_Out = 0;
  double CteVel = 0.0;
  double Systole = 0.0;
  double Diastole = 0.0;
  int NumCycle = 0;
  int MaxIter = 0;
  int IterToWrite = 0;
  std::vector<Equation> equations;
  std::vector<PlaneSpec> inlets;
  std::vector<PlaneSpec> outlets;
};

// Throws std::runtime_error when the file cannot be opened (LBM.h:83-85).
// Numeric conversion errors propagate as std::invalid_argument /
// std::out_of_range exactly like std::stod/std::stoi in the reference.
void parseInputFile(const std::string& fileName, SimulationParameters& p);

std::string trim(const std::string& s);
std::vector<double> parseRange(const std::string& value);

}  // namespace lbm
