// config.txt parser; semantics of LBM.h:61-188 (see config.h for the quirks).
#include "config.h"

#include <cmath>
#include <fstream>
#include <limits>
#include <stdexcept>

namespace lbm {

std::string trim(const std::string& s) {
  // Only blanks and tabs are stripped (LBM.h:61-65).
  const size_t a = s.find_first_not_of(" \t");
  const size_t b = s.find_last_not_of(" \t");
  if (a == std::string::npos || b == std::string::npos) return "";
  return s.substr(a, b - a + 1);
}

// Splits like repeated std::getline(stream, tok, ','): interior empty tokens
// are produced, a trailing empty token after the last comma is not, and an
// empty string yields nothing.
static std::vector<std::string> split_commas(const std::string& s) {
  std::vector<std::string> out;
  size_t pos = 0;
  while (pos < s.size()) {
    const size_t c = s.find(',', pos);
    if (c == std::string::npos) {
      out.push_back(s.substr(pos));
      break;
    }
    out.push_back(s.substr(pos, c - pos));
    pos = c + 1;
  }
  return out;
}

std::vector<double> parseRange(const std::string& value) {
  // LBM.h:67-79: numbers between the first '[' and the first ']'.
  std::vector<double> r;
  const size_t a = value.find('[');
  const size_t b = value.find(']');
  if (a == std::string::npos || b == std::string::npos) return r;
  for (const std::string& tok : split_commas(value.substr(a + 1, b - a - 1)))
    r.push_back(std::stod(trim(tok)));
  return r;
}

static void strip_quotes(std::string& v) {
  std::string o;
  for (char ch : v)
    if (ch != '"') o.push_back(ch);
  v.swap(o);
}

static void parse_parameter(const std::string& line, SimulationParameters& p) {
  // getline(iss, key, ':') && getline(iss, value) (LBM.h:102-104).
  const size_t colon = line.find(':');
  if (colon == std::string::npos) return;          // second getline fails
  if (colon + 1 >= line.size()) return;            // nothing left to read
  const std::string key = trim(line.substr(0, colon));
  std::string value = trim(line.substr(colon + 1));
  if (key == "time_resolution") p.dt = std::stod(value);
  else if (key == "density") p.rho = std::stod(value);
  else if (key == "kinematic_viscosity") p.nu_blood = std::stod(value);
  else if (key == "relaxation_factor") p.tau = std::stod(value);
  else if (key == "grid_size") p.dx = std::stod(value);
  else if (key == "pressure") p.pressure = std::stod(value);
  else if (key == "CteVel") p.CteVel = std::stod(value);
  else if (key == "reynolds_number") p.reynoldsNumber = std::stod(value);
  else if (key == "number_of_outlets") p.num_Out = std::stoi(value);
  else if (key == "IterToWrite") p.IterToWrite = std::stoi(value);
  else if (key == "MaxIter") p.MaxIter = std::stoi(value);
  else if (key == "Systole") p.Systole = std::stod(value);
  else if (key == "Diastole") p.Diastole = std::stod(value);
  else if (key == "NumCycle") p.NumCycle = static_cast<int>(std::stod(value));
  else if (key == "FlowType") { strip_quotes(value); p.FlowType = value; }
  else if (key == "stl_file") { strip_quotes(value); p.stlFile = value; }
}

void parseInputFile(const std::string& fileName, SimulationParameters& p) {
  std::ifstream file(fileName);
  if (!file.is_open()) throw std::runtime_error("Unable to open input file: " + fileName);

  const double inf = std::numeric_limits<double>::infinity();
  std::string raw, section;
  while (std::getline(file, raw)) {
    std::string line = trim(raw);
    if (line.empty() || line[0] == '#') continue;
    if (line.back() == ':') {
      section = line.substr(0, line.size() - 1);
      continue;
    }
    if (section == "parameters") {
      parse_parameter(line, p);
    } else if (section == "inlets" || section == "outlets") {
      PlaneSpec blk;
      blk.coordinate = std::numeric_limits<double>::quiet_NaN();
      // do { ... } while (getline(file, line) && !line.empty());  LBM.h:134-163
      bool more = true;
      std::string cur = line;
      while (more) {
        const std::string t = trim(cur);
        if (!t.empty()) {
          const size_t pos = t.find(':');
          if (pos != std::string::npos) {
            const std::string key = trim(t.substr(0, pos));
            const std::string value = trim(t.substr(pos + 1));
            if (key == "plane") blk.plane = value;
            else if (key == "coordinate") blk.coordinate = std::stod(value);
            else if (key == "range_x") blk.range_x = parseRange(value);
            else if (key == "range_y") blk.range_y = parseRange(value);
            else if (key == "range_z") blk.range_z = parseRange(value);
            else if (key == "normal") blk.normal = parseRange(value);
          }
        }
        more = static_cast<bool>(std::getline(file, cur)) && !cur.empty();
      }
      if (blk.range_x.empty()) blk.range_x = {-inf, inf};
      if (blk.range_y.empty()) blk.range_y = {-inf, inf};
      if (blk.range_z.empty()) blk.range_z = {-inf, inf};
      if (section == "inlets") p.inlets.push_back(blk);
      else p.outlets.push_back(blk);
    } else if (section == "equations") {
      const size_t pos = line.find(':');
      if (pos != std::string::npos) {
        Equation e;
        e.name = trim(line.substr(0, pos));
        e.coeffs = parseRange(trim(line.substr(pos + 1)));
        p.equations.push_back(e);
      }
    }
  }
}

}  // namespace lbm
