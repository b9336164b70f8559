// Node classification and connectivity on the GPU (device/classify.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "config.h"
#include "geometry.h"

namespace lbm {

struct DeviceTables {
  int64_t N = 0;
  std::vector<int32_t> ijk;        // 3N, node order
  std::vector<uint8_t> kind;       // N
  std::vector<int32_t> outlet_id;  // N
  std::vector<int32_t> nb;         // Q*N push table
  int64_t n_boundary_points = 0;
};

// Boundary points, node kinds, node order and the push table of the labelled
// grid, computed on HIP device `device`; the same tables as the host chain.
bool device_classify(const GridSpec& g, const std::vector<uint8_t>& fluid, const std::vector<PlaneSpec>& inlets,
                     const std::vector<PlaneSpec>& outlets, int Q, const int periodic[3], int device,
                     DeviceTables& out, std::string* err);

}  // namespace lbm
