// Per-triangle terms of the near-wall normal sum (Subdomain.cpp:530-557) and
// the device evaluation of that sum.
#pragma once

#include <string>
#include <vector>

namespace lbm {

struct TriTerms {
  double nx, ny, nz;  // unit normal from the vertex cross product (0 if degenerate)
  double area;        // 0.5 * |cross|
  double cx, cy, cz;  // centroid
  double pad;
};

// Normals of `pts` (3 per point) on HIP device `device`; false + err on a
// HIP error (no device, allocation).
bool device_boundary_normals(const std::vector<double>& pts, const std::vector<TriTerms>& tt,
                             std::vector<double>& out, int device, std::string* err);

}  // namespace lbm
