// Device evaluation of the ray-parity labels (device/voxel.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "stl.h"

namespace lbm {

// Labels of the NX*NY*NZ grid points (index (i*NY + j)*NZ + k; NX = rx.size(),
// NY = ry.size()): ray x per i, ray y per j (after the bounding-box offsets),
// candidate triangles of column c in cand[start[c] .. start[c+1]), points
// z_k = min_z + k*spacing. False + err on a HIP error.
bool device_label_points(const std::vector<Triangle>& tris, const std::vector<int32_t>& cand,
                         const std::vector<int64_t>& start, const std::vector<double>& rx,
                         const std::vector<double>& ry, double min_z, double max_z, double spacing, int NZ,
                         std::vector<uint8_t>& fluid, int device, std::string* err);

}  // namespace lbm
