// Private definitions shared by the C ABI translation units.
#pragma once

#include <string>
#include <vector>

#include "../../include/lbm_c_api.h"
#include "host/config.h"
#include "host/geometry.h"

struct lbm_geometry {
  lbm::Tables t;
  std::vector<lbm::Triangle> tris;
};

namespace lbm {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
}  // namespace lbm
