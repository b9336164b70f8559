// ASCII STL reader with the reference's acceptance rules.
//
// Follows Reader::readSTLASCII (GeometryReader.cpp:74-158): a facet is kept
// only when the five lines `facet normal nx ny nz` / `outer loop` /
// 3 x `vertex x y z` / `endloop` / `endfacet` parse in sequence; a block that
// breaks the pattern is skipped by resuming the scan at the next line. The
// triangle record is reused across facets, so fields a malformed block did not
// overwrite keep the previous facet's values, as in the reference. Numbers are
// read with strtod (correctly rounded, the same as iostream extraction for the
// decimal forms STL writers produce).
#pragma once

#include <string>
#include <vector>

namespace lbm {

struct Triangle {
  double x1, y1, z1;
  double x2, y2, z2;
  double x3, y3, z3;
  double nx, ny, nz;  // normal as written in the file (used by the ray test)
};

// Returns false (and an empty list) when the file cannot be opened; the
// reference prints an error and returns an empty list in that case
// (GeometryReader.cpp:77-80).
bool readSTLASCII(const std::string& filename, std::vector<Triangle>& out);

// Same parser over an in-memory buffer (used by tests and the C ABI).
void parseSTLASCII(const std::string& text, std::vector<Triangle>& out);

}  // namespace lbm
