#include "stl.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace lbm {

namespace {

// Whitespace-token cursor over one line, mimicking `ss >> token` / `ss >> x`.
struct Cursor {
  const std::string& s;
  size_t i = 0;
  bool failed = false;
  explicit Cursor(const std::string& line) : s(line) {}
  void skip_ws() {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
  }
  std::string word() {
    if (failed) return "";
    skip_ws();
    if (i >= s.size()) { failed = true; return ""; }
    const size_t a = i;
    while (i < s.size() && !std::isspace(static_cast<unsigned char>(s[i]))) ++i;
    return s.substr(a, i - a);
  }
  bool number(double& v) {
    if (failed) return false;
    skip_ws();
    if (i >= s.size()) { failed = true; return false; }
    const char c = s[i];
    // iostream extraction accepts sign, digits and '.', not inf/nan/hex.
    if (!(std::isdigit(static_cast<unsigned char>(c)) || c == '+' || c == '-' || c == '.')) {
      failed = true;
      return false;
    }
    if ((s.compare(i, 2, "0x") == 0) || (s.compare(i, 2, "0X") == 0)) {
      failed = true;
      return false;
    }
    const char* begin = s.c_str() + i;
    char* end = nullptr;
    const double x = std::strtod(begin, &end);
    if (end == begin) { failed = true; return false; }
    i += static_cast<size_t>(end - begin);
    v = x;
    return true;
  }
};

}  // namespace

void parseSTLASCII(const std::string& text, std::vector<Triangle>& out) {
  std::vector<std::string> lines;
  {
    std::istringstream in(text);
    std::string l;
    while (std::getline(in, l)) lines.push_back(l);
  }
  const std::string empty;
  size_t li = 0;
  auto next = [&]() -> const std::string& { return li < lines.size() ? lines[li++] : (++li, empty); };

  Triangle tri{};  // persists across facets like the reference's local
  while (li < lines.size()) {
    {
      Cursor c(lines[li++]);
      const std::string facet = c.word();
      const std::string normal = c.word();
      c.number(tri.nx) && c.number(tri.ny) && c.number(tri.nz);
      if (c.failed || facet != "facet" || normal != "normal") continue;
    }
    {
      Cursor c(next());
      const std::string outer = c.word();
      const std::string loop = c.word();
      if (c.failed || outer != "outer" || loop != "loop") continue;
    }
    for (int v = 0; v < 3; ++v) {
      Cursor c(next());
      const std::string vertex = c.word();
      double x = 0, y = 0, z = 0;
      c.number(x) && c.number(y) && c.number(z);
      if (c.failed || vertex != "vertex") break;
      if (v == 0) { tri.x1 = x; tri.y1 = y; tri.z1 = z; }
      if (v == 1) { tri.x2 = x; tri.y2 = y; tri.z2 = z; }
      if (v == 2) { tri.x3 = x; tri.y3 = y; tri.z3 = z; }
    }
    {
      Cursor c(next());
      const std::string endloop = c.word();
      if (c.failed || endloop != "endloop") continue;
    }
    {
      Cursor c(next());
      const std::string endfacet = c.word();
      if (c.failed || endfacet != "endfacet") continue;
    }
    out.push_back(tri);
  }
}

bool readSTLASCII(const std::string& filename, std::vector<Triangle>& out) {
  out.clear();
  std::ifstream f(filename, std::ios::binary);
  if (!f.is_open()) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  parseSTLASCII(ss.str(), out);
  return true;
}

}  // namespace lbm
