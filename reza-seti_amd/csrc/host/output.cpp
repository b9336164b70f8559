// Output writers (see output.h for the reference map).
#include "output.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace lbm {

namespace {

void put_block(std::string& blob, const void* data, uint64_t bytes) {
  blob.append(reinterpret_cast<const char*>(&bytes), sizeof(bytes));
  blob.append(static_cast<const char*>(data), bytes);
}

}  // namespace

bool write_vtu(const Tables& t, const std::vector<double>& rho, const std::vector<double>& u,
               const std::string& path, std::string* err) {
  const int64_t N = t.N;
  const int NX = t.grid.NX, NY = t.grid.NY, NZ = t.grid.NZ;
  if (static_cast<int64_t>(rho.size()) != N || static_cast<int64_t>(u.size()) != 3 * N) {
    if (err) *err = "write_vtu: field sizes do not match the node count";
    return false;
  }
  // Nodes in the writer's z-y-x loop order (OutputFile.cpp:80-112).
  std::vector<int32_t> order(N);
  for (int64_t n = 0; n < N; ++n) order[n] = static_cast<int32_t>(n);
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    const int32_t* A = &t.ijk[3 * size_t(a)];
    const int32_t* B = &t.ijk[3 * size_t(b)];
    if (A[2] != B[2]) return A[2] < B[2];
    if (A[1] != B[1]) return A[1] < B[1];
    return A[0] < B[0];
  });
  std::unordered_map<int64_t, int64_t> pid;  // grid index -> point id
  pid.reserve(static_cast<size_t>(N) * 2);
  std::vector<double> pts(3 * N), vel(3 * N), den(N), mag(N);
  for (int64_t p = 0; p < N; ++p) {
    const int32_t n = order[p];
    const int32_t* c = &t.ijk[3 * size_t(n)];
    pid[t.grid.gidx(c[0], c[1], c[2])] = p;
    for (int a = 0; a < 3; ++a) {
      pts[3 * p + a] = static_cast<double>(c[a]);
      vel[3 * p + a] = u[3 * size_t(n) + a];
    }
    den[p] = rho[n];
    const double* v = &u[3 * size_t(n)];
    mag[p] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  }
  // Hexahedra over cubes whose corners are all nodes (OutputFile.cpp:180-219).
  std::vector<int64_t> conn, offs;
  static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int z = 0; z < NZ - 2; ++z)
    for (int y = 0; y < NY - 2; ++y)
      for (int x = 0; x < NX - 2; ++x) {
        int64_t ids[8];
        bool ok = true;
        for (int c = 0; c < 8 && ok; ++c) {
          auto it = pid.find(t.grid.gidx(x + corner[c][0], y + corner[c][1], z + corner[c][2]));
          if (it == pid.end()) ok = false;
          else ids[c] = it->second;
        }
        if (!ok) continue;
        conn.insert(conn.end(), ids, ids + 8);
        offs.push_back(static_cast<int64_t>(conn.size()));
      }
  const std::vector<uint8_t> types(offs.size(), 12);  // VTK_HEXAHEDRON
  std::string blob;
  std::vector<uint64_t> at;
  auto add = [&](const void* d, uint64_t bytes) {
    at.push_back(blob.size());
    put_block(blob, d, bytes);
  };
  add(vel.data(), vel.size() * 8);
  add(den.data(), den.size() * 8);
  add(mag.data(), mag.size() * 8);
  add(pts.data(), pts.size() * 8);
  add(conn.data(), conn.size() * 8);
  add(offs.data(), offs.size() * 8);
  add(types.data(), types.size());
  std::ofstream f(path, std::ios::binary);
  if (!f) {
    if (err) *err = "cannot open " + path;
    return false;
  }
  f << "<?xml version=\"1.0\"?>\n"
    << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
    << "  <UnstructuredGrid>\n"
    << "    <Piece NumberOfPoints=\"" << N << "\" NumberOfCells=\"" << offs.size() << "\">\n"
    << "      <PointData>\n"
    << "        <DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"appended\" offset=\""
    << at[0] << "\"/>\n"
    << "        <DataArray type=\"Float64\" Name=\"Density\" format=\"appended\" offset=\"" << at[1] << "\"/>\n"
    << "        <DataArray type=\"Float64\" Name=\"VelocityMagnitude\" format=\"appended\" offset=\"" << at[2]
    << "\"/>\n"
    << "      </PointData>\n"
    << "      <Points>\n"
    << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" << at[3] << "\"/>\n"
    << "      </Points>\n"
    << "      <Cells>\n"
    << "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"appended\" offset=\"" << at[4] << "\"/>\n"
    << "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" offset=\"" << at[5] << "\"/>\n"
    << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" << at[6] << "\"/>\n"
    << "      </Cells>\n"
    << "    </Piece>\n"
    << "  </UnstructuredGrid>\n"
    << "  <AppendedData encoding=\"raw\">\n   _";
  f.write(blob.data(), static_cast<std::streamsize>(blob.size()));
  f << "\n  </AppendedData>\n</VTKFile>\n";
  if (!f) {
    if (err) *err = "write failed: " + path;
    return false;
  }
  return true;
}

bool write_surface_csv(const Tables& t, const std::vector<double>& wss, const std::vector<double>& osi,
                       const std::vector<double>& tawss, const std::vector<double>& velmag, bool pulsatile,
                       const std::string& path, std::string* err) {
  const size_t n = t.nearwall.size();
  if (wss.size() != n || velmag.size() != n || (pulsatile && (osi.size() != n || tawss.size() != n))) {
    if (err) *err = "write_surface_csv: field sizes do not match the near-wall count";
    return false;
  }
  std::ofstream f(path);
  if (!f) {
    if (err) *err = "cannot open " + path;
    return false;
  }
  f << "X,Y,Z,WSS";
  if (pulsatile) f << ",OSI,TAWSS";
  f << ",VelMag\n";
  for (size_t i = 0; i < n; ++i) {
    const int32_t* c = &t.ijk[3 * size_t(t.nearwall[i])];
    f << t.grid.px(c[0]) << "," << t.grid.py(c[1]) << "," << t.grid.pz(c[2]) << "," << wss[i];
    if (pulsatile) f << "," << osi[i] << "," << tawss[i];
    f << "," << velmag[i] << "\n";
  }
  return static_cast<bool>(f);
}

bool smooth_surface_csv(const std::string& in_path, const std::string& out_path, double radius, bool pulsatile,
                        std::string* err) {
  struct Row {
    double x, y, z, wss, osi, tawss, vm;
  };
  std::ifstream in(in_path);
  std::string header, line;
  if (!in || !std::getline(in, header)) {
    if (err) *err = "cannot read " + in_path;
    return false;
  }
  std::vector<Row> rows;
  while (std::getline(in, line)) {
    std::stringstream ss(line);
    std::string tok;
    Row r{};
    try {
      std::getline(ss, tok, ',');
      r.x = std::stod(tok);
      std::getline(ss, tok, ',');
      r.y = std::stod(tok);
      std::getline(ss, tok, ',');
      r.z = std::stod(tok);
      std::getline(ss, tok, ',');
      r.wss = std::stod(tok);
      if (pulsatile) {
        std::getline(ss, tok, ',');
        r.osi = std::stod(tok);
        std::getline(ss, tok, ',');
        r.tawss = std::stod(tok);
      }
      std::getline(ss, tok, ',');
      r.vm = std::stod(tok);
    } catch (const std::exception&) {
      continue;  // a row that does not parse is skipped (OutputFile.cpp:449-451)
    }
    rows.push_back(r);
  }
  // Neighbours within the radius through a uniform cell grid; each mean sums
  // its neighbours in row order, as the reference's full scan does.
  const double cell = radius > 0 ? radius : 1.0;
  auto key = [&](double v) { return static_cast<int64_t>(std::floor(v / cell)); };
  auto pack = [](int64_t a, int64_t b, int64_t c) { return (a * 73856093) ^ (b * 19349663) ^ (c * 83492791); };
  std::unordered_map<int64_t, std::vector<int32_t>> grid;
  for (size_t i = 0; i < rows.size(); ++i) grid[pack(key(rows[i].x), key(rows[i].y), key(rows[i].z))].push_back(int32_t(i));
  std::vector<Row> out = rows;
  std::vector<int32_t> nb;
  for (size_t i = 0; i < rows.size(); ++i) {
    const Row& p = rows[i];
    nb.clear();
    const int64_t cx = key(p.x), cy = key(p.y), cz = key(p.z);
    for (int64_t a = cx - 1; a <= cx + 1; ++a)
      for (int64_t b = cy - 1; b <= cy + 1; ++b)
        for (int64_t c = cz - 1; c <= cz + 1; ++c) {
          auto it = grid.find(pack(a, b, c));
          if (it == grid.end()) continue;
          for (int32_t j : it->second) {
            const Row& q = rows[j];
            // OutputFile::calculateDistance (:356-359), then 0 < d <= radius.
            const double d = std::sqrt(std::pow(q.x - p.x, 2) + std::pow(q.y - p.y, 2) + std::pow(q.z - p.z, 2));
            if (d > 0 && d <= radius) nb.push_back(j);
          }
        }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    if (!nb.empty()) {
      double sw = 0.0, so = 0.0, st = 0.0;
      for (int32_t j : nb) {
        sw += rows[j].wss;
        so += rows[j].osi;
        st += rows[j].tawss;
      }
      out[i].wss = sw / nb.size();
      if (pulsatile) {
        out[i].osi = so / nb.size();
        out[i].tawss = st / nb.size();
      }
    }
  }
  std::ofstream f(out_path);
  if (!f) {
    if (err) *err = "cannot open " + out_path;
    return false;
  }
  f << header << "\n";
  for (const Row& r : out) {
    f << r.x << "," << r.y << "," << r.z << "," << r.wss;
    if (pulsatile) f << "," << r.osi << "," << r.tawss;
    f << "," << r.vm << "\n";
  }
  return static_cast<bool>(f);
}

}  // namespace lbm
