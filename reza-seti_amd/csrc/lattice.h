// Lattice velocity sets shared by host code and HIP kernels.
//
// Index order, weights and opposite maps follow the reference exactly so that
// population dumps line up entry for entry:
//   D3Q19: LatticeModel.cpp:92-147 (weights/velocities), :149-191 (opposite)
//   D3Q27: LatticeModel.cpp:249-313 (weights/velocities), :316-374 (opposite)
// D3Q19 keeps the rest population LAST (index 18); D3Q27 keeps it FIRST (0).
#pragma once

#if defined(__HIPCC__)
#define LBM_HD __host__ __device__
#else
#define LBM_HD
#endif

namespace lbm {

struct D3Q19 {
  static constexpr int Q = 19;
  static constexpr int REST = 18;
  static constexpr int cx[19] = {1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
  static constexpr int cy[19] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0};
  static constexpr int cz[19] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0};
  static constexpr int opp[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};
  // Same double expressions as the reference table (2/36, 1/36, 12/36).
  static constexpr double w[19] = {
      2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      12.0 / 36.0};
};

struct D3Q27 {
  static constexpr int Q = 27;
  static constexpr int REST = 0;
  static constexpr int cx[27] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, 1, -1,
                                 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
  static constexpr int cy[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0,
                                 1, 1, -1, -1, 1, -1, 1, -1, -1, 1, -1, 1};
  static constexpr int cz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1,
                                 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, -1, 1};
  static constexpr int opp[27] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11,
                                  18, 17, 16, 15, 20, 19, 22, 21, 24, 23, 26, 25};
  static constexpr double w[27] = {
      8.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0,
      2.0 / 27.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
      1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
      1.0 / 54.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
      1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
};

// Runtime view of a velocity set for host-side table builders.
struct LatticeView {
  int Q;
  int rest;
  const int* cx;
  const int* cy;
  const int* cz;
  const int* opp;
  const double* w;
};

inline LatticeView lattice_view(int q) {
  if (q == 27) return {27, D3Q27::REST, D3Q27::cx, D3Q27::cy, D3Q27::cz, D3Q27::opp, D3Q27::w};
  return {19, D3Q19::REST, D3Q19::cx, D3Q19::cy, D3Q19::cz, D3Q19::opp, D3Q19::w};
}

}  // namespace lbm
