// The reference's +z ray test, shared by the host voxeliser (geometry.cpp,
// g++) and its device version (voxel.hip, hipcc): one source, so both
// evaluate the same binary64 expressions in the same order (both files are
// built with -ffp-contract=off; division is IEEE on both).
#pragma once

#include <cmath>

#include "stl.h"

#if defined(__HIPCC__)
#define LBM_HOST_DEVICE __host__ __device__
#else
#define LBM_HOST_DEVICE
#endif

namespace lbm {

// Geometry::isPointInsideTriangle (ComplexGeometry.cpp:208-262) for the +z
// ray direction used by labelPoints.
LBM_HOST_DEVICE inline bool inside_triangle_zray(double x, double y, double z, const Triangle& T) {
  const double x1 = T.x1, y1 = T.y1, z1 = T.z1;
  const double x2 = T.x2, y2 = T.y2, z2 = T.z2;
  const double x3 = T.x3, y3 = T.y3, z3 = T.z3;
  const double dirX = 0.0, dirY = 0.0, dirZ = 1.0;
  const double uX = x2 - x1, uY = y2 - y1, uZ = z2 - z1;
  const double vX = x3 - x1, vY = y3 - y1, vZ = z3 - z1;
  const double nX = uY * vZ - uZ * vY;
  const double nY = uZ * vX - uX * vZ;
  const double nZ = uX * vY - uY * vX;
  const double nLengthSq = nX * nX + nY * nY + nZ * nZ;
  const double tol = 1e-16;
  if (nLengthSq < tol * tol) return false;
  const double w0X = x - x1, w0Y = y - y1, w0Z = z - z1;
  const double a = -(nX * w0X + nY * w0Y + nZ * w0Z);
  const double b = nX * dirX + nY * dirY + nZ * dirZ;
  if (fabs(b) < tol) return false;
  const double pos = a / b;
  const double ix = x + pos * dirX;
  const double iy = y + pos * dirY;
  const double iz = z + pos * dirZ;
  const double uu = uX * uX + uY * uY + uZ * uZ;
  const double uv = uX * vX + uY * vY + uZ * vZ;
  const double vv = vX * vX + vY * vY + vZ * vZ;
  const double wX = ix - x1, wY = iy - y1, wZ = iz - z1;
  const double wu = wX * uX + wY * uY + wZ * uZ;
  const double wv = wX * vX + wY * vY + wZ * vZ;
  const double D = uv * uv - uu * vv;
  const double s = (uv * wv - vv * wu) / D;
  if (s < 0.0 || s > 1.0) return false;
  const double t = (uv * wu - uu * wv) / D;
  if (t < 0.0 || (s + t) > 1.0) return false;
  return true;
}

// One triangle of Geometry::countIntersectinsTriangle (ComplexGeometry.cpp:265-310)
// for a ray from (px,py,pz) to (ex,ey,ez).
LBM_HOST_DEVICE inline bool segment_counts(double px, double py, double pz, double ex, double ey, double ez,
                           const Triangle& T) {
  const double threshold = 1e-16;
  const double dx = ex - px, dy = ey - py, dz = ez - pz;
  const double numerator = (T.x1 - px) * T.nx + (T.y1 - py) * T.ny + (T.z1 - pz) * T.nz;
  const double denominator = dx * T.nx + dy * T.ny + dz * T.nz;
  if (fabs(denominator) < threshold) return false;
  const double t = numerator / denominator;
  if (!(t >= 0.0 && t <= 1.0)) return false;
  const double ix = px + t * dx;
  const double iy = py + t * dy;
  const double iz = pz + t * dz;
  return inside_triangle_zray(ix, iy, iz, T);
}

}  // namespace lbm
