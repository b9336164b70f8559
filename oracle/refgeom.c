/*
 * ORACLE (test infrastructure; see oracle.h). Brute-force restatement of the
 * reference set-up chain: every grid point is tested against every triangle
 * (ComplexGeometry.cpp:440-514), boundary points by the 26-neighbour test
 * (:517-564), inlet/outlet/wall sets (:567-662), node kinds with
 * CreateNodes precedence (Subdomain.cpp:364-442) and push connectivity
 * (:447-471). Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

void oracle_grid(const oracle_triangle* t, int64_t nt, double grid_size, double* bbox, int32_t* dims,
                 double* spacing) {
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < nt; ++i) {
    mn[0] = dmin(mn[0], dmin(t[i].x1, dmin(t[i].x2, t[i].x3)));
    mn[1] = dmin(mn[1], dmin(t[i].y1, dmin(t[i].y2, t[i].y3)));
    mn[2] = dmin(mn[2], dmin(t[i].z1, dmin(t[i].z2, t[i].z3)));
    mx[0] = dmax(mx[0], dmax(t[i].x1, dmax(t[i].x2, t[i].x3)));
    mx[1] = dmax(mx[1], dmax(t[i].y1, dmax(t[i].y2, t[i].y3)));
    mx[2] = dmax(mx[2], dmax(t[i].z1, dmax(t[i].z2, t[i].z3)));
  }
  bbox[0] = mn[0]; bbox[1] = mx[0]; bbox[2] = mn[1]; bbox[3] = mx[1]; bbox[4] = mn[2]; bbox[5] = mx[2];
  double s = grid_size;
  s *= 2; /* ComplexGeometry.cpp:157 */
  *spacing = s;
  dims[0] = (int)ceil((mx[0] - mn[0]) / s) + 1;
  dims[1] = (int)ceil((mx[1] - mn[1]) / s) + 1;
  dims[2] = (int)ceil((mx[2] - mn[2]) / s) + 1;
}

/* Geometry::isPointInsideTriangle (ComplexGeometry.cpp:208-262). */
static int inside(double x, double y, double z, const oracle_triangle* T) {
  double x1 = T->x1, y1 = T->y1, z1 = T->z1;
  double x2 = T->x2, y2 = T->y2, z2 = T->z2;
  double x3 = T->x3, y3 = T->y3, z3 = T->z3;
  double dirX = 0.0, dirY = 0.0, dirZ = 1.0;
  double uX = x2 - x1, uY = y2 - y1, uZ = z2 - z1;
  double vX = x3 - x1, vY = y3 - y1, vZ = z3 - z1;
  double nX = uY * vZ - uZ * vY;
  double nY = uZ * vX - uX * vZ;
  double nZ = uX * vY - uY * vX;
  double nLengthSq = nX * nX + nY * nY + nZ * nZ;
  const double tol = 1e-16;
  if (nLengthSq < tol * tol) return 0;
  double w0X = x - x1, w0Y = y - y1, w0Z = z - z1;
  double a = -(nX * w0X + nY * w0Y + nZ * w0Z);
  double b = nX * dirX + nY * dirY + nZ * dirZ;
  if (fabs(b) < tol) return 0;
  double pos = a / b;
  double ix = x + pos * dirX;
  double iy = y + pos * dirY;
  double iz = z + pos * dirZ;
  double uu = uX * uX + uY * uY + uZ * uZ;
  double uv = uX * vX + uY * vY + uZ * vZ;
  double vv = vX * vX + vY * vY + vZ * vZ;
  double wX = ix - x1, wY = iy - y1, wZ = iz - z1;
  double wu = wX * uX + wY * uY + wZ * uZ;
  double wv = wX * vX + wY * vY + wZ * vZ;
  double D = uv * uv - uu * vv;
  double s = (uv * wv - vv * wu) / D;
  if (s < 0.0 || s > 1.0) return 0;
  double t = (uv * wu - uu * wv) / D;
  if (t < 0.0 || (s + t) > 1.0) return 0;
  return 1;
}

void oracle_label_points(const oracle_triangle* tr, int64_t nt, const double* bbox, const int32_t* dims,
                         double spacing, uint8_t* fluid, int32_t nthreads) {
  const double rayOffset = 0.01;
  const double rayLength = bbox[5] - bbox[4];
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
  for (int i = 0; i < dims[0]; ++i)
    for (int j = 0; j < dims[1]; ++j)
      for (int k = 0; k < dims[2]; ++k) {
        /* generatePointsInBoundingBox + labelPoints case 0 + adjustLineCoordinates case 0 */
        const double px0 = bbox[0] + i * spacing, py0 = bbox[2] + j * spacing, pz0 = bbox[4] + k * spacing;
        double sx = px0, sy = py0, sz = pz0, ex = px0, ey = py0, ez = pz0 + rayLength + 0.1;
        if (sx == bbox[0]) { sx += rayOffset; ex += rayOffset; }
        if (sx == bbox[1]) { sx -= rayOffset; ex -= rayOffset; }
        if (sy == bbox[2]) { sy += rayOffset; ey += rayOffset; }
        if (sy == bbox[3]) { sy -= rayOffset; ey -= rayOffset; }
        if (sz == bbox[4]) sz += rayOffset;
        if (ez == bbox[5]) ez -= rayOffset;
        /* countIntersectinsTriangle (ComplexGeometry.cpp:265-310) */
        int count = 0;
        for (int64_t ti = 0; ti < nt; ++ti) {
          const oracle_triangle* T = tr + ti;
          const double threshold = 1e-16;
          double dx = ex - sx, dy = ey - sy, dz = ez - sz;
          double numerator = (T->x1 - sx) * T->nx + (T->y1 - sy) * T->ny + (T->z1 - sz) * T->nz;
          double denominator = dx * T->nx + dy * T->ny + dz * T->nz;
          if (fabs(denominator) < threshold) continue;
          double t = numerator / denominator;
          if (t >= 0.0 && t <= 1.0) {
            double ix = sx + t * dx, iy = sy + t * dy, iz = sz + t * dz;
            if (inside(ix, iy, iz, T)) count++;
          }
        }
        fluid[((int64_t)i * dims[1] + j) * dims[2] + k] = (count % 2 == 1) ? 1 : 0;
      }
}

static int block_match(const oracle_block* b, double x, double y, double z) {
  const long double tolerance = 2.5e-2;
  if (strcmp(b->plane, "x") == 0 && fabsl((long double)(x - b->coordinate)) <= tolerance)
    return y >= b->range_y[0] && y <= b->range_y[1] && z >= b->range_z[0] && z <= b->range_z[1];
  if (strcmp(b->plane, "y") == 0 && fabsl((long double)(y - b->coordinate)) <= tolerance)
    return x >= b->range_x[0] && x <= b->range_x[1] && z >= b->range_z[0] && z <= b->range_z[1];
  if (strcmp(b->plane, "z") == 0 && fabsl((long double)(z - b->coordinate)) <= tolerance)
    return x >= b->range_x[0] && x <= b->range_x[1] && y >= b->range_y[0] && y <= b->range_y[1];
  return 0;
}

int64_t oracle_classify(const uint8_t* fl, const int32_t* d, const double* bbox, double sp,
                        const oracle_block* inlets, int32_t n_in, const oracle_block* outlets, int32_t n_out,
                        int32_t Q, int32_t* ijk, uint8_t* kind, int32_t* outlet_id, int32_t* nb) {
  const int64_t P = (int64_t)d[0] * d[1] * d[2];
  int32_t* map = (int32_t*)malloc(sizeof(int32_t) * P);
  int64_t N = 0;
  int32_t c[81], opp[27];
  double w[27];
  oracle_lattice(Q, c, opp, w);
#define G(i, j, k) (((int64_t)(i) * d[1] + (j)) * d[2] + (k))
  for (int i = 0; i < d[0]; ++i)
    for (int j = 0; j < d[1]; ++j)
      for (int k = 0; k < d[2]; ++k) {
        const int64_t g = G(i, j, k);
        map[g] = -1;
        if (!fl[g]) continue;
        map[g] = (int32_t)N;
        ijk[3 * N] = i;
        ijk[3 * N + 1] = j;
        ijk[3 * N + 2] = k;
        /* boundary point? (26 neighbours inside the box, or a box face) */
        int bp = 0;
        for (int dx = -1; dx <= 1 && !bp; ++dx)
          for (int dy = -1; dy <= 1 && !bp; ++dy)
            for (int dz = -1; dz <= 1 && !bp; ++dz) {
              const int a = i + dx, b = j + dy, cc = k + dz;
              if (a >= 0 && a < d[0] && b >= 0 && b < d[1] && cc >= 0 && cc < d[2] && !fl[G(a, b, cc)]) bp = 1;
            }
        if (i == 0 || i == d[0] - 1 || j == 0 || j == d[1] - 1 || k == 0 || k == d[2] - 1) bp = 1;
        kind[N] = 0;
        outlet_id[N] = -1;
        if (bp) {
          const double x = bbox[0] + i * sp, y = bbox[2] + j * sp, z = bbox[4] + k * sp;
          int in = 0;
          for (int b = 0; b < n_in; ++b) in = in || block_match(inlets + b, x, y, z);
          int out = -1;
          for (int b = 0; b < n_out && out < 0; ++b)
            if (block_match(outlets + b, x, y, z)) out = b;
          if (in) kind[N] = 2;
          else if (out >= 0) {
            kind[N] = 3;
            outlet_id[N] = out;
          } else kind[N] = 1;
        }
        N++;
      }
  for (int64_t n = 0; n < N; ++n)
    for (int q = 0; q < Q; ++q) {
      const int a = ijk[3 * n] + c[3 * q], b = ijk[3 * n + 1] + c[3 * q + 1], cc = ijk[3 * n + 2] + c[3 * q + 2];
      int32_t v = -1;
      if (a >= 0 && a < d[0] && b >= 0 && b < d[1] && cc >= 0 && cc < d[2]) v = map[G(a, b, cc)];
      nb[(int64_t)q * N + n] = v;
    }
#undef G
  free(map);
  return N;
}

void oracle_normal(const oracle_triangle* tr, int64_t nt, double x, double y, double z, double* out) {
  double n0 = 0.0, n1 = 0.0, n2 = 0.0;
  for (int64_t i = 0; i < nt; ++i) {
    const oracle_triangle* T = tr + i;
    double ux = T->x2 - T->x1, uy = T->y2 - T->y1, uz = T->z2 - T->z1;
    double vx = T->x3 - T->x1, vy = T->y3 - T->y1, vz = T->z3 - T->z1;
    double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    double norm = sqrt(nx * nx + ny * ny + nz * nz);
    if (norm > 0) { nx /= norm; ny /= norm; nz /= norm; }
    double area = 0.5 * norm;
    double cX = (T->x1 + T->x2 + T->x3) / 3.0, cY = (T->y1 + T->y2 + T->y3) / 3.0, cZ = (T->z1 + T->z2 + T->z3) / 3.0;
    double distance = sqrt((x - cX) * (x - cX) + (y - cY) * (y - cY) + (z - cZ) * (z - cZ));
    double weight = area / (distance * distance);
    n0 += nx * weight;
    n1 += ny * weight;
    n2 += nz * weight;
  }
  double norm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  if (norm > 0) { n0 /= norm; n1 /= norm; n2 /= norm; }
  out[0] = n0;
  out[1] = n1;
  out[2] = n2;
}
