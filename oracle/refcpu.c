/*
 * ORACLE (test infrastructure; see oracle.h). Restates one reference time
 * step (Simulation::Iterate body, Simulation.cpp:553-568) in plain C with the
 * reference's operation order, on node-major FP64 arrays in reference node
 * order. Build with -ffp-contract=off (no FMA), no fast-math.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* LatticeModel.cpp:92-191 (D3Q19) and :249-374 (D3Q27). */
static const int C19[19][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
                               {0, 0, -1}, {1, 1, 0},  {1, -1, 0}, {1, 0, 1},  {1, 0, -1},
                               {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1}, {0, 1, 1},
                               {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {0, 0, 0}};
static const int OPP19[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};
static const int C27[27][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
                               {0, 0, -1},  {1, 1, 0},   {1, -1, 0}, {-1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                               {-1, 0, 1},  {1, 0, -1},  {-1, 0, -1}, {0, 1, 1},  {0, 1, -1},  {0, -1, 1},
                               {0, -1, -1}, {1, 1, 1},   {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},
                               {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
static const int OPP27[27] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15,
                              20, 19, 22, 21, 24, 23, 26, 25};

static void weights(int q, double* w) {
  if (q == 19) {
    for (int i = 0; i < 6; ++i) w[i] = 2.0 / 36.0;
    for (int i = 6; i < 18; ++i) w[i] = 1.0 / 36.0;
    w[18] = 12.0 / 36.0;
  } else {
    w[0] = 8.0 / 27.0;
    for (int i = 1; i < 7; ++i) w[i] = 2.0 / 27.0;
    for (int i = 7; i < 19; ++i) w[i] = 1.0 / 54.0;
    for (int i = 19; i < 27; ++i) w[i] = 1.0 / 216.0;
  }
}

int oracle_lattice(int q, int32_t* c, int32_t* opp, double* w) {
  if (q != 19 && q != 27) return 1;
  for (int i = 0; i < q; ++i) {
    for (int a = 0; a < 3; ++a) c[3 * i + a] = q == 19 ? C19[i][a] : C27[i][a];
    opp[i] = q == 19 ? OPP19[i] : OPP27[i];
  }
  weights(q, w);
  return 0;
}

struct oracle_state {
  oracle_setup s;
  int Q;
  int64_t N;
  int c[27][3];
  int opp[27];
  double w[27];
  double omega;
  int32_t* ijk;
  uint8_t* kind;
  int32_t* outlet_id;
  double* outlet_normals;
  int32_t* nb;      /* N*Q push table */
  int32_t* up;      /* N: node behind an inlet/outlet along its normal, or -1 */
  int32_t* inlets;  /* inlet nodes, creation order */
  int64_t n_in;
  int32_t* outs;    /* outlet nodes, outlet block order then creation order */
  int64_t n_out;
  int outs_serial;  /* some outlet node copies from an outlet node: keep the sequential order */
  int32_t* walls;   /* wall nodes, creation order */
  int64_t n_wall;
  double* prof;     /* inlet profile factor per inlet node */
  int32_t* nw;
  double* nrm;
  int64_t n_nw;
  double *f, *fs, *feq, *rho, *u, *tmp;
  double *wss, *acc, *accm;
};

static int64_t gidx(const int32_t* g, int i, int j, int k) { return ((int64_t)i * g[1] + j) * g[2] + k; }

static int wrap(int v, int n, int per, int* ok) {
  if (v >= 0 && v < n) return v;
  if (per) return (v + n) % n;
  *ok = 0;
  return v;
}

oracle_state* oracle_create(const oracle_setup* s) {
  oracle_state* st = (oracle_state*)calloc(1, sizeof(oracle_state));
  st->s = *s;
  st->Q = s->Q;
  st->N = s->N;
  const int Q = s->Q;
  const int64_t N = s->N;
  int32_t o27[27];
  int32_t c81[81];
  oracle_lattice(Q, c81, o27, st->w);
  for (int q = 0; q < Q; ++q) {
    for (int a = 0; a < 3; ++a) st->c[q][a] = c81[3 * q + a];
    st->opp[q] = o27[q];
  }
  st->omega = 1.0 / s->tau;
  st->ijk = (int32_t*)malloc(sizeof(int32_t) * 3 * N);
  memcpy(st->ijk, s->ijk, sizeof(int32_t) * 3 * N);
  st->kind = (uint8_t*)malloc(N);
  memcpy(st->kind, s->kind, N);
  st->outlet_id = (int32_t*)malloc(sizeof(int32_t) * N);
  memcpy(st->outlet_id, s->outlet_id, sizeof(int32_t) * N);
  st->outlet_normals = (double*)malloc(sizeof(double) * 3 * (s->n_outlets + 1));
  if (s->n_outlets) memcpy(st->outlet_normals, s->outlet_normals, sizeof(double) * 3 * s->n_outlets);

  /* Dense index map -> push neighbours (SubDomain::Connectivity). */
  const int32_t* g = s->grid;
  const int64_t P = (int64_t)g[0] * g[1] * g[2];
  int32_t* map = (int32_t*)malloc(sizeof(int32_t) * P);
  for (int64_t p = 0; p < P; ++p) map[p] = -1;
  for (int64_t n = 0; n < N; ++n) map[gidx(g, st->ijk[3 * n], st->ijk[3 * n + 1], st->ijk[3 * n + 2])] = (int32_t)n;
  st->nb = (int32_t*)malloc(sizeof(int32_t) * N * Q);
  for (int64_t n = 0; n < N; ++n)
    for (int q = 0; q < Q; ++q) {
      int ok = 1;
      const int a = wrap(st->ijk[3 * n] + st->c[q][0], g[0], s->periodic[0], &ok);
      const int b = wrap(st->ijk[3 * n + 1] + st->c[q][1], g[1], s->periodic[1], &ok);
      const int cc = wrap(st->ijk[3 * n + 2] + st->c[q][2], g[2], s->periodic[2], &ok);
      st->nb[n * Q + q] = ok ? map[gidx(g, a, b, cc)] : -1;
    }
  /* Node behind inlet/outlet nodes (Simulation.cpp:655-738 else-if order). */
  st->up = (int32_t*)malloc(sizeof(int32_t) * N);
  for (int64_t n = 0; n < N; ++n) {
    st->up[n] = -1;
    const double* nv = NULL;
    if (st->kind[n] == 3) nv = st->outlet_normals + 3 * st->outlet_id[n];
    if (st->kind[n] == 2) nv = s->inlet_normal;
    if (!nv) continue;
    int d[3] = {0, 0, 0};
    if (nv[0] > 0) d[0] = -1;
    else if (nv[0] < 0) d[0] = 1;
    else if (nv[1] > 0) d[1] = -1;
    else if (nv[1] < 0) d[1] = 1;
    else if (nv[2] > 0) d[2] = -1;
    else if (nv[2] < 0) d[2] = 1;
    else continue;
    int ok = 1;
    const int a = wrap(st->ijk[3 * n] + d[0], g[0], s->periodic[0], &ok);
    const int b = wrap(st->ijk[3 * n + 1] + d[1], g[1], s->periodic[1], &ok);
    const int cc = wrap(st->ijk[3 * n + 2] + d[2], g[2], s->periodic[2], &ok);
    if (ok) st->up[n] = map[gidx(g, a, b, cc)];
  }
  free(map);
  st->inlets = (int32_t*)malloc(sizeof(int32_t) * (N + 1));
  st->n_in = 0;
  for (int64_t n = 0; n < N; ++n)
    if (st->kind[n] == 2) st->inlets[st->n_in++] = (int32_t)n;
  st->outs = (int32_t*)malloc(sizeof(int32_t) * (N + 1));
  st->n_out = 0;
  for (int k = 0; k < s->n_outlets; ++k)
    for (int64_t n = 0; n < N; ++n)
      if (st->kind[n] == 3 && st->outlet_id[n] == k) st->outs[st->n_out++] = (int32_t)n;
  st->outs_serial = 0;
  for (int64_t i = 0; i < st->n_out; ++i)
    if (st->up[st->outs[i]] >= 0 && st->kind[st->up[st->outs[i]]] == 3) st->outs_serial = 1;
  st->walls = (int32_t*)malloc(sizeof(int32_t) * (N + 1));
  st->n_wall = 0;
  for (int64_t n = 0; n < N; ++n)
    if (st->kind[n] == 1) st->walls[st->n_wall++] = (int32_t)n;
  st->prof = (double*)malloc(sizeof(double) * (st->n_in + 1));
  for (int64_t i = 0; i < st->n_in; ++i) st->prof[i] = s->inlet_profile ? s->inlet_profile[i] : 1.0;
  st->n_nw = s->n_nearwall;
  st->nw = (int32_t*)malloc(sizeof(int32_t) * (st->n_nw + 1));
  st->nrm = (double*)malloc(sizeof(double) * 3 * (st->n_nw + 1));
  if (st->n_nw) {
    memcpy(st->nw, s->nearwall, sizeof(int32_t) * st->n_nw);
    memcpy(st->nrm, s->normals, sizeof(double) * 3 * st->n_nw);
  }
  st->f = (double*)malloc(sizeof(double) * N * Q);
  st->fs = (double*)malloc(sizeof(double) * N * Q);
  st->feq = (double*)malloc(sizeof(double) * N * Q);
  st->tmp = (double*)malloc(sizeof(double) * Q * (st->n_in + st->n_out + 1));
  st->rho = (double*)malloc(sizeof(double) * N);
  st->u = (double*)malloc(sizeof(double) * 3 * N);
  st->wss = (double*)malloc(sizeof(double) * (st->n_nw + 1));
  st->acc = (double*)malloc(sizeof(double) * 3 * (st->n_nw + 1));
  st->accm = (double*)malloc(sizeof(double) * (st->n_nw + 1));
  oracle_reset(st);
  return st;
}

void oracle_destroy(oracle_state* st) {
  if (!st) return;
  free(st->ijk); free(st->kind); free(st->outlet_id); free(st->outlet_normals); free(st->nb); free(st->up);
  free(st->inlets); free(st->outs); free(st->walls); free(st->prof); free(st->nw); free(st->nrm); free(st->f); free(st->fs); free(st->feq);
  free(st->tmp); free(st->rho); free(st->u); free(st->wss); free(st->acc); free(st->accm);
  free(st);
}

void oracle_reset(oracle_state* st) {
  const int Q = st->Q;
  for (int64_t n = 0; n < st->N; ++n)
    for (int q = 0; q < Q; ++q) {
      st->f[n * Q + q] = st->w[q];  /* Node::Initialize (Node.cpp:71-78) */
      st->fs[n * Q + q] = 0.0;
      st->feq[n * Q + q] = 0.0;
    }
  for (int64_t i = 0; i < st->n_nw; ++i) {
    st->wss[i] = 0.0;
    st->acc[3 * i] = st->acc[3 * i + 1] = st->acc[3 * i + 2] = 0.0;
    st->accm[i] = 0.0;
  }
}

void oracle_set_f(oracle_state* st, const double* f) { memcpy(st->f, f, sizeof(double) * st->N * st->Q); }
void oracle_get_f(const oracle_state* st, double* f) { memcpy(f, st->f, sizeof(double) * st->N * st->Q); }
void oracle_get_fs(const oracle_state* st, double* fs) { memcpy(fs, st->fs, sizeof(double) * st->N * st->Q); }
void oracle_get_feq(const oracle_state* st, double* feq) { memcpy(feq, st->feq, sizeof(double) * st->N * st->Q); }

void oracle_get_macro(const oracle_state* st, double* rho, double* u) {
  memcpy(rho, st->rho, sizeof(double) * st->N);
  memcpy(u, st->u, sizeof(double) * 3 * st->N);
}

void oracle_get_wss(const oracle_state* st, double* wss, double* acc, double* accm) {
  if (wss) memcpy(wss, st->wss, sizeof(double) * st->n_nw);
  if (acc) memcpy(acc, st->acc, sizeof(double) * 3 * st->n_nw);
  if (accm) memcpy(accm, st->accm, sizeof(double) * st->n_nw);
}

/* Simulation.cpp:100-134 (D3Q19) and :149-207 (D3Q27). */
static void ref_velocity(int Q, const double* f, double rho, double* u) {
  if (Q == 19) {
    u[0] = ((f[0] + f[6] + f[7] + f[8] + f[9]) - (f[1] + f[10] + f[11] + f[12] + f[13])) / rho;
    u[1] = ((f[2] + f[6] + f[10] + f[14] + f[15]) - (f[3] + f[7] + f[11] + f[16] + f[17])) / rho;
    u[2] = ((f[4] + f[8] + f[12] + f[14] + f[16]) - (f[5] + f[9] + f[13] + f[15] + f[17])) / rho;
  } else {
    u[0] = ((f[1] + f[7] + f[8] + f[11] + f[13] + f[19] + f[21] + f[23] + f[25]) -
            (f[2] + f[9] + f[10] + f[12] + f[14] + f[20] + f[22] + f[24] + f[26])) / rho;
    u[1] = ((f[3] + f[7] + f[9] + f[15] + f[16] + f[19] + f[21] + f[24] + f[26]) -
            (f[4] + f[8] + f[10] + f[17] + f[18] + f[20] + f[22] + f[23] + f[25])) / rho;
    u[2] = ((f[5] + f[11] + f[12] + f[15] + f[17] + f[19] + f[22] + f[23] + f[26]) -
            (f[6] + f[13] + f[14] + f[16] + f[18] + f[20] + f[21] + f[24] + f[25])) / rho;
  }
}

static double feq_ref(const oracle_state* st, int q, double rho, const double* u) {
  const double du = u[0] * st->c[q][0] + u[1] * st->c[q][1] + u[2] * st->c[q][2];
  const double u_sqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  return st->w[q] * rho * (1.0 + 3.0 * du + 4.5 * du * du - 1.5 * u_sqr);
}

static int has_force(const oracle_state* st) {
  return st->s.force[0] != 0.0 || st->s.force[1] != 0.0 || st->s.force[2] != 0.0;
}

/* Simulation::Collision (Simulation.cpp:77-232); with a body force, Guo's
 * scheme (extension, no reference counterpart). */
static void collide_node(oracle_state* st, int64_t n) {
  const int Q = st->Q;
  const double* f = st->f + n * Q;
  double* fs = st->fs + n * Q;
  double* feq = st->feq + n * Q;
  double rho = 0.0;
  for (int q = 0; q < Q; ++q) rho += f[q];
  double u[3];
  const int wall0 = st->kind[n] == 1 && st->s.wall_mode == 0;
  if (!has_force(st)) {
    if (wall0) u[0] = u[1] = u[2] = 0.0;
    else ref_velocity(Q, f, rho, u);
    for (int q = 0; q < Q; ++q) feq[q] = feq_ref(st, q, rho, u);
    for (int q = 0; q < Q; ++q) fs[q] = f[q] + st->omega * (feq[q] - f[q]);
  } else {
    double j[3] = {0, 0, 0};
    for (int q = 0; q < Q; ++q)
      for (int a = 0; a < 3; ++a) j[a] += st->c[q][a] * f[q];
    const double* F = st->s.force;
    for (int a = 0; a < 3; ++a) u[a] = wall0 ? 0.0 : (j[a] + 0.5 * F[a]) / rho;
    const double kf = 1.0 - 0.5 * st->omega;
    for (int q = 0; q < Q; ++q) {
      feq[q] = feq_ref(st, q, rho, u);
      const double cu = st->c[q][0] * u[0] + st->c[q][1] * u[1] + st->c[q][2] * u[2];
      const double cF = st->c[q][0] * F[0] + st->c[q][1] * F[1] + st->c[q][2] * F[2];
      const double uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
      fs[q] = f[q] + st->omega * (feq[q] - f[q]) + kf * st->w[q] * (3.0 * (cF - uF) + 9.0 * cu * cF);
    }
  }
  st->rho[n] = rho;
  st->u[3 * n] = u[0];
  st->u[3 * n + 1] = u[1];
  st->u[3 * n + 2] = u[2];
}

void oracle_collide(oracle_state* st) {
  for (int64_t n = 0; n < st->N; ++n) collide_node(st, n);
}

/* Boundary::ApplyBc (Boundary.cpp:111-289), D3Q19, all six branches. */
void oracle_zouhe(double* m, double vel, const double* normal) {
  double v[3];
  if (normal[0] > 0) {
    v[0] = vel; v[1] = 0.0; v[2] = 0.0;
    double rhoN = (m[2] + m[3] + m[4] + m[5] + m[14] + m[15] + m[16] + m[17] + m[18] +
                   2 * (m[0] + m[6] + m[7] + m[8] + m[9])) / (1 + v[0]);
    double Nxy = (1.0 / 2) * (m[2] + m[14] + m[15] - m[3] - m[16] - m[17]) - (1.0 / 3) * (rhoN * v[1]);
    double Nxz = (1.0 / 2) * (m[4] + m[16] + m[14] - m[5] - m[15] - m[17]) - (1.0 / 3) * (rhoN * v[2]);
    m[1] = m[0] - (1 / 3.0) * rhoN * v[0];
    m[10] = m[7] + (1 / 6.0) * rhoN * (-v[0] + v[1]) - Nxy;
    m[11] = m[6] + (1 / 6.0) * rhoN * (-v[0] - v[1]) + Nxy;
    m[13] = m[8] + (1 / 6.0) * rhoN * (-v[0] - v[2]) + Nxz;
    m[12] = m[9] + (1 / 6.0) * rhoN * (-v[0] + v[2]) - Nxz;
  }
  if (normal[0] < 0) {
    v[0] = vel; v[1] = 0.0; v[2] = 0.0;
    double rhoN = (m[2] + m[3] + m[4] + m[5] + m[14] + m[15] + m[16] + m[17] + m[18] +
                   2 * (m[1] + m[10] + m[11] + m[12] + m[13])) / (1 - v[0]);
    double Nxy = (1.0 / 2) * (m[2] + m[14] + m[15] - m[3] - m[16] - m[17]) - (1.0 / 3) * (rhoN * v[1]);
    double Nxz = (1.0 / 2) * (m[4] + m[16] + m[14] - m[5] - m[15] - m[17]) - (1.0 / 3) * (rhoN * v[2]);
    m[0] = m[1] + (1 / 3.0) * rhoN * v[0];
    m[7] = m[10] + (1 / 6.0) * rhoN * (v[0] - v[1]) + Nxy;
    m[6] = m[11] + (1 / 6.0) * rhoN * (v[0] + v[1]) - Nxy;
    m[8] = m[13] + (1 / 6.0) * rhoN * (v[0] + v[2]) - Nxz;
    m[9] = m[12] + (1 / 6.0) * rhoN * (v[0] - v[2]) + Nxz;
  }
  if (normal[1] > 0) {
    v[0] = 0.0; v[1] = 1.0; v[2] = 0.0;
    double rhoN = (m[0] + m[1] + m[4] + m[5] + m[8] + m[9] + m[12] + m[13] + m[18] +
                   2 * (m[2] + m[6] + m[10] + m[14] + m[15])) / (1 + v[1]);
    double Nyx = (1.0 / 2) * (m[0] + m[8] + m[9] - m[1] - m[12] - m[13]) - (1.0 / 3) * (rhoN * v[0]);
    double Nyz = (1.0 / 2) * (m[4] + m[8] + m[12] - m[5] - m[9] - m[13]) - (1.0 / 3) * (rhoN * v[2]);
    m[3] = m[2] - (1 / 3.0) * rhoN * v[1];
    m[11] = m[6] + (1 / 6.0) * rhoN * (-v[0] - v[1]) + Nyx;
    m[7] = m[10] + (1 / 6.0) * rhoN * (-v[1] + v[0]) - Nyx;
    m[17] = m[14] + (1 / 6.0) * rhoN * (-v[1] - v[2]) + Nyz;
    m[16] = m[15] + (1 / 6.0) * rhoN * (-v[1] + v[2]) - Nyz;
  }
  if (normal[1] < 0) {
    v[0] = 0.0; v[1] = 1.0; v[2] = 0.0;
    double rhoN = (m[0] + m[1] + m[4] + m[5] + m[8] + m[9] + m[12] + m[13] + m[18] +
                   2 * (m[3] + m[7] + m[11] + m[16] + m[17])) / (1 - v[1]);
    double Nyx = (1.0 / 2) * (m[0] + m[8] + m[9] - m[1] - m[12] - m[13]) - (1.0 / 3) * (rhoN * v[0]);
    double Nyz = (1.0 / 2) * (m[4] + m[8] + m[12] - m[5] - m[9] - m[13]) - (1.0 / 3) * (rhoN * v[2]);
    m[2] = m[3] + (1 / 3.0) * rhoN * v[1];
    m[6] = m[11] + (1 / 6.0) * rhoN * (v[0] + v[1]) - Nyx;
    m[10] = m[7] + (1 / 6.0) * rhoN * (v[1] - v[0]) + Nyx;
    m[14] = m[17] + (1 / 6.0) * rhoN * (v[1] + v[2]) - Nyz;
    m[15] = m[16] + (1 / 6.0) * rhoN * (v[1] - v[2]) + Nyz;
  }
  if (normal[2] > 0) {
    v[0] = 0.0; v[1] = 0.0; v[2] = vel;
    double rhoN = (m[0] + m[1] + m[4] + m[5] + m[8] + m[9] + m[12] + m[13] + m[18] +
                   2 * (m[3] + m[7] + m[11] + m[16] + m[17])) / (1 + v[2]);
    double Nzx = (1.0 / 2) * (m[0] + m[6] + m[7] - m[1] - m[10] - m[11]) - (1.0 / 3) * (rhoN * v[0]);
    double Nzy = (1.0 / 2) * (m[2] + m[6] + m[10] - m[3] - m[7] - m[11]) - (1.0 / 3) * (rhoN * v[1]);
    m[5] = m[4] - (1 / 3.0) * rhoN * v[2];
    m[9] = m[12] + (1 / 6.0) * rhoN * (v[0] - v[2]) - Nzx;
    m[13] = m[8] + (1 / 6.0) * rhoN * (-v[0] - v[2]) + Nzx;
    m[15] = m[16] + (1 / 6.0) * rhoN * (v[1] - v[2]) - Nzy;
    m[17] = m[14] + (1 / 6.0) * rhoN * (-v[1] - v[2]) + Nzy;
  }
  if (normal[2] < 0) {
    v[0] = 0.0; v[1] = 0.0; v[2] = vel;
    double rhoN = (m[0] + m[1] + m[2] + m[3] + m[6] + m[7] + m[10] + m[11] + m[18] +
                   2 * (m[5] + m[9] + m[13] + m[15] + m[17])) / (1 - v[2]);
    double Nzx = (1.0 / 2) * (m[0] + m[6] + m[7] - m[1] - m[10] - m[11]) - (1.0 / 3) * (rhoN * v[0]);
    double Nzy = (1.0 / 2) * (m[2] + m[6] + m[10] - m[3] - m[7] - m[11]) - (1.0 / 3) * (rhoN * v[1]);
    m[4] = m[5] + (1 / 3.0) * rhoN * v[2];
    m[8] = m[13] + (1 / 6.0) * rhoN * (v[0] + v[2]) - Nzx;
    m[12] = m[9] + (1 / 6.0) * rhoN * (-v[0] + v[2]) + Nzx;
    m[14] = m[17] + (1 / 6.0) * rhoN * (v[1] + v[2]) - Nzy;
    m[16] = m[15] + (1 / 6.0) * rhoN * (-v[1] + v[2]) + Nzy;
  }
}

/* Pulsatile::InletVelequation1..3 (pulsatile.h:55-69) and
 * Simulation::getInletVelocity (Simulation.cpp:519-534). */
double oracle_inlet_velocity(double x, double period) {
  double vel = 0.0;
  const double t = fmod(x, period);
  if (t >= 0 && t < (0.1 * period))
    vel = (-8.579e+08 * pow(t, 8) + 4.059e+08 * pow(t, 7) - 7.817e+07 * pow(t, 6) + 7.994e+06 * pow(t, 5) -
           4.835e+05 * pow(t, 4) + 1.782e+04 * pow(t, 3) - 321.6 * pow(t, 2) + 3.215 * t + 0.1423);
  else if (t >= (0.1 * period) && t < (0.3 * period))
    vel = (-4.32e+06 * pow(t, 8) + 7.94e+06 * pow(t, 7) - 6.298e+06 * pow(t, 6) + 2.82e+06 * pow(t, 5) -
           7.809e+05 * pow(t, 4) + 1.37e+05 * pow(t, 3) - 1.484e+04 * pow(t, 2) + 901.9 * t - 22.87);
  else if (t >= (0.3 * period) && t < (0.9 * period))
    vel = (-5.123 * pow(t, 4) + 14.15 * pow(t, 3) - 14.21 * pow(t, 2) + 5.865 * t - 0.5696);
  return vel;
}

/* Density/velocity/equilibrium helpers for the extension BCs. */
static void rho_u_of(const oracle_state* st, const double* f, double* rho, double* u) {
  double r = 0.0, j[3] = {0, 0, 0};
  for (int q = 0; q < st->Q; ++q) {
    r += f[q];
    for (int a = 0; a < 3; ++a) j[a] += st->c[q][a] * f[q];
  }
  *rho = r;
  for (int a = 0; a < 3; ++a) u[a] = j[a] / r;
}

static void nee(const oracle_state* st, const double* G, double rho_b, const double* u_b, int use_G_u, double* out) {
  double rg, ug[3], ub[3];
  rho_u_of(st, G, &rg, ug);
  for (int a = 0; a < 3; ++a) ub[a] = use_G_u ? ug[a] : u_b[a];
  if (rho_b <= 0) rho_b = rg;
  for (int q = 0; q < st->Q; ++q) out[q] = feq_ref(st, q, rho_b, ub) + (G[q] - feq_ref(st, q, rg, ug));
}

/* Simulation::ComputeWallShearStress body for one node (Simulation.cpp:775-797):
 * tau from f, f_eq and the node normal. */
void oracle_wss_tau(int Q, const double* f, const double* feq, const double* nv, double C, double* tau) {
  int32_t c[81], opp[27];
  double w[27];
  oracle_lattice(Q, c, opp, w);
  for (int a = 0; a < 3; ++a) {
    double sum_alpha = 0.0;
    for (int al = 0; al < Q; ++al) {
      double sum_k = 0.0;
      for (int k = 0; k < 3; ++k) sum_k += c[3 * al + k] * nv[k] * nv[a];
      double sum_j = 0.0;
      for (int j = 0; j < 3; ++j) sum_j += c[3 * al + j] * nv[j] * (c[3 * al + a] - sum_k);
      const double f_ne = f[al] - feq[al];
      sum_alpha += f_ne * sum_j;
    }
    tau[a] = C * sum_alpha;
  }
}

static void wss_node(oracle_state* st, int64_t i) {
  const int Q = st->Q;
  const int64_t n = st->nw[i];
  double tau_i[3];
  oracle_wss_tau(Q, st->f + n * Q, st->feq + n * Q, st->nrm + 3 * i, st->s.wss_C, tau_i);
  const double W = sqrt(tau_i[0] * tau_i[0] + tau_i[1] * tau_i[1] + tau_i[2] * tau_i[2]);
  st->wss[i] = W;
  for (int a = 0; a < 3; ++a) st->acc[3 * i + a] += tau_i[a];
  st->accm[i] += W;
}

static void one_step(oracle_state* st, double U, int nthreads) {
  const int Q = st->Q;
  const int64_t N = st->N;
  const int ref_walls = st->s.wall_mode == 0;
  (void)nthreads;
  /* 1. Collision (Simulation.cpp:553). */
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
  for (int64_t n = 0; n < N; ++n) collide_node(st, n);
  /* 2. Streaming, push (Simulation.cpp:238-263). Each target (nb, q) has one
   *    source, so the loop order does not matter. The f* != 0 skip only
   *    exists in the reference rules. */
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
  for (int64_t n = 0; n < N; ++n)
    for (int q = 0; q < Q; ++q) {
      const int32_t m = st->nb[n * Q + q];
      const double v = st->fs[n * Q + q];
      if (m >= 0 && (!ref_walls || v != 0.0)) st->f[(int64_t)m * Q + q] = v;
    }
  /* Half-way bounce-back (Node::BounceBack, Node.cpp:144-148) on every link
   * whose push neighbour is missing: f[opp](x) = f*[q](x). Each node writes
   * only its own populations, so every boundary loop below runs its nodes in
   * parallel; the sequential order only matters for an outlet node that
   * copies from another outlet node (outs_serial). */
  if (!ref_walls) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t n = 0; n < N; ++n)
      for (int q = 0; q < Q; ++q)
        if (st->nb[n * Q + q] < 0) st->f[n * Q + st->opp[q]] = st->fs[n * Q + q];
  }
  /* 3-4. Inlet (Simulation.cpp:558-565). */
  if (st->s.inlet_mode == 0) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t i = 0; i < st->n_in; ++i)
      oracle_zouhe(st->f + (int64_t)st->inlets[i] * Q, U * st->prof[i], st->s.inlet_normal);
  } else if (st->s.inlet_mode == 1) {
    /* Every inlet node reads its upstream node as streamed, before any
     * inlet node is rewritten: copy those first. */
    for (int64_t i = 0; i < st->n_in; ++i) {
      const int32_t u = st->up[st->inlets[i]];
      if (u >= 0) memcpy(st->tmp + i * Q, st->f + (int64_t)u * Q, sizeof(double) * Q);
    }
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t i = 0; i < st->n_in; ++i) {
      const int64_t n = st->inlets[i];
      if (st->up[n] < 0) continue;
      const double Ui = U * st->prof[i];
      const double ub[3] = {-st->s.inlet_normal[0] * Ui, -st->s.inlet_normal[1] * Ui, -st->s.inlet_normal[2] * Ui};
      nee(st, st->tmp + i * Q, -1.0, ub, 0, st->f + n * Q);
    }
  }
  /* 5. Outlet (Simulation.cpp:650-741): outlet blocks in order, nodes in
   *    creation order, every q: f[opp q](x) = f[opp q](behind). */
  if (st->s.outlet_mode == 0) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1 && !st->outs_serial)
    for (int64_t i = 0; i < st->n_out; ++i) {
      const int64_t n = st->outs[i];
      const int32_t u = st->up[n];
      if (u < 0) continue;
      for (int q = 0; q < Q; ++q) {
        const int o = st->opp[q];
        st->f[n * Q + o] = st->f[(int64_t)u * Q + o];
      }
    }
  } else if (st->s.outlet_mode == 1) {
    double* t = st->tmp + st->n_in * Q;
    for (int64_t i = 0; i < st->n_out; ++i) {
      const int32_t u = st->up[st->outs[i]];
      if (u >= 0) memcpy(t + i * Q, st->f + (int64_t)u * Q, sizeof(double) * Q);
    }
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t i = 0; i < st->n_out; ++i) {
      const int64_t n = st->outs[i];
      if (st->up[n] < 0) continue;
      const double z[3] = {0, 0, 0};
      nee(st, t + i * Q, st->s.outlet_rho, z, 1, st->f + n * Q);
    }
  }
  /* 6. Walls (Simulation.cpp:746-763): q in order, f[opp] = f[q] where the
   *    push neighbour is missing. */
  if (ref_walls) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t i = 0; i < st->n_wall; ++i) {
      const int64_t n = st->walls[i];
      for (int q = 0; q < Q; ++q)
        if (st->nb[n * Q + q] < 0) st->f[n * Q + st->opp[q]] = st->f[n * Q + q];
    }
  }
  /* 7. Wall shear stress (Simulation.cpp:568). */
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
  for (int64_t i = 0; i < st->n_nw; ++i) wss_node(st, i);
}

int oracle_step(oracle_state* st, int32_t nsteps, const double* u_in, int32_t nthreads) {
  for (int32_t t = 0; t < nsteps; ++t) one_step(st, u_in ? u_in[t] : 0.0, nthreads);
  return 0;
}
