/*
 * ORACLE — test infrastructure only. Never linked into, called by or shipped
 * with the product library. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker / CPU
 * baseline.
 *
 * A plain-C restatement of the reference solver (Resacd/Reza-Seti) on a
 * single rank:
 *   refcpu.c  — one time step exactly as Simulation::Iterate runs it
 *               (Simulation.cpp:553-568), push streaming on node-major FP64
 *               arrays, with the reference's arithmetic order.
 *   refgeom.c — the setup chain by brute force: every grid point against
 *               every triangle (ComplexGeometry.cpp:83-662), node kinds and
 *               connectivity by dense grid lookup (Subdomain.cpp:364-611).
 *
 * Pinning (what checks each function, besides the GPU comparison):
 *   lattice tables      closed forms (sum w = 1, sum w c = 0, sum w cc = I/3,
 *                       opp(opp(q)) = q, c_opp = -c)
 *   collision           f = f_eq fixed point; moment identities sum f_eq = rho,
 *                       sum c f_eq = rho u; brute force vs a direct expansion
 *   streaming           mass/momentum conservation in a closed periodic box;
 *                       f_q(x + c_q) = f*_q(x) on a periodic box
 *                       (tests/test_oracle_rules.py)
 *   Zou-He inlet        after the rule, rho and u_z of a -z inlet node equal
 *                       rhoN and U (Zou-He identities); compiled reference
 *                       Boundary::ApplyBc where oracle/_ref builds
 *   outlet / wall copy, hand-made node sets with answers derived by hand
 *   stale links, 0.0    (tests/test_oracle_rules.py): a lone wall node, a
 *   push rule           fluid-fluid-outlet and a fluid-wall-outlet column, a
 *                       zero wall node beside a fluid node
 *   WSS                 plane shear closed form (tau tangential, analytic)
 *   voxeliser/classes   compiled reference Geometry/SubDomain where
 *                       oracle/_ref builds; analytic tube/sphere counts
 *   normals             radial for a cylinder (tolerance); compiled reference
 *   extension modes     (half-way bounce-back, NEE inlet/outlet, Guo force)
 *                       have no reference counterpart: pinned by Poiseuille
 *                       closed form, boundary identities (NEE moments and
 *                       non-equilibrium parts, tests/test_oracle_rules.py)
 *                       and conservation.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ lattices -- */
int oracle_lattice(int q, int32_t* c /* 3q */, int32_t* opp, double* w);

/* --------------------------------------------------------------- solver -- */
typedef struct {
  int32_t Q;
  int64_t N;
  int32_t grid[3];
  int32_t periodic[3];
  const int32_t* ijk;         /* 3N, reference node order */
  const uint8_t* kind;        /* N: 0 fluid, 1 wall, 2 inlet, 3 outlet */
  const int32_t* outlet_id;   /* N */
  int32_t n_outlets;
  const double* outlet_normals; /* 3 per outlet */
  double inlet_normal[3];
  int64_t n_nearwall;
  const int32_t* nearwall;    /* node indices */
  const double* normals;      /* 3 per near-wall node */
  double tau;
  double wss_C;
  int32_t wall_mode;          /* 0 reference copy, 1 half-way bounce-back */
  int32_t inlet_mode;         /* 0 reference Zou-He, 1 NEE velocity, 2 none */
  int32_t outlet_mode;        /* 0 reference copy, 1 NEE pressure, 2 none */
  double outlet_rho;
  double force[3];            /* Guo body force (extension) */
  const double* inlet_profile;  /* per inlet node (creation order) factor, NULL = 1 */
} oracle_setup;

typedef struct oracle_state oracle_state;

oracle_state* oracle_create(const oracle_setup* s);
void oracle_destroy(oracle_state* st);
/* f = w_q everywhere, f* = 0, accumulators 0 (Node::Initialize). */
void oracle_reset(oracle_state* st);
void oracle_set_f(oracle_state* st, const double* f);    /* node-major N*Q */
void oracle_get_f(const oracle_state* st, double* f);    /* m_distributions */
void oracle_get_fs(const oracle_state* st, double* fs);  /* m_newDistributions */
void oracle_get_feq(const oracle_state* st, double* feq);
/* Density and velocity set by the last collision (node->density, node->velocity
 * of Simulation::Collision, Simulation.cpp:77-139): N and 3N values. */
void oracle_get_macro(const oracle_state* st, double* rho, double* u);
/* wss, acc (3 per node), acc_mag in the setup's near-wall order. */
void oracle_get_wss(const oracle_state* st, double* wss, double* acc, double* acc_mag);
/* Runs nsteps iterations with inlet speed u_in[t]; nthreads <= 1 is serial.
 * Returns 0. */
int oracle_step(oracle_state* st, int32_t nsteps, const double* u_in, int32_t nthreads);
/* Simulation::Collision alone (used by the fixed-point tests). */
void oracle_collide(oracle_state* st);
/* Boundary::ApplyBc restated, on one population vector (D3Q19). */
void oracle_zouhe(double* f, double vel, const double* normal);
/* Wall shear stress vector of one node (Simulation.cpp:775-797). */
void oracle_wss_tau(int Q, const double* f, const double* feq, const double* n3, double C, double* tau3);
/* Simulation::getInletVelocity(x, period). */
double oracle_inlet_velocity(double x, double period);

/* ------------------------------------------------------------- geometry -- */
typedef struct {
  double x1, y1, z1, x2, y2, z2, x3, y3, z3, nx, ny, nz;
} oracle_triangle;

typedef struct {
  char plane[16];
  double coordinate;
  double range_x[2], range_y[2], range_z[2];
  double normal[3];
} oracle_block;

/* Grid from triangles and config grid_size: bbox[6], dims[3], spacing. */
void oracle_grid(const oracle_triangle* t, int64_t nt, double grid_size, double* bbox, int32_t* dims,
                 double* spacing);
/* Ray-parity labels, every point against every triangle. */
void oracle_label_points(const oracle_triangle* t, int64_t nt, const double* bbox, const int32_t* dims,
                         double spacing, uint8_t* fluid, int32_t nthreads);
/* Node classification from labels. Outputs sized by the caller with N =
 * number of fluid points: ijk (3N), kind (N), outlet_id (N), nb (Q*N push
 * table). Returns N. */
int64_t oracle_classify(const uint8_t* fluid, const int32_t* dims, const double* bbox, double spacing,
                        const oracle_block* inlets, int32_t n_in, const oracle_block* outlets, int32_t n_out,
                        int32_t Q, int32_t* ijk, uint8_t* kind, int32_t* outlet_id, int32_t* nb);
/* Near-wall normal of one point (Subdomain.cpp:519-570). */
void oracle_normal(const oracle_triangle* t, int64_t nt, double x, double y, double z, double* n3);

#ifdef __cplusplus
}
#endif

#endif
