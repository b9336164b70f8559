/*
 * C ABI of the MI355X lattice-Boltzmann collide-stream path.
 *
 * The reference (Resacd/Reza-Seti) has no FFI; its "operator API" is the C++
 * class surface of Simulation / SubDomain / Geometry / Boundary /
 * LatticeModel. Each entry point below names the reference interface it
 * replaces. Plain pointers and sizes only; no C++ or torch types cross this
 * boundary; no C++ exception escapes it. Every call returns LBM_OK (0) or an
 * LBM_ERR_* code, and lbm_last_error() returns the calling thread's message.
 *
 * Node order: every per-node array is in the reference's node order, i.e. the
 * order of SubDomain::Involved_Lat_ (Subdomain.cpp:598-611: grid index i
 * outermost, k innermost, fluid points only). Populations are node-major,
 * f[n * Q + q], with the reference's q order (LatticeModel.cpp:92-191 for
 * D3Q19, :249-374 for D3Q27).
 *
 * A solver handle owns one HIP stream; calls on one handle are not
 * re-entrant, distinct handles are independent.
 */
#ifndef LBM_C_API_H
#define LBM_C_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_ARG 1          /* invalid argument or inconsistent sizes */
#define LBM_ERR_IO 2           /* config / STL could not be read, output could not be written */
#define LBM_ERR_HIP 3          /* HIP runtime error (no device, OOM, launch) */
#define LBM_ERR_UNSUPPORTED 4  /* mode not available for this lattice/precision */
#define LBM_ERR_PARITY 5       /* a reference-mode assumption was violated (see lbm_step) */
#define LBM_ERR_STATE 6        /* call out of order (e.g. step before create) */

/* Node kinds (Struct.h:53-67 BoundaryType, plus plain fluid). */
#define LBM_KIND_FLUID 0
#define LBM_KIND_WALL 1
#define LBM_KIND_INLET 2
#define LBM_KIND_OUTLET 3

/* Population storage / arithmetic. */
#define LBM_PREC_F64 0          /* f stored and computed in FP64, unshifted (reference parity) */
#define LBM_PREC_F32_SHIFTED 1  /* f - w stored in FP32, FP32 collision */
#define LBM_PREC_F16_SHIFTED 2  /* f - w stored in FP16, FP32 collision */

/* Wall treatment. */
#define LBM_WALL_REFERENCE 0    /* Simulation::ApplyWallBc copy rule (Simulation.cpp:746-763) */
#define LBM_WALL_HALFWAY_BB 1   /* Node::BounceBack half-way rule on every missing link (Node.cpp:144-148) */

/* Inlet treatment. */
#define LBM_INLET_REFERENCE_ZOUHE 0  /* Boundary::ApplyBc (Boundary.cpp:111-289), D3Q19 only */
#define LBM_INLET_NEE_VELOCITY 1     /* non-equilibrium extrapolation, per-node velocity profile */
#define LBM_INLET_NONE 2             /* inlet nodes behave as plain fluid nodes */

/* Outlet treatment. */
#define LBM_OUTLET_REFERENCE 0       /* Simulation::ApplyOutletBc copy (Simulation.cpp:650-741) */
#define LBM_OUTLET_NEE_PRESSURE 1    /* non-equilibrium extrapolation at fixed density */
#define LBM_OUTLET_NONE 2

/* Inlet speed source for lbm_step. */
#define LBM_FLOW_CONSTANT 0   /* CteVel every step (Simulation.cpp:558-559) */
#define LBM_FLOW_PULSATILE 1  /* Simulation::getInletVelocity(t, 0.9) (Simulation.cpp:519-534, pulsatile.h) */

/* Replaces Simulation::getInletVelocity (Simulation.cpp:519-534): the
 * piecewise polynomial Pulsatile::InletVelequation1..3 (pulsatile.h:55-69)
 * of fmod(t, period) on [0, 0.1p), [0.1p, 0.3p), [0.3p, 0.9p), else 0. This
 * host function is the inlet speed lbm_step uses for LBM_FLOW_PULSATILE. */
double lbm_inlet_velocity(double t, double period);

/* Lattice tables: VelocitySet::GetDirection / GetWeight and
 * LatticeModel::OppositeDirection (VelocitySet.h:60-65, LatticeModel.h:50-54).
 * c: 3 ints per direction; q is 19 or 27. */
int lbm_lattice(int32_t q, int32_t* c, int32_t* opp, double* w);

/* ------------------------------------------------------------------------ */
/* Configuration                                                             */
/* ------------------------------------------------------------------------ */

#define LBM_MAX_BLOCKS 16

typedef struct {
  char plane[16];
  double coordinate;
  double range_x[2], range_y[2], range_z[2];
  double normal[3];
  int32_t normal_len;
} lbm_plane_block;

/* Mirrors struct SimulationParameters (SimulationParameters.h:49-89). */
typedef struct {
  char stl_file[512];
  char flow_type[32];
  double dt, rho, nu_blood, tau, dx, pressure, reynolds_number, cte_vel;
  double systole, diastole;
  int32_t num_out, num_cycle, max_iter, iter_to_write;
  int32_t n_inlets, n_outlets, n_equations;
  lbm_plane_block inlets[LBM_MAX_BLOCKS];
  lbm_plane_block outlets[LBM_MAX_BLOCKS];
} lbm_config;

/* Replaces parseInputFile (LBM.h:81-188). LBM_ERR_IO if the file cannot be
 * opened; LBM_ERR_ARG on a numeric conversion error (the reference throws). */
int lbm_config_parse(const char* path, lbm_config* out);

/* ------------------------------------------------------------------------ */
/* Geometry and node tables (host)                                           */
/* ------------------------------------------------------------------------ */

typedef struct lbm_geometry lbm_geometry;

typedef struct {
  int32_t q;               /* 19 or 27 */
  int32_t periodic[3];     /* extension: periodic axes (reference: none) */
  int32_t nthreads;        /* host threads, 0 = all */
  int32_t verify_stride;   /* >0: re-check every n-th column against all triangles */
  int32_t compute_normals; /* near-wall normals (Subdomain.cpp:499-572): 0 none, 1 host, 2 GPU */
  int32_t device;          /* HIP device for compute_normals = 2, label_device = 1 and classify_device = 1 */
  int32_t label_device;    /* ray labels (ComplexGeometry.cpp:440-500): 0 host threads, 1 GPU (same labels) */
  int32_t classify_device; /* boundary points, node kinds, order and push table (ComplexGeometry.cpp:517-662,
                              Subdomain.cpp:364-471): 0 host threads, 1 GPU (same tables) */
} lbm_geometry_options;

void lbm_geometry_options_default(lbm_geometry_options* o);

/* Replaces the Geometry + SubDomain constructors (ComplexGeometry.cpp:48-73,
 * Subdomain.cpp:43-62) for one rank: STL -> grid -> ray labels -> boundary
 * sets -> nodes -> connectivity -> near-wall set and normals. The STL path in
 * the config is taken relative to the config file's directory. */
int lbm_geometry_from_config(const char* config_path, const lbm_geometry_options* opts,
                             lbm_geometry** out);

/* Same chain starting from a fluid label per grid point (uint8, index
 * (i*ny + j)*nz + k), skipping STL reading and ray casting. origin/spacing
 * give point coordinates origin + index*spacing (spacing is the lattice
 * pitch itself, not config grid_size). tris (12 doubles per triangle:
 * 3 vertices then file normal) is only used for near-wall normals and may be
 * NULL with ntris = 0. */
int lbm_geometry_from_labels(const uint8_t* fluid, int32_t nx, int32_t ny, int32_t nz,
                             const double origin[3], double spacing,
                             const lbm_plane_block* inlets, int32_t n_inlets,
                             const lbm_plane_block* outlets, int32_t n_outlets,
                             const double* tris, int64_t ntris,
                             const lbm_geometry_options* opts, lbm_geometry** out);

/* The node tables of a reference SubDomain, handed over as they stand
 * (Subdomain.h:66-118, Node.h:53-70), so that the reference's own set-up can
 * drive the GPU path without a second voxelisation. Every array is borrowed
 * for the call and copied. */
typedef struct {
  int32_t q;                     /* 19 or 27 (_nLatNodes) */
  int64_t n_nodes;               /* N = Involved_Lat_.size() */
  int32_t grid[3];               /* Geometry::pointsBox extents (nx+1, ny+1, nz+1) */
  double origin[3];              /* coordinates of grid index (0,0,0): the bounding-box minimum */
  double spacing;                /* Geometry::spacing, the lattice pitch (twice grid_size) */
  const int32_t* ijk;            /* 3N: Node::x_position, y_position, z_position, Involved_Lat_ order */
  const uint8_t* kind;           /* N: LBM_KIND_* (Node::NodeType_ and BoundaryType_) */
  const int32_t* outlet_id;      /* N: the OutletBoundaryNodes_ block holding the node, else -1 (NULL: no outlets) */
  const int32_t* nb;             /* Q*N: Node::m_neighbours[q] as a node index, nb[q*N + n], -1 for none */
  int32_t has_inlet_normal;      /* InletBoundaryNode_.second is set */
  double inlet_normal[3];        /* InletBoundaryNode_.second */
  int32_t n_outlets;             /* OutletBoundaryNodes_.size() */
  const double* outlet_normals;  /* 3 per outlet block: OutletBoundaryNodes_[k].second */
  int64_t n_nearwall;            /* nextToTheWall.size() */
  const int32_t* nearwall;       /* plain fluid node indices, in the order results are to be reported */
  const double* normals;         /* 3 per near-wall node: Node::normal */
} lbm_tables;

/* Replaces the Geometry + SubDomain constructors with tables the caller built
 * (the reference's own SubDomain, for one): checks them (sizes, kinds, outlet
 * blocks, node positions inside the grid and unique, every neighbour entry at
 * x + c_q, near-wall nodes plain fluid) and derives what the solver needs:
 * the creation-order node lists, and the node behind each inlet/outlet node
 * along its block normal (the +x,-x,+y,-y,+z,-z else-if of
 * Simulation::ApplyOutletBc, Simulation.cpp:650-741, read from nb). The
 * near-wall order given is the order lbm_get_wss and the output files use.
 * No triangles: lbm_geometry_get_triangles returns none. */
int lbm_geometry_from_tables(const lbm_tables* t, lbm_geometry** out);

typedef struct {
  int64_t n_nodes, n_fluid, n_wall, n_inlet, n_outlet, n_nearwall, n_boundary_points;
  int64_t n_triangles, upstream_out_of_box;
  int32_t grid[3];
  int32_t q;
  int32_t n_outlets;
  int32_t has_inlet_normal;
  double spacing;
  double bbox[6];          /* min_x, max_x, min_y, max_y, min_z, max_z */
  double inlet_normal[3];
} lbm_geometry_info;

int lbm_geometry_get_info(const lbm_geometry* g, lbm_geometry_info* out);
/* Grid labels (Point::Type Fluid/NotFluid after labelPoints). */
int lbm_geometry_get_labels(const lbm_geometry* g, uint8_t* fluid);
/* ijk: 3 per node; kind: LBM_KIND_*; outlet_id: outlet block or -1. Any may be NULL. */
int lbm_geometry_get_nodes(const lbm_geometry* g, int32_t* ijk, uint8_t* kind, int32_t* outlet_id);
/* Push neighbour table Node::m_neighbours (Subdomain.cpp:447-471): nb[q*N + n]
 * = node index at x + c_q, or -1. */
int lbm_geometry_get_neighbours(const lbm_geometry* g, int32_t* nb);
/* Near-wall nodes (ascending node index) and their normals (3 per node). */
int lbm_geometry_get_nearwall(const lbm_geometry* g, int32_t* nodes, double* normals);
/* Node behind each inlet/outlet node along its block normal, or -1. */
int lbm_geometry_get_bc_neighbour(const lbm_geometry* g, int32_t* bc_nbr);
/* Outlet block normals (3 per outlet). */
int lbm_geometry_get_outlet_normals(const lbm_geometry* g, double* normals);
/* Triangles as read from the STL (Geometry::globalTriangles,
 * GeometryReader.cpp:74-158): 12 doubles each, 3 vertices then the file
 * normal. n_triangles in lbm_geometry_info. */
int lbm_geometry_get_triangles(const lbm_geometry* g, double* tris);
void lbm_geometry_destroy(lbm_geometry* g);

/* ------------------------------------------------------------------------ */
/* Solver (device)                                                           */
/* ------------------------------------------------------------------------ */

typedef struct lbm_solver lbm_solver;

typedef struct {
  int32_t precision;       /* LBM_PREC_* */
  int32_t wall_mode;       /* LBM_WALL_* */
  int32_t inlet_mode;      /* LBM_INLET_* */
  int32_t outlet_mode;     /* LBM_OUTLET_* */
  int32_t flow_type;       /* LBM_FLOW_* */
  int32_t wss_every;       /* accumulate WSS every n-th step (reference: 1); 0 = never */
  int32_t parity_checks;   /* FP64 reference mode: detect the f*==0 streaming skip */
  int32_t device;          /* HIP device ordinal */
  double tau;              /* relaxation time (Simulation::SetRelaxation) */
  double dt;               /* time step (physical seconds) */
  double cte_vel;          /* inlet speed, lattice units */
  double spacing;          /* lattice pitch used in the WSS constant (0 = from geometry) */
  double outlet_rho;       /* LBM_OUTLET_NEE_PRESSURE density */
  double force[3];         /* extension: constant body-force density (Guo forcing) */
} lbm_params;

/* Reference defaults: FP64, reference walls/inlet/outlet, constant flow,
 * WSS every step, parity checks on, device 0, outlet_rho 1. */
void lbm_params_default(lbm_params* p);

/* Replaces the state set-up of the Simulation constructor
 * (Simulation.cpp:44-58) and SubDomain::Initialize (Subdomain.cpp:578-594):
 * populations f = w_q, accumulators zero, and the first collision done. */
int lbm_create(const lbm_geometry* g, const lbm_params* p, lbm_solver** out);

/* Per-inlet-node velocity profile: u_in(node, t) = sum_m coef_m(t) * basis[m][i]
 * for inlet node i (creation order). Default: one mode, basis 1, coef = the
 * lbm_step inlet speed (the reference's uniform plug). n_modes <= 16. */
int lbm_set_inlet_profile(lbm_solver* s, int32_t n_modes, const double* basis);

/* Extension: Brinkman damping field (>= 0) per node; NULL clears it. Before
 * the first step the initial collision is redone with the new field. */
int lbm_set_porosity(lbm_solver* s, const float* alpha);

/* Advance nsteps reference iterations (Simulation::Iterate loop body,
 * Simulation.cpp:553-568: Collision, Communicate, Streaming, inlet speed,
 * ApplyZouHeBc, ApplyOutletBc, ApplyWallBc, ComputeWallShearStress), each one
 * fused into a single kernel launch. The inlet speed follows params.flow_type
 * with the reference's accumulated time (t += dt per step).
 * In FP64 reference mode with parity_checks, returns LBM_ERR_PARITY if the
 * f*==0.0 streaming skip (Simulation.cpp:245-248) hit a node whose previous
 * populations are not kept (a plain fluid node), which would break parity. */
int lbm_step(lbm_solver* s, int32_t nsteps);

/* Same, with explicit profile coefficients per step: coef[t * n_modes + m]. */
int lbm_step_coefficients(lbm_solver* s, int32_t nsteps, const double* coef);

/* Post-BC populations f(t) (Node::m_distributions after the last iteration),
 * FP64, node-major, reference node order. */
int lbm_get_populations(lbm_solver* s, double* f);
/* Post-collision populations of the next iteration (m_newDistributions of the
 * collision that follows), FP64, same layout. */
int lbm_get_post_collision(lbm_solver* s, double* fs);
/* Overwrite f(t) of every node (Node::m_distributions) and redo the collision.
 * f = NULL sets the rest state f = w_q that lbm_create starts from (no host
 * copy of the state is kept, so a following adjoint pass starts from w_q as
 * after lbm_create). */
int lbm_set_populations(lbm_solver* s, const double* f);

/* Near-wall WSS in lbm_geometry_get_nearwall order: instantaneous |tau|
 * (Node::wall_shear_stress), accumulated tau (3 per node) and accumulated
 * |tau| (Simulation.cpp:767-799). Any pointer may be NULL. */
int lbm_get_wss(lbm_solver* s, double* wss, double* acc_tau, double* acc_mag);

/* TAWSS and OSI from the accumulators (Simulation::ComputeOSIAndTAWSS,
 * Simulation.cpp:804-826). */
int lbm_get_osi_tawss(lbm_solver* s, double* osi, double* tawss);

/* Density and velocity of f(t) per node (the values Simulation::Collision
 * computes at the start of the next iteration; wall velocity is 0 in
 * reference wall mode). These are the raw moments rho = sum f, u = sum c f / rho,
 * as in the reference, which has no forcing. With the Guo force or a Brinkman
 * field (extension modes) the physical velocity also takes half the step's
 * force density, (sum c f + F/2) / rho; neither this call nor the VTU of
 * lbm_write_output adds it. */
int lbm_get_macroscopic(lbm_solver* s, double* rho, double* u);

/* Density and velocity (sum f, sum c f / sum f, FP64 sums) of the
 * post-collision populations f* that the next step streams (the state
 * lbm_get_post_collision returns), at `count` given nodes only: a probe of a
 * state too large to read back whole. Unforced BGK leaves a node's density
 * and momentum as its collision found them (to rounding), so away from
 * reference-rule walls these are the density and velocity of f(t). */
int lbm_get_post_collision_moments(lbm_solver* s, const int32_t* nodes, int64_t count, double* rho, double* u);

/* Simulation::WriteOutput (Simulation.cpp:835-856) with `dir` standing for
 * the working directory: dir/Fluid_Variables/<iter>.vtu
 * (OutputFile::WriteVTK_Collective, OutputFile.cpp:60-231: nodes in z-y-x
 * order at their grid indices, Velocity / Density / VelocityMagnitude that
 * the last iteration's collision set, i.e. of the populations before the
 * last lbm_step, as WriteOutput(iter) sees them after iteration iter;
 * hexahedra over all-node cubes), dir/<iter>_FlowHemodynamics.csv
 * (WriteSurfaceAttribute, :234-316: X,Y,Z,WSS[,OSI,TAWSS],VelMag per
 * near-wall node; VelMag is 0 on D3Q27, where the reference never sets it)
 * and its copy smoothed over a 0.25 radius
 * (processFlowData / smoothData, :369-483) in
 * dir/FlowHemodynamics/<phase_name>_<iter>_FlowHemodynamics.csv. OSI and
 * TAWSS columns when `pulsatile`. The VTU is XML with appended raw binary
 * data (the reference's VTK writer picks its own encoding). LBM_ERR_ARG if
 * no step ran since lbm_create or lbm_set_populations. */
int lbm_write_output(const lbm_geometry* g, lbm_solver* s, int32_t iter, const char* phase_name, const char* dir,
                     int32_t pulsatile);
/* The two host-side pieces of lbm_write_output on their own: the VTU of given
 * node fields (rho: N, u: 3 per node, reference node order) and the
 * smoothing of a surface CSV (radius 0.25 in the reference). */
int lbm_write_vtu(const lbm_geometry* g, const double* rho, const double* u, const char* path);
int lbm_smooth_surface_csv(const char* in_path, const char* out_path, double radius, int32_t pulsatile);

typedef struct {
  double mass;             /* sum over nodes of rho, FP64 accumulation */
  double momentum[3];      /* sum of rho*u */
  double wss_l2;           /* sqrt(sum of instantaneous |tau|^2) over near-wall nodes */
  double wss_max;
  int64_t iteration;
  double time;             /* reference total_runtime */
} lbm_stats;

/* Replaces Boundary::ApplyBc(double& vel, std::vector<double>& normal)
 * (Boundary.h:61, Boundary.cpp:111-289) for `count` independent D3Q19
 * population vectors f[i*19 + q] (FP64, in place) with inlet speeds vel[i]
 * and one normal (InletBoundaryNode_.second), evaluated on HIP device
 * `device` by the device code the fused step kernel runs for the reference
 * inlet rule. */
int lbm_zouhe(int32_t device, int64_t count, double* f, const double* vel, const double normal[3]);

/* Wavefront-reduction kernel over the current post-collision state. */
int lbm_reduce(lbm_solver* s, lbm_stats* out);

/* lbm_reduce without device memory of the solver's own: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc)
 * and do not want the reduction tied to a workspace sized at lbm_create. The
 * partial sums go to one static device array of 64 MiB per GPU (held from
 * the library's load on that GPU until the process ends); the call allocates
 * nothing and gives the same values as lbm_reduce, bit for bit. A solver
 * whose reduction needs more than the array (about 4 * N / 256 + 2 * N_nw / 256
 * doubles, i.e. beyond roughly 500M nodes) gets LBM_ERR_ARG. Every call on
 * one GPU uses the same array and it keeps the last call's contents: calls
 * that overlap on one GPU (from two host threads) race and must be
 * serialised by the caller. */
int lbm_reduce_noalloc(lbm_solver* s, lbm_stats* out);

/* Run nsteps and report the device time of the step kernels, measured with
 * HIP events on the solver's stream (milliseconds, all steps). */
int lbm_step_timed(lbm_solver* s, int32_t nsteps, double* ms);

/* Device time (HIP events on the solver's stream) of the most recent
 * lbm_step / lbm_step_coefficients / lbm_step_timed batch, and its step
 * count. Waits for that batch to finish. */
int lbm_last_batch_time(lbm_solver* s, double* ms, int32_t* nsteps);
/* Split timing (no reference counterpart; measurement only): with it on,
 * every later batch also records an event pair around each of its first 16
 * launches that fold the stress ring into the FP64 accumulators, and
 * lbm_last_batch_kernel_time returns the time of all those launches (their
 * timed mean times their count, other_ms) and the rest of the batch's time
 * (the step kernel's share, step_ms). */
int lbm_set_kernel_timing(lbm_solver* s, int32_t enable);
/* Column family: take the 64-bit element offsets that states of 4 GiB or
 * more need on a state of any size (no reference counterpart; it lets a
 * small case check that path against the 32-bit one, bit for bit). */
int lbm_set_wide_offsets(lbm_solver* s, int32_t enable);
int lbm_last_batch_kernel_time(lbm_solver* s, double* step_ms, double* other_ms);

int lbm_synchronize(lbm_solver* s);
void lbm_destroy(lbm_solver* s);

typedef struct {
  /* n_special: nodes on the general path (reference wall rule: special
   * slots; half-way bounce-back: nodes with an explicit pull row);
   * n_bulk: the others. */
  int64_t n_nodes, n_special, n_bulk, n_pad;
  int64_t device_bytes;
  int32_t q, precision;
  double bytes_per_update;  /* algorithmic: 2 * Q * sizeof(population) */
} lbm_solver_info;
int lbm_get_solver_info(const lbm_solver* s, lbm_solver_info* out);

/* Plain device copy of `bytes` (read once, written once) repeated `iters`
 * times: the achievable HBM bandwidth reference (GB/s, read + write bytes)
 * and a launch of known size for calibrating traffic counters
 * (kernel copy_probe_kernel). No reference counterpart. */
int lbm_copy_probe(int32_t device, int64_t bytes, int32_t iters, double* gbps);

/* ------------------------------------------------------------- adjoint
 * No counterpart in the reference (it has no differentiation). Discrete
 * adjoint of the half-way bounce-back step (LBM_WALL_HALFWAY_BB) with NEE or
 * no inlet/outlet rule, for the objective
 *   J = (1/T) sum_{t=1..T} [ wss_weight * sum_{x in segment} |tau_t(x)|^2
 *                            + dp_weight * (rho_in(t) - rho_out(t)) / 3 ]
 * where tau is the wall shear stress of Simulation::ComputeWallShearStress
 * (Simulation.cpp:767-799) and rho_in / rho_out are the mean densities of the
 * inlet / outlet nodes after their boundary rule. */
typedef struct {
  double wss_weight;
  double dp_weight;
  const uint8_t* wss_mask; /* per near-wall node (lbm_geometry_get_nearwall order), 1 = in the segment; NULL = all */
} lbm_objective;

/* Forward pass of nsteps from the state set by lbm_create / lbm_set_populations,
 * with inlet coefficients coef[t * n_modes + m] (NULL: the lbm_step inlet
 * schedule). Stores the state every checkpoint_every steps and returns J. */
int lbm_adjoint_forward(lbm_solver* s, int32_t nsteps, const double* coef, const lbm_objective* obj,
                        int32_t checkpoint_every, double* J);
/* Reverse pass of the last lbm_adjoint_forward: dJ/dcoef[t * n_modes + m]
 * and dJ/dalpha per node (either may be NULL). Inside each checkpoint
 * interval the states are recomputed on a binomial (revolve) schedule with
 * `slots` stored states (<= 0: as many as fit, at most checkpoint_every - 1).
 * The solver state afterwards is the forward pass's final state. The record
 * stays valid for further backward calls until lbm_step*, lbm_set_populations,
 * lbm_set_inlet_profile or lbm_set_porosity runs; after any of them (or with
 * no forward pass) this returns LBM_ERR_STATE. */
int lbm_adjoint_backward(lbm_solver* s, int32_t slots, double* grad_coef, double* grad_porosity);
/* Moments tape of the reverse pass (default on): while recomputing a
 * checkpoint interval the step kernel stores the density and momentum sums
 * of every node's pulled populations, and the reverse step reads those 4
 * values instead of re-pulling 19 from the recomputed state (K * 4 values
 * per node of device memory, taken only where it fits in a quarter of the
 * free memory). The gradient is the same bit for bit either way. */
int lbm_adjoint_set_tape(lbm_solver* s, int32_t enable);
/* The binomial schedule used inside one interval of nsteps steps with
 * `slots` buffers: 5 int32 per action (op 0 advance / 1 reverse, from, to,
 * buffer from, buffer to; buffer -1 = the interval's stored start state,
 * -2 = scratch). actions may be NULL to query the count. */
int lbm_checkpoint_schedule(int32_t nsteps, int32_t slots, int32_t* actions, int64_t capacity, int64_t* count);

const char* lbm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LBM_C_API_H */
