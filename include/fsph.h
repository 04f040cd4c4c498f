/*
 * fsph.h -- C ABI of the MI355X (gfx950) per-step SPH particle pipeline.
 *
 * One fsph_ctx holds one MPI rank's share of particles (the state that
 * flecsi::topology::tree_topology::entities_ holds in the reference) as
 * device-resident structure-of-arrays, plus the octree, the SPH groups and
 * the neighbour tiles built from it.  Every entry point below replaces one
 * call of the reference's body_system / physics:: surface; the citation in
 * each comment names the reference function (file:line) whose effect the
 * call reproduces.
 *
 * Conventions
 *   - Every function returns FSPH_OK (0) or a negative FSPH_E_* code; the
 *     message of the last failure is fsph_last_error(ctx).  No C++ exception
 *     crosses this boundary.
 *   - The caller owns every host buffer; the context owns all device memory.
 *   - All work of a context is ordered on the context's own HIP stream and
 *     every call returns after that work has completed.
 *   - A context is not thread-safe: one host thread per context, as the
 *     reference runs one thread per MPI rank.
 *   - Particle arrays are in the context's current order.  After
 *     fsph_update_iteration that order is ascending (Morton key, id), the
 *     order psort leaves bodies in (mpisph/psort.h:247-253).  Periodic WALL
 *     copies (include/physics/boundary.h:121-250) follow the real particles
 *     in that sorted order and are told apart by id >= n_real.
 */
#ifndef FSPH_H
#define FSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define FSPH_OK 0
#define FSPH_E_ARG (-1)     /* bad argument or call order                   */
#define FSPH_E_HIP (-2)     /* HIP runtime error                            */
#define FSPH_E_OOM (-3)     /* device allocation failed                     */
#define FSPH_E_DUPKEY (-4)  /* two particles share a Morton key; the
                               reference's build_tree has undefined
                               behaviour here (tree_topology.h:957-971)     */
#define FSPH_E_NAN (-5)     /* NaN in state (default_physics.h:919-945)     */
#define FSPH_E_NEG (-6)     /* non-positive density/energy/h
                               (default_physics.h:950-979)                  */
#define FSPH_E_CAP (-7)     /* an internal capacity was exceeded            */
#define FSPH_E_NODEV (-8)   /* no HIP device / library built without one    */

/* ---- parameter keywords (include/params.h:128-152) --------------------- */
#define FSPH_KERNEL_CUBIC_SPLINE 0 /* kernels.h:106-153                    */
#define FSPH_KERNEL_WENDLAND_C6 1  /* kernels.h:502-566                    */

#define FSPH_EOS_IDEAL 0           /* eos.h:148-209                        */
#define FSPH_EOS_POLYTROPIC 1      /* eos.h:57-145                         */

#define FSPH_PREC_FP32 0 /* pair sums in FP32 with compensated summation    */
#define FSPH_PREC_FP64 1 /* pair sums in FP64 (the reference's arithmetic)  */

#define FSPH_GRAV_NONE 0
#define FSPH_GRAV_FMM 1  /* body_system::gravitation_fmm, bodies_system.h:282 */

/* Parameters of one context.  Field names follow the reference's param::
   names (include/params.h) where one exists. */
typedef struct fsph_params {
  int32_t kernel;      /* sph_kernel                                        */
  int32_t eos;         /* eos_type                                          */
  int32_t precision;   /* FSPH_PREC_*                                       */
  int32_t adaptive_dt; /* adaptive_timestep                                 */
  int32_t gravity;     /* enable_fmm                                        */
  int32_t group_max;   /* tree_topology::sub_entities_ (128)                */
  int32_t leaf_size;   /* FMM leaf bucket size                              */
  int32_t periodic[3]; /* periodic_boundary_x/y/z                           */
  int32_t n_target;    /* target neighbour count (smoothing-length solvers) */
  int32_t fmm_order;   /* 1: monopole sources, first-order local field (the
                          reference's scheme, fmm.h:37-159, with leaf_size
                          1); 2: quadrupole sources, second-order local
                          expansion                                         */
  double gamma;        /* poly_gamma                                        */
  double poly_K;       /* adiabatic constant of the polytropic EOS          */
  double eta;          /* sph_eta                                           */
  double visc_alpha;   /* sph_viscosity_alpha                               */
  double visc_beta;    /* sph_viscosity_beta                                */
  double visc_eps;     /* sph_viscosity_epsilon                             */
  double cfl;          /* timestep_cfl_factor                               */
  double initial_dt;   /* initial_dt                                        */
  double G;            /* gravitational_constant                            */
  double theta;        /* fmm_macangle                                      */
  double box[3];       /* box_length, box_width, box_height (periodic box
                          centred on the origin, boundary.h:134-144)        */
  /* smoothing-length update at the end of a step (fsph_update_h) */
  int32_t h_mode;      /* FSPH_H_CLOSED | FSPH_H_NEWTON | FSPH_H_LEARNED    */
  int32_t newton_maxit;/* Newton-h iteration cap (20)                       */
  double newton_tol;   /* stop when |dh|/h < newton_tol (1e-4)              */
  double h_search;     /* gather radius of the Newton solve, in units of
                          the current h (1.3); the solution is clamped to
                          it and counted as unconverged                     */
  int32_t learn_steps; /* FSPH_H_LEARNED: steps of plain Newton-h whose
                          histories train the predictor online (50)         */
  int32_t n_tol;       /* FSPH_H_LEARNED: a proposal whose kernel-weighted
                          neighbour count is off the target by more than
                          this falls back to plain Newton-h (5)             */
  int32_t thermokinetic; /* thermokinetic_formulation (params.h:590-592):
                          1 evolves the specific thermokinetic energy
                          e = u + v^2/2 with compute_dedt and leapfrog_kick_e
                          and recovers u from it (default_physics.h:127-
                          145, 503-566; integration.h:95-98); 0 evolves u
                          with compute_dudt                                 */
  int32_t viscosity;   /* sph_viscosity: FSPH_VISC_CONSTANT or
                          FSPH_VISC_CULLEN (viscosity.h:128-155)            */
  double visc_alpha_max; /* sph_viscosity_alpha_max (Cullen-Dehnen)         */
  double visc_l;       /* sph_viscosity_l (Cullen-Dehnen decay length)      */
  double visc_delta;   /* sph_viscosity_delta (Cullen-Dehnen)               */
  int32_t evolve_internal_energy; /* evolve_internal_energy (params.h:583-
                          585): 0 leaves out the energy pass and the energy
                          kicks (hydro/main_driver.cc:145,175,214) and the
                          checks of u, as the reference's polytropic-star
                          configurations do (data/mesa_nx20.par:52)        */
} fsph_params;

#define FSPH_VISC_CONSTANT 0 /* Monaghan alpha, beta (viscosity.h:128-137) */
#define FSPH_VISC_CULLEN 1   /* Cullen & Dehnen 2010 (viscosity.h:139-363) */

#define FSPH_H_CLOSED 0  /* h = eta kw (m/rho)^(1/3), default_physics.h:826 */
#define FSPH_H_NEWTON 1  /* Newton on rho~(h) h^3 = m (eta kw)^3            */
#define FSPH_H_LEARNED 2 /* MLP proposal, Newton fallback                   */

/* Fills p with the reference's defaults (params.h DECLARE_PARAM values,
   kernel = cubic spline, eos = ideal, gamma 1.4, group_max 128). */
void fsph_default_params(fsph_params *p);

/* Host view of particle input, one array per field (SoA).  h is the
   smoothing length in the reference's sense: the kernel support radius
   (body_u::radius, body.h).  rho/P/c may be NULL (then zero). */
typedef struct fsph_particles {
  int64_t n;
  const double *x, *y, *z;
  const double *vx, *vy, *vz;
  const double *u;
  const double *h;
  const double *m;
  const int64_t *id;
  const double *rho, *P, *c;
} fsph_particles;

/* Field selectors for fsph_download (each field is n doubles, except
   FSPH_F_ID (int64) and FSPH_F_KEY (uint64)). */
enum fsph_field {
  FSPH_F_X = 0, FSPH_F_Y, FSPH_F_Z,
  FSPH_F_VX, FSPH_F_VY, FSPH_F_VZ,
  FSPH_F_VHX, FSPH_F_VHY, FSPH_F_VHZ,
  FSPH_F_U, FSPH_F_H, FSPH_F_M,
  FSPH_F_RHO, FSPH_F_P, FSPH_F_C, FSPH_F_VSIG,
  FSPH_F_AX, FSPH_F_AY, FSPH_F_AZ,
  FSPH_F_DUDT,
  FSPH_F_GX, FSPH_F_GY, FSPH_F_GZ, FSPH_F_PHI,
  FSPH_F_ID, FSPH_F_KEY, FSPH_F_NNGB,
  FSPH_F_HNGB,   /* #{b : r_ab < h_a} after the last Newton/learned update */
  FSPH_F_HITER,  /* Newton iterations of the last update (0: accepted
                    learned proposal)                                      */
  FSPH_F_E,      /* thermokinetic energy e (body_u::getTotalenergy)        */
  FSPH_F_DEDT,   /* de/dt (body_u::getDedt)                                */
  FSPH_F_ALPHA,  /* viscosity alpha of the particle (body_u::getAlpha)     */
  FSPH_F_DIVV,   /* velocity divergence (body_u::getDivergenceV)           */
  FSPH_NFIELDS
};

typedef struct fsph_ctx fsph_ctx;

/* ---- lifetime ---------------------------------------------------------- */
/* Creates a context on HIP device `device` (replaces constructing
   body_system<double,3>, mpisph/bodies_system.h:38-90). */
int fsph_create(const fsph_params *p, int device, fsph_ctx **out);
void fsph_destroy(fsph_ctx *ctx);
const char *fsph_last_error(const fsph_ctx *ctx);

/* Copies host particles to the device (replaces body_system::read_bodies,
   bodies_system.h:92-98, minus the HDF5 reader).  Resets time and dt. */
int fsph_upload(fsph_ctx *ctx, const fsph_particles *parts);

/* body_system::write_bodies (bodies_system.h:100-102 -> io.h:873-1081)
   and read_bodies (bodies_system.h:92-98 -> io.h:632-871) in the H5part
   step layout of doc/H5_STRUCTURE.md, stored without HDF5: <prefix>.h5part
   is a directory holding header.json ({"format": "fsph-h5raw-1",
   "dimension": 3}) and one Step#<k> directory per output with
   attributes.json ({"time", "iteration", "timestep", "nparticles"}) and
   one raw little-endian array per dataset, named as the reference names
   them: x y z vx vy vz ax ay az h rho u P m alpha divergenceV dDivVdt e
   (.f8, double), id (.i8, int64), type (.i4, int32: 0 NORMAL, 1 WALL).
   write_bodies appends a step holding every particle (WALL copies
   included, as the reference writes tree entities) with the current
   time, iteration and dt; with separate_iterations each output goes to
   <prefix>_<%05d output>.h5part as Step#0.  read_bodies uploads the
   particles of Step#0 (start_iteration 0) or of the step whose iteration
   attribute is start_iteration (in <prefix>.h5part, else in
   <prefix>_<%05d start_iteration>.h5part), reading x y z vx vy vz m h id
   and, when present, u, rho and P; for start_iteration != 0 the context's
   dt, time and iteration become the step's, and the next fsph_step is the
   driver's initial iteration at that iteration (main_driver.cc:101). */
int fsph_write_bodies(fsph_ctx *ctx, const char *prefix,
                      int separate_iterations);
int fsph_read_bodies(fsph_ctx *ctx, const char *prefix,
                     int64_t start_iteration);

/* Number of particles currently held: *n_total includes periodic WALL
   copies, *n_real does not (body_system::getNLocalBodies). */
int fsph_count(const fsph_ctx *ctx, int64_t *n_total, int64_t *n_real);

/* Copies one field of all n_total particles to host memory dst. */
int fsph_download(fsph_ctx *ctx, int field, void *dst);

/* Overwrites one field of all n_total particles from host doubles src (in
   the context's current order), the write-back half of running a functor
   on the host: the reference's functors read and write body_u fields
   directly (tree_topology.h:406).  Ids, keys and neighbour counts cannot
   be set; setting x, y, z or h invalidates the tree until the next
   fsph_update_iteration.  Derived fields are rounded to the context's
   precision. */
int fsph_set_field(fsph_ctx *ctx, int field, const double *src);

/* ---- the per-step pipeline (app/drivers/hydro/main_driver.cc:98-279) -- */

/* body_system::update_iteration (bodies_system.h:152-229): removes and
   regenerates periodic WALL copies (boundary.h:46-250), computes the range
   of x+-h (bodies_system.h:231-268), the Morton keys (tree_topology.h:
   828-832, filling_curve.h:503-525), sorts by (key,id) (psort.h:247),
   builds the octree with its centre-of-mass data (tree_topology.h:879-997,
   default_physics.h:60-119), selects the <=group_max groups
   (tree_topology.h:222-235) and builds the neighbour tiles that
   traversal_sph would visit (tree_topology.h:213-442).  Positions and h
   do not change between this call and the end of the step's SPH passes,
   so the tiles serve all of them (the reference rebuilds the tree in
   reset_ghosts, bodies_system.h:273-275, with identical result). */
int fsph_update_iteration(fsph_ctx *ctx);

/* apply_in_smoothinglength(physics::compute_density_pressure_soundspeed)
   (default_physics.h:310-320): density, EOS pressure and sound speed,
   signal speed. */
int fsph_density_eos(fsph_ctx *ctx);

/* eos::init over all particles (eos.h:155-159): for the ideal EOS nothing
   that the pipeline reads; kept for call-order fidelity. */
int fsph_eos_init(fsph_ctx *ctx);

/* apply_in_smoothinglength(physics::compute_acceleration)
   (default_physics.h:349-408): hydro acceleration; zeroes g and phi. */
int fsph_acceleration(fsph_ctx *ctx);

/* body_system::gravitation_fmm (bodies_system.h:282-288 ->
   tree_topology.h:449-795 with the callbacks of fmm.h:37-159): FMM
   self-gravity of the current tree into g and phi, with opening angle
   theta, bucket size leaf_size and multipole order fmm_order.  Call it
   after fsph_acceleration, which zeroes g and phi as compute_acceleration
   does (default_physics.h:406-407). */
int fsph_gravity(fsph_ctx *ctx);

/* Sizes of the last fsph_gravity (introspection for measurement; no
   reference counterpart): out[0] FMM cells (internal nodes with more than
   leaf_size particles, and buckets), out[1] buckets, out[2] M2L cell
   pairs, out[3] P2P bucket pairs, out[4] particle-particle interactions of
   those pairs (self pairs excluded), out[5] levels of the pair recursion
   or depths of the sink-major walk, out[6] the list builder that ran
   (FSPH_FMM_WALK or FSPH_FMM_PAIRS), out[7] list-pool chunks the walk
   used.  All zero before the first gravity evaluation. */
int fsph_gravity_stats(fsph_ctx *ctx, int64_t out[8]);

/* FMM interaction-list builders (fsph_set_fmm_builder).  Both build the
   pair set of the dual-tree recursion (tree_topology.h:449-795):
   FSPH_FMM_WALK walks it sink by sink, depth by depth, and sums M2L and
   L2L in the walk; FSPH_FMM_PAIRS runs it level by level over (sink,
   source) pairs and buckets the lists by sink. */
#define FSPH_FMM_WALK 0
#define FSPH_FMM_PAIRS 1

/* List builder of the next fsph_gravity calls (no reference counterpart;
   testing and measurement), FSPH_FMM_WALK by default.  stack_limit in
   [1, 1024] caps the walk's per-sink LDS stack of cells still to split
   (0: 1024); an evaluation in which a sink would pass it is redone by the
   pair recursion (gravity_stats out[6] says which ran). */
int fsph_set_fmm_builder(fsph_ctx *ctx, int builder, int stack_limit);

/* Capacities (entries) that the next fsph_gravity starts from (no
   reference counterpart): cap_q sizes the walk's list pool (in 64-entry
   chunks) and the pair recursion's queue, cap_m2l and cap_p2p the pair
   recursion's M2L and P2P lists.  They grow whenever an evaluation
   overflows them, so any positive value gives the same result; small
   values make the overflow-and-regrow path run, for testing. */
int fsph_set_fmm_capacity(fsph_ctx *ctx, int64_t cap_q, int64_t cap_m2l,
                          int64_t cap_p2p);

/* apply_in_smoothinglength(physics::compute_dudt) (default_physics.h:
   422-487) followed by apply_all(recompute_pressure_soundspeed)
   (default_physics.h:154-161) with the context's current dt. */
int fsph_dudt(fsph_ctx *ctx);

/* apply_in_smoothinglength(physics::compute_dedt) (default_physics.h:
   503-566) followed by apply_all(recompute_pressure_soundspeed_
   thermokinetic) (default_physics.h:170-180) with the context's current
   dt: de/dt, u recovered from e and v, P and c at the half step.  Needs
   thermokinetic = 1. */
int fsph_dedt(fsph_ctx *ctx);

/* apply_in_smoothinglength(viscosity::compute_alpha) (viscosity.h:
   318-356): the Cullen-Dehnen switch sets each particle's alpha from the
   trigger -d(div v)/dt, the limiter xi and the signal speed (compute_divv,
   default_physics.h:274-300, runs inside fsph_density_eos with this
   viscosity).  Needs viscosity = FSPH_VISC_CULLEN. */
int fsph_compute_alpha(fsph_ctx *ctx);

/* apply_all(viscosity::initialize_alpha) (viscosity.h:164-171): alpha = 0
   for the Cullen-Dehnen viscosity, sph_viscosity_alpha otherwise. */
int fsph_initialize_alpha(fsph_ctx *ctx);

/* apply_all(physics::set_total_energy) (default_physics.h:328-335):
   e = v^2/2 + u (no external potential).  The driver calls it once, after
   eos::init (hydro/main_driver.cc:108-111); later density passes recover
   u from e (default_physics.h:313-314).  Needs thermokinetic = 1. */
int fsph_set_total_energy(fsph_ctx *ctx);

/* Leapfrog pieces (include/physics/integration.h:53-110), with the
   context's current dt.  `what` is a bitmask of FSPH_KICK_*.  A drift
   moves the particles away from the tree and neighbour tiles, so the
   calls that use them (density, forces, gravity, Newton-h and learned-h
   updates, neighbour lists) then need fsph_update_iteration first, as in
   hydro/main_driver.cc:187-193. */
#define FSPH_KICK_V 1u      /* leapfrog_kick_v   integration.h:66-70     */
#define FSPH_KICK_U 2u      /* leapfrog_kick_u   integration.h:81-84     */
#define FSPH_SAVE_VHALF 4u  /* save_velocityhalf integration.h:53-55     */
#define FSPH_DRIFT 8u       /* leapfrog_drift    integration.h:107-110   */
#define FSPH_KICK_E 16u     /* leapfrog_kick_e   integration.h:95-98     */
int fsph_integrate(fsph_ctx *ctx, uint32_t what);

/* get_all(physics::compute_smoothinglength) (default_physics.h:826-848). */
int fsph_update_h(fsph_ctx *ctx);

/* apply_all(physics::compute_dt) + get_all(set_adaptive_timestep)
   (default_physics.h:603-695, 762-823): sets the context's dt and
   returns it in *dt (may be NULL). */
int fsph_timestep(fsph_ctx *ctx, double *dt);

/* physics::advance_time (default_physics.h:855-862). */
int fsph_advance_time(fsph_ctx *ctx);

/* One whole iteration of the hydro driver loop body
   (app/drivers/hydro/main_driver.cc:98-277; with gravity, the newtonian
   driver app/drivers/newtonian/main_driver.cc:103-296), including
   advance_time.  The first call after fsph_upload runs the initial
   iteration (no integration). */
int fsph_step(fsph_ctx *ctx);

/* apply_all(physics::check_nans) and apply_all(physics::check_negativity)
   (default_physics.h:918-979) over every particle: FSPH_E_NAN if a field
   the reference checks holds a NaN, FSPH_E_NEG for a negative density,
   pressure or internal energy (the message names the particle's id).
   fsph_step runs the same check after the smoothing-length update, as the
   driver does (hydro/main_driver.cc:263-265), and returns its code. */
int fsph_check_state(fsph_ctx *ctx);
/* Turns fsph_step's state check off or on (default on). */
int fsph_set_checks(fsph_ctx *ctx, int on);

/* Current time-step state: iteration, dt, totaltime
   (physics::iteration/dt/totaltime, default_physics.h:31-40). */
int fsph_time(const fsph_ctx *ctx, int64_t *iteration, double *dt,
              double *totaltime);
/* Overrides dt (e.g. a fixed-step run). */
int fsph_set_dt(fsph_ctx *ctx, double dt);

/* Global FP64 reductions over real particles (analysis.h:57-219):
   out[0] mass, out[1..3] linear momentum, out[4] kinetic energy,
   out[5] internal energy, out[6] gravitational energy (1/2 sum m phi),
   out[7..9] angular momentum. */
int fsph_reductions(fsph_ctx *ctx, double out[10]);

/* fsph_reductions that allocates nothing.  No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).
   The partial sums live in a static device array of the library, one copy
   per device, set up by fsph_create: at most 512 partial rows of 10
   doubles (41,040 bytes), which serves any particle count (the reduction
   never launches more than 512 workgroups); a call that would need more
   returns FSPH_E_CAP.  The result equals fsph_reductions bit for bit.
   Every call on one device shares that array: calls that overlap on one
   GPU (contexts of one device driven from several host threads) race. */
int fsph_reductions_noalloc(fsph_ctx *ctx, double out[10]);

/* ---- parity/introspection ---------------------------------------------- */

/* Range of the last update_iteration: lo[3], hi[3]. */
int fsph_get_range(fsph_ctx *ctx, double lo[3], double hi[3]);

/* Octree of the last update_iteration.  *n_nodes and *n_groups are
   always set; the arrays may be NULL.  node_key: the node's key with its
   leading 1 (filling_curve.h:65-67 convention, depth = (63-clz)/3);
   node_start/node_count: its particle range in sorted order;
   node_parent: parent node index (-1 for the root). */
int fsph_get_tree(fsph_ctx *ctx, int64_t *n_nodes, uint64_t *node_key,
                  int32_t *node_start, int32_t *node_count,
                  int32_t *node_parent, int64_t *n_groups,
                  int32_t *group_start, int32_t *group_count);

/* Neighbour lists that traversal_sph hands to the physics functors
   (tree_topology.h:372-379, 406), in sorted particle order: the ids of the
   neighbours of particle i are ids[off[i]..off[i+1]) in ascending sorted
   index.  off has n_total+1 entries; ids may be NULL to query the total
   (returned in off[n_total]).  cap is the length of ids. */
int fsph_neighbor_lists(fsph_ctx *ctx, int64_t *off, int64_t *ids,
                        int64_t cap);

/* ---- learned smoothing-length predictor (new; no reference counterpart:
   the reference updates h in closed form, default_physics.h:826-848) ---- */

/* The predictor is a 6 -> 64 -> 64 -> 1 tanh MLP (BF16 weights on MFMA,
   FP32 accumulation and master weights): y = w3 . tanh(W2^T tanh(W1^T x +
   b1) + b2) + b3.  In FSPH_H_LEARNED mode its inputs are the features
   [ln((4 pi / 3) rho h^3 / (m (N_target - 1))), (N_ngb - N_target) /
   N_target, dt div(v), dt vsig / h, ln(h / h_old) of the previous update,
   |a| dt^2 / h], normalised with the mean and standard deviation of the
   first update's sample, and its output is ln(h_new / h) times a scale
   fixed at the first update.  Parameters are a flat FP32 vector of
   fsph_mlp_num_params() entries: W1 [8][64] (input rows 6 and 7 unused and
   kept zero), b1 [64], W2 [64][64], b2 [64], w3 [64], b3.  Setting the
   parameters also resets the Adam moments. */
int fsph_mlp_num_params(void);
int fsph_mlp_get_params(fsph_ctx *ctx, float *p, int64_t n);
int fsph_mlp_set_params(fsph_ctx *ctx, const float *p, int64_t n);
/* y[B] = MLP(x[B][6]) for host arrays (features already normalised). */
int fsph_mlp_forward(fsph_ctx *ctx, const float *x, int64_t B, float *y);
/* Gradient of L = mean_b (MLP(x_b) - t_b)^2 with respect to the flat
   parameters (grad has fsph_mlp_num_params() entries) and L itself. */
int fsph_mlp_gradient(fsph_ctx *ctx, const float *x, const float *t,
                      int64_t B, float *grad, float *loss);
/* `iters` Adam steps (learning rate lr) on one host batch; final loss. */
int fsph_mlp_train(fsph_ctx *ctx, const float *x, const float *t, int64_t B,
                   int iters, float lr, float *loss);
/* Statistics of the last smoothing-length update: particles that fell
   back from the learned proposal to plain Newton-h, and particles whose
   root lay beyond the gather radius (clamped); training steps done and the
   last training loss. */
int fsph_h_stats(fsph_ctx *ctx, int64_t *fallback, int64_t *clamped,
                 int32_t *trained_steps, float *last_loss);

/* Gather set of the last smoothing-length update (Newton-h or learned-h):
   *kappa is the bracket radius of its first solve in units of the
   particle's h (h_search for Newton-h, the narrowed radius around the
   proposal for learned-h), *mask_words the neighbour-mask words of the
   gather tiles the update built (rows times 32-candidate words: the pair
   tests of those tiles).  Learned-h solves first on the step's SPH tiles,
   which it reuses (its bracket there stays within the current h), and
   builds gather tiles only for the tiles it could not settle there.  Both
   0 after a closed-form update. */
int fsph_h_gather(fsph_ctx *ctx, double *kappa, int64_t *mask_words);

/* Milliseconds spent in each phase of the last fsph_step, measured with
   HIP events on the context's stream around the phase's kernels (summed
   when a phase runs more than once in a step); see FSPH_T_* for the slots.
   Slots marked (1 kernel) time exactly one kernel launch. */
enum fsph_timer {
  FSPH_T_INTEGRATE = 0, /* leapfrog kernels                              */
  FSPH_T_BOUNDARY,      /* periodic WALL copies                          */
  FSPH_T_RANGE_KEYS,    /* range reduction + Morton keys                 */
  FSPH_T_SORT,          /* radix sort + permutation                      */
  FSPH_T_TREE,          /* octree, aggregates, groups                    */
  FSPH_T_TRAVERSE,      /* neighbour-tile traversal (1 kernel)           */
  FSPH_T_MASKS,         /* neighbour bitmasks (1 kernel)                 */
  FSPH_T_DENSITY,       /* density + EOS (1 kernel)                      */
  FSPH_T_VSIG,          /* signal speed (1 kernel)                       */
  FSPH_T_ACCEL,         /* hydro acceleration (1 kernel)                 */
  FSPH_T_GRAVITY,       /* self-gravity                                  */
  FSPH_T_DUDT,          /* du/dt + P, c at the half step (1 kernel)      */
  FSPH_T_H,             /* smoothing-length update                       */
  FSPH_T_DT,            /* time-step reduction                           */
  FSPH_T_GRAPH,         /* graph replay of a batched step's tail (pair
                           passes, kicks, h, dt minimum)                 */
  FSPH_T_GRAV_PAIRS,    /* FMM interaction lists: dual-tree pair levels
                           and their bucketing by sink (inside
                           FSPH_T_GRAVITY)                               */
  FSPH_T_M2L,           /* multipole-to-local (1 kernel, inside
                           FSPH_T_GRAVITY)                               */
  FSPH_T_P2P,           /* local-expansion evaluation and direct sums
                           (1 kernel, inside FSPH_T_GRAVITY)             */
  FSPH_NTIMERS
};
int fsph_phase_times(fsph_ctx *ctx, float ms[FSPH_NTIMERS]);
/* Turns the per-phase HIP-event timing on or off (default off). */
int fsph_enable_timing(fsph_ctx *ctx, int on);

/* ---- batched mode (new; no reference counterpart: the reference runs one
   simulation per MPI job) ----------------------------------------------

   One context advances n_inst (1..256) independent simulations that share
   one parameter set: the instances' particles are concatenated, and every
   kernel of a step runs once for all of them.  Each instance evolves
   exactly as it would in a context of its own (same range, keys, tree,
   neighbours, periodic copies and time step).  fsph_step advances every
   instance by one iteration of its own dt; fsph_update_iteration,
   fsph_density_eos, fsph_acceleration, fsph_dudt, fsph_integrate and
   fsph_update_h act on all instances.  Calls with one-simulation meaning
   (fsph_upload, fsph_timestep, fsph_time, fsph_set_dt, fsph_reductions,
   fsph_get_range, fsph_gravity) return FSPH_E_ARG on a batched context;
   fsph_count and fsph_download cover the concatenated arrays.  Self-gravity
   and FSPH_H_LEARNED are not available in batched mode. */
int fsph_create_batch(const fsph_params *p, int device, int n_inst,
                      fsph_ctx **out);
/* parts[i] is instance i (ids in [0, parts[i].n)); rho, P, c for all
   instances or none. */
int fsph_upload_batch(fsph_ctx *ctx, const fsph_particles *parts,
                      int n_inst);
int fsph_instance_count(const fsph_ctx *ctx, int inst, int64_t *n_total,
                        int64_t *n_real);
/* One field (as fsph_download) of instance inst's particles, in the
   instance's current (sorted) order. */
int fsph_instance_download(fsph_ctx *ctx, int inst, int field, void *dst);
/* Boundary of instance inst (new): periodic = 1 (the default) gives the
   context's periodic boundaries and box (periodic[], box[]); 0 gives open
   boundaries (no wrapping, no WALL copies), so one batch can hold periodic
   Sedov blasts and open Sod tubes.  Takes effect at the next
   fsph_update_iteration. */
int fsph_instance_set_periodic(fsph_ctx *ctx, int inst, int periodic);
int fsph_instance_time(const fsph_ctx *ctx, int inst, int64_t *iteration,
                       double *dt, double *totaltime);

/* Graph mode of a batched context (new): the sync-free tail of every step
   after the neighbour tiles (density, signal speed, acceleration, kick v,
   du/dt, kick u, closed-form h, per-instance dt minimum) runs as one
   replayed hipGraph of persistent kernels.  Needs h_mode FSPH_H_CLOSED,
   no gravity and the constant viscosity; results are those of the eager
   path. */
int fsph_enable_graph(fsph_ctx *ctx, int on);

/* Autotuner (new): times `steps` calls of fsph_step for each candidate
   tile width of the pair kernels (64 or 128 rows per workgroup: a group of
   more than 64 particles is processed as two tiles of its rows, each with
   the group's candidates that neighbour one of its rows) and candidate
   chunk (64, 128, 256 candidates staged in LDS per round), keeps the
   fastest and returns it.  Candidates are compared on the time of the
   phases the tuning changes (FSPH_T_MASKS, _DENSITY, _VSIG, _ACCEL,
   _DUDT, _H and _GRAPH, by HIP events), not of whole steps.  group_max,
   the reference's grouping
   (tree_topology::sub_entities_, tree_topology.h:222-235,2121), is not
   searched: the traversal's node test uses the whole group
   (tree_topology.h:323-355), so the neighbour lists depend on it.  The
   tile width and the chunk only cut and re-associate the sums: the
   neighbour lists are those of the configured group_max for every choice.
   Every candidate is timed from the same state: the particle fields,
   clocks and instance clocks are saved first (one device copy of every
   field) and restored before each candidate and at the end, so the call
   leaves the simulation where it was (after the initial iteration, which it
   runs if it has not run).  The learned predictor of FSPH_H_LEARNED keeps
   what it learns during the timed steps. */
int fsph_autotune(fsph_ctx *ctx, int steps, int32_t *tile_rows,
                  int32_t *chunk);

/* Sets the two tuning choices the autotuner searches (new): tile_rows (64
   or 128) and chunk (64, 128 or 256).  tile_rows takes effect at the next
   fsph_update_iteration (the current tiles keep the width they were built
   with); chunk at the next pair pass. */
int fsph_set_tuning(fsph_ctx *ctx, int32_t tile_rows, int32_t chunk);

#ifdef __cplusplus
}
#endif
#endif /* FSPH_H */
