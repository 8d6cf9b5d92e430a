/*
 * rh3t.h -- C ABI of the MI355X (gfx950) implementation of the semi-discrete
 * operator l() and the SSP-RK3 stage rk3_ssp() of the 2D three-temperature
 * radiation-hydrodynamics solver colinhu123/3T_RH.
 *
 * The reference is a single Rust binary crate with no plugin or FFI layer:
 * l() and rk3_ssp() are private functions (src/main.rs:135, src/main.rs:351)
 * over the Field / GhostGrid / BCType / Scratch data model.  Each entry point
 * below names the reference item it replaces.  All pointers are host pointers;
 * every device buffer is owned by an rh3t_ctx.  No call throws; each returns
 * an rh3t_status.  A problem or a context is not thread-safe: use one per
 * thread.  Field state crosses the boundary as the reference's AoS layout:
 * 6 doubles (rho, mom_x, mom_y, ee, ei, er) per grid point at l = i*ny + j
 * (src/field1.rs:83-85, src/field1.rs:204-215).
 *
 * A Rust crate binds this header 1:1 through an extern "C" block; the
 * Python mirror lives in the 3t_rh_amd package.
 */
#ifndef RH3T_H
#define RH3T_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RH3T_NFIELD 6      /* rho, mom_x, mom_y, ee, ei, er */
#define RH3T_NOFFSETS 17   /* radius-4 cross, src/ghost.rs:482-493 */
#define RH3T_HALO 4        /* stencil radius */

typedef enum {
  RH3T_OK = 0,
  RH3T_ERR_ARG = 1,          /* bad argument / size / null pointer            */
  RH3T_ERR_HIP = 2,          /* a HIP runtime call failed                     */
  RH3T_ERR_NONPHYSICAL = 3,  /* assert_admissible failed, src/main.rs:24-33   */
  RH3T_ERR_GHOST = 4,        /* non-finite ghost state, src/ghost.rs:229-263  */
  RH3T_ERR_GEOMETRY = 5,     /* ghost build failed (uncached stencil target,
                                no boundary side, substencil cannot fit)      */
  RH3T_ERR_UNSUPPORTED = 6,  /* feature outside this build (an unknown
                                TimeDependent profile; reverse mode of the
                                heavy and TimeDependent ghost kinds)          */
  RH3T_ERR_NO_DEVICE = 7,    /* no HIP device visible                         */
  RH3T_ERR_SINGULAR = 8,     /* singular 6x6 Wall-ILW system, src/bc1.rs:468  */
  RH3T_ERR_DT = 9,           /* non-positive dt, src/main.rs:464              */
  RH3T_ERR_SIZE = 10         /* input beyond a fixed-capacity entry point     */
} rh3t_status;

/* BCType, src/bc1.rs:16-28 (tagged union flattened to POD). */
typedef enum {
  RH3T_BC_REFLECTIVE_WALL = 0,
  RH3T_BC_WALL = 1,
  RH3T_BC_PERIODIC = 2,
  RH3T_BC_CONSTANT = 3,
  RH3T_BC_TIME_DEPENDENT = 4,
  RH3T_BC_OUTFLOW = 5,
  RH3T_BC_ZEROTH_ORDER = 6,
  RH3T_BC_FARFIELD = 7
} rh3t_bc_kind;

/* TimeDependent(f), src/bc1.rs:634-636: the reference's BC is a host
 * closure f(P0, n, t) -> State of the boundary point P0, its normal and the
 * stage time.  Here the closure is one of a registry of built-in profiles,
 * evaluated on the device at the stage time (t, t + dt, t + dt/2 in
 * rk3_ssp's three l() calls, src/main.rs:351-373):
 *   RH3T_TD_MOVING_FRONT  a planar front normal to x at front0 + speed * t:
 *                         state where P0.x <= front0 + speed * t, state2
 *                         elsewhere -- with state = post-shock, state2 =
 *                         pre-shock, speed = 10, front0 = 0 the moving
 *                         Mach-10 shock closure of init_double_mach
 *                         (src/init.rs:61-65). */
#define RH3T_TD_MOVING_FRONT 1

typedef struct {
  int32_t kind;        /* rh3t_bc_kind                                        */
  int32_t profile;     /* TimeDependent: RH3T_TD_* profile; 0 otherwise       */
  double state[6];     /* Constant(State) / FarField(State) / TimeDependent
                          state behind the front                             */
  double p_inf;        /* Outflow { p_inf, sigma, l_domain }                  */
  double sigma;
  double l_domain;
  double state2[6];    /* TimeDependent: state ahead of the front             */
  double speed;        /* TimeDependent: front speed along x                  */
  double front0;       /* TimeDependent: front position at t = 0              */
} rh3t_bc;

/* Physics constants, src/constant.rs:1-25, as a runtime struct.  The
 * DIFFUSION_ACTIVE / SOURCE_ACTIVE gates are derived exactly as the
 * reference derives them (any kappa != 0, any omega != 0). */
typedef struct {
  double gamma_e, gamma_i, gamma_r;
  double kappa_e, kappa_i, kappa_r;
  double cve, cvi, a_rad;
  double omega_ei, omega_er;
  double lambda;       /* CFL number LAMBDA                                   */
  double weno_q;       /* WENO_Q, exponent of the boundary WENO weights       */
  double eps;          /* DEFAULT_EPS                                         */
} rh3t_physics;

/* GridInfo, src/field1.rs:5-73 */
typedef struct {
  int32_t nx, ny;
  double dx, dy, x0, y0;
} rh3t_grid;

/* One cached ghost (GhostInfo, src/ghost.rs:18-28) plus the static source
 * index that the cheap BC branches resolve per call in the reference
 * (reflective_wall_value bc1.rs:930-1037, zeroth_order_value bc1.rs:471-545,
 * periodic_value bc1.rs:547-607).  src_i/src_j = -1 when not applicable. */
typedef struct {
  int32_t i, j;
  int32_t boundary;     /* 0 = outer, 1 = inner                               */
  int32_t side_id;
  int32_t bc_kind;      /* rh3t_bc_kind of the selected side                  */
  int32_t heavy_id;     /* index into the GhostBC tables, -1 for cheap BCs    */
  int32_t src_i, src_j; /* static source cell of a cheap BC                   */
  int32_t nearest_i, nearest_j; /* GhostInfo::nearest_idx (raw projection)   */
  double p0x, p0y;      /* final projection (arc override applied)           */
  double nx, ny;
  double distance;
} rh3t_ghost;

/* GhostBC, src/bc1.rs:49-57: per heavy ghost WENO-extrapolation tables. */
typedef struct {
  int32_t nearest_i, nearest_j;
  double normal[2];
  double d[5];
  uint32_t stencil[55];  /* linear fluid indices i*ny+j, r-major             */
  double s[979];         /* S_r blocks at offsets 0,1,17,98,354              */
  double w[225];         /* w[r][k] blocks at offsets 0,1,9,36,100           */
} rh3t_ghost_bc;

typedef struct {
  int32_t kind;          /* rh3t_status of the failure                        */
  int32_t stage;         /* 1..3 = "RK1/RK2/RK3 state", 0 = l()/ghost update */
  int32_t i, j;          /* grid point (ghost index for RH3T_ERR_GHOST)       */
  int32_t ghost_id;      /* -1 unless kind == RH3T_ERR_GHOST                  */
  int32_t reserved;
  double state[6];
  char message[256];
} rh3t_error;

typedef struct rh3t_problem rh3t_problem;
typedef struct rh3t_ctx rh3t_ctx;

/* ------------------------------------------------------------------------ */
/* Host side: Field / GhostGrid construction (geometry is integer-exact).   */
/* ------------------------------------------------------------------------ */

/* constant.rs defaults. */
void rh3t_physics_default(rh3t_physics* p);

/* Field::new, src/field1.rs:95-136.  Polygons are (x,y) pairs.  The outer
 * polygon has the fluid inside, the inner polygon the fluid outside;
 * bc_outer[k] / bc_inner[k] belong to side k (point k -> point k+1). */
int rh3t_problem_create(const rh3t_grid* grid,
                        const double* outer_xy, int32_t n_outer,
                        const rh3t_bc* bc_outer,
                        const double* inner_xy, int32_t n_inner,
                        const rh3t_bc* bc_inner,
                        rh3t_problem** out);
void rh3t_problem_destroy(rh3t_problem* pb);

/* ArcOverride registration (Field::outer_arcs / inner_arcs, src/field1.rs:
 * 89-92) with CircularArc::from_three_points (src/geometry.rs:493-551);
 * fluid_outside selects FluidSide::Outside (the cylinder of init.rs:336-344)
 * or FluidSide::Inside for the arc's outward normal. */
int rh3t_problem_add_arc(rh3t_problem* pb, int32_t boundary,
                         int32_t side_start, int32_t side_end,
                         int32_t fluid_outside,
                         const double start_xy[2], const double mid_xy[2],
                         const double end_xy[2]);

/* GhostGrid::build with default_stencil_offsets(), src/ghost.rs:50-118 and
 * src/ghost.rs:482-493, including precompute_ghost_bc (src/bc1.rs:108-188)
 * for Wall / Outflow / FarField ghosts. */
int rh3t_problem_build_ghosts(rh3t_problem* pb);

/* init::init_cylinder, src/init.rs:116-510, generalised as the survey of the
 * batched workload requires: nx in {121,161,201,241,...} with h = 3/(nx-1)
 * and ny = round(12/h)+1, freestream Mach number `mach` (3.0 reproduces the
 * reference exactly).  The returned problem has its ghosts built. */
int rh3t_problem_init_cylinder(int32_t nx, double mach, rh3t_problem** out);

/* init::init_double_mach, src/init.rs:26-113, with the mesh width h as a
 * parameter (1/320 reproduces the reference): the Mach-10 double Mach
 * reflection polygon with its Constant / Wall / ZerothOrder sides.  The
 * returned problem has its ghosts built. */
int rh3t_problem_init_double_mach(double h, rh3t_problem** out);

/* init_double_mach with the BC of its top side D -> E given by the caller
 * (top == NULL: the reference's ZerothOrder).  The reference defines the
 * moving Mach-10 shock closure for that side (src/init.rs:61-65) but
 * assigns ZerothOrder (src/init.rs:67-73); passing TimeDependent with
 * RH3T_TD_MOVING_FRONT (state = post, state2 = pre, speed = 10,
 * front0 = 0) runs the closure.  The returned problem has its ghosts
 * built. */
int rh3t_problem_init_double_mach_top(double h, const rh3t_bc* top,
                                      rh3t_problem** out);

/* Initial state of init_double_mach (src/init.rs:97-104: post-shock where
 * x <= 0, pre-shock elsewhere, on fluid cells; zero elsewhere), AoS
 * nx*ny*6. */
int rh3t_problem_double_mach_state(const rh3t_problem* pb, double* u_aos);

/* Copy of a problem (fluid mask, BCs, ghost list).  The heavy-ghost tables
 * depend only on geometry and grid and are shared with the source, on the
 * host and, inside one rh3t_batch, on the device. */
int rh3t_problem_clone(const rh3t_problem* src, rh3t_problem** out);

/* Replace the state of every FarField (kind = RH3T_BC_FARFIELD) or
 * Constant (RH3T_BC_CONSTANT) side, e.g. the inflow Mach number of the
 * cylinder.  Ghost tables are unaffected. */
int rh3t_problem_set_bc_state(rh3t_problem* pb, int32_t kind,
                              const double state[6]);

/* Initial state of init_cylinder (uniform freestream on fluid cells, zero
 * elsewhere), AoS nx*ny*6. */
int rh3t_problem_cylinder_state(const rh3t_problem* pb, double mach,
                                double* u_aos);

int rh3t_problem_grid(const rh3t_problem* pb, rh3t_grid* out);
int rh3t_problem_counts(const rh3t_problem* pb, int64_t* n_fluid,
                        int64_t* n_ghost, int64_t* n_heavy);
int rh3t_problem_fluid_mask(const rh3t_problem* pb, uint8_t* mask);
int rh3t_problem_ghosts(const rh3t_problem* pb, rh3t_ghost* out);
int rh3t_problem_ghost_bc(const rh3t_problem* pb, int32_t heavy_id,
                          rh3t_ghost_bc* out);
int rh3t_problem_bc(const rh3t_problem* pb, int32_t boundary, int32_t side,
                    rh3t_bc* out);
/* The reference's per-fluid-cell access table (src/ghost.rs:68-107):
 * access[l*17+k] = target fluid index, or nx*ny + ghost_id.  Rows of
 * non-fluid cells are 0.  Size nx*ny*17 u32. */
int rh3t_problem_access_table(const rh3t_problem* pb, uint32_t* access);

/* ------------------------------------------------------------------------ */
/* Device side: the l() / rk3_ssp() hot path on one MI355X.                 */
/* ------------------------------------------------------------------------ */

int rh3t_device_count(int32_t* n);

/* Uploads the problem's fluid mask and ghost tables; allocates the state,
 * stage and scratch buffers (main.rs:439-445 Scratch + u1/u2/u3). */
int rh3t_ctx_create(const rh3t_problem* pb, const rh3t_physics* phys,
                    int32_t device, rh3t_ctx** out);
void rh3t_ctx_destroy(rh3t_ctx* ctx);

/* Field.value / Field.time in and out (only fluid cells are read). */
int rh3t_set_state(rh3t_ctx* ctx, const double* u_aos, double time);
int rh3t_get_state(rh3t_ctx* ctx, double* u_aos, double* time);

/* l(u, ghosts, s), src/main.rs:135-305: writes Scratch::rhs (AoS nx*ny*6,
 * zero on non-fluid cells).  ghost_vals (n_ghost*6, may be NULL) receives
 * GhostGrid::values of this evaluation. */
int rh3t_l(rh3t_ctx* ctx, double* rhs_aos, double* ghost_vals);

/* rk3_ssp, src/main.rs:351-373: one SSP-RK3 step of size dt in place. */
int rh3t_rk3_step(rh3t_ctx* ctx, double dt);

/* calc_global_dt, src/main.rs:375-408 (min of dt::get_local_dt). */
int rh3t_calc_global_dt(rh3t_ctx* ctx, double* dt_cfl);

/* The time loop body of main(), src/main.rs:459-467, kept on the device:
 * nsteps x { dt = cfl_factor * calc_global_dt; rk3_ssp(dt) } with no host
 * round trip; the step is captured once in a hipGraph and replayed.  If
 * dt_fixed > 0 it is used instead of the CFL rule.  On return *time holds
 * Field.time. */
int rh3t_advance(rh3t_ctx* ctx, int32_t nsteps, double cfl_factor,
                 double dt_fixed, double* time);

/* Arithmetic of the WENO interface flux and assembly kernels.  EXACT (the
 * default of every context and batch) keeps the reference's rounding
 * sequence (literal divisions, left-to-right sums, no FMA) and agrees with
 * the reference's l() and rk3_ssp() to 1e-12 on every problem.  FAST uses
 * the same algebra with fewer instructions (one reciprocal per shared
 * denominator, FMA chains, WENO5 on first differences, flux projections
 * from state projections): agreement with EXACT to roundoff, within 1e-12
 * on smooth periodic boxes (6.4e-13 at 2048^2) but not on the cylinder's
 * first full RK3 step (1.4e-12), where the boundary WENO weights
 * (eps + beta)^-10 (src/bc1.rs:241) amplify last-bit differences.  No
 * counterpart in the reference (which has one arithmetic). */
#define RH3T_ARITH_EXACT 0
#define RH3T_ARITH_FAST 1
int rh3t_ctx_set_arith(rh3t_ctx* ctx, int32_t mode);

/* Replaces the physics constants (src/constant.rs:1-25 as a runtime
 * struct) of an existing context, e.g. to evaluate the operator at other
 * exchange or diffusion coefficients.  Problem tables are kept; the next
 * rh3t_advance recomputes dt from the state with the new constants. */
int rh3t_ctx_set_physics(rh3t_ctx* ctx, const rh3t_physics* phys);

/* Workgroup shape of the WENO flux / assembly kernels (bx lanes along y,
 * by rows along x; bx a multiple of 16, bx*by <= 256 and a multiple of 64)
 * and heavy ghosts per workgroup (1..4, one wavefront each).  Other values
 * return RH3T_ERR_ARG and keep the current shape.  Default 64 x 1 with one
 * ghost per workgroup; the batched runs' shapes are chosen by the autotuner
 * in 3t_rh_amd/autotune.py.  No counterpart in the reference. */
int rh3t_ctx_set_launch(rh3t_ctx* ctx, int32_t bx, int32_t by,
                        int32_t ghosts_per_block);
/* Workgroup shape of the x-direction flux kernel alone (same limits), which
 * otherwise follows rh3t_ctx_set_launch; (0, 0) returns it to that.  The
 * fused y-flux + assembly pass needs rows of 16, 32 or 64 lanes, the x flux
 * does not.  No counterpart in the reference. */
int rh3t_ctx_set_launch_x(rh3t_ctx* ctx, int32_t bx, int32_t by);

/* First failure of the last call (main.rs:24-33 / ghost.rs:229-263). */
int rh3t_last_error(const rh3t_ctx* ctx, rh3t_error* err);

/* Per-kernel timing of the last rh3t_advance / rh3t_rk3_step, measured with
 * hipEvents on the context's stream.  names: NUL-separated list written
 * into `names` (capacity names_cap); avg_ms[k], launches[k] per kernel. */
int rh3t_profile(rh3t_ctx* ctx, int32_t enable);
int rh3t_profile_read(rh3t_ctx* ctx, int32_t max_kernels, char* names,
                      int32_t names_cap, double* avg_ms, int64_t* launches,
                      int32_t* n_kernels);

/* calc_global_dt (src/main.rs:375-408) of a state that already lives in
 * device memory, on the caller's stream, allocating nothing.  No
 * counterpart in the reference: it is the variant of rh3t_calc_global_dt
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops) and hold the state themselves.  Device
 * pointers, on the current device: u_aos (nx*ny*6 doubles, AoS at
 * l = i*ny + j), fluid (nx*ny bytes, nonzero = fluid cell), dt_cfl (one
 * double, written).  grid and phys (NULL: src/constant.rs defaults) are
 * host pointers; stream is a hipStream_t (NULL: the null stream).  The
 * per-workgroup minima live in a fixed device array of the library (one
 * per GPU, held until the process ends): nx*ny above
 * RH3T_NOALLOC_MAX_CELLS returns RH3T_ERR_SIZE and launches nothing.
 * Calls that overlap on one GPU share that array and race; a call's result
 * is complete when its stream has reached it. */
#define RH3T_NOALLOC_MAX_CELLS (1 << 25)
int rh3t_calc_global_dt_noalloc(const double* u_aos, const uint8_t* fluid,
                                const rh3t_grid* grid,
                                const rh3t_physics* phys, double* dt_cfl,
                                void* stream);

/* ------------------------------------------------------------------------ */
/* Batched solves: many independent problems (uneven grids, different      */
/* inflow states) advanced together.  No counterpart in the reference,     */
/* which runs one Field per process; each solve behaves exactly as an      */
/* rh3t_ctx on its own problem (same kernels, per-solve dt).               */
/* ------------------------------------------------------------------------ */

typedef struct rh3t_batch rh3t_batch;

int rh3t_batch_create(const rh3t_problem* const* problems, int32_t n,
                      const rh3t_physics* phys, int32_t device,
                      rh3t_batch** out);
void rh3t_batch_destroy(rh3t_batch* b);
int rh3t_batch_set_state(rh3t_batch* b, int32_t solve, const double* u_aos,
                         double time);
int rh3t_batch_get_state(rh3t_batch* b, int32_t solve, double* u_aos,
                         double* time);
/* nsteps x { per solve: dt_s = cfl_factor * calc_global_dt_s; rk3_ssp(dt_s) }
 * for all solves at once: one hipGraph of batched kernels per step, every
 * grid-shaped kernel covering the tiles of all solves.  times[n] (may be
 * NULL) receives each solve's Field.time. */
int rh3t_batch_advance(rh3t_batch* b, int32_t nsteps, double cfl_factor,
                       double* times);
/* rh3t_ctx_set_arith for every solve of the batch (EXACT by default). */
int rh3t_batch_set_arith(rh3t_batch* b, int32_t mode);
int rh3t_batch_set_launch(rh3t_batch* b, int32_t bx, int32_t by,
                          int32_t ghosts_per_block);
int rh3t_batch_set_launch_x(rh3t_batch* b, int32_t bx, int32_t by);
int rh3t_batch_last_error(const rh3t_batch* b, int32_t solve,
                          rh3t_error* err);
int rh3t_batch_profile(rh3t_batch* b, int32_t enable);
int rh3t_batch_profile_read(rh3t_batch* b, int32_t max_kernels, char* names,
                            int32_t names_cap, double* avg_ms,
                            int64_t* launches, int32_t* n_kernels);

/* ------------------------------------------------------------------------ */
/* Reverse mode.  No counterpart in the reference (which has no            */
/* derivatives): vector-Jacobian products of l() and of rk3_ssp() with     */
/* respect to the state and to the parameters                               */
/*   pgrad[5] = (omega_ei, omega_er, kappa_e, kappa_i, kappa_r)             */
/* (src/constant.rs:17-25), and a checkpoint-and-recompute adjoint of a     */
/* sequence of RK3 steps with a frozen dt schedule.  Covered ghost kinds:  */
/* every kind but TimeDependent (which returns RH3T_ERR_UNSUPPORTED);     */
/* single-solve contexts only.  Cotangent arrays                            */
/* use the state's AoS layout; entries of non-fluid cells are ignored on   */
/* input and zero on output.  pgrad may be NULL.                            */
/* ------------------------------------------------------------------------ */

/* ubar = (d l / d u)^T rbar at the context's state (Scratch::rhs cotangent
 * rbar, src/main.rs:135-305). */
int rh3t_l_vjp(rh3t_ctx* ctx, const double* rbar, double* ubar,
               double* pgrad);

/* ubar = (d u^{n+1} / d u^n)^T lam for one rk3_ssp step of size dt from the
 * context's state (src/main.rs:351-373); the state is not advanced. */
int rh3t_rk3_vjp(rh3t_ctx* ctx, double dt, const double* lam, double* ubar,
                 double* pgrad);

typedef struct rh3t_adjoint rh3t_adjoint;

/* Adjoint of nsteps RK3 steps of ctx (which must outlive it): the forward
 * keeps the state every checkpoint_every steps; the backward recomputes
 * each segment from its checkpoint (taping the stage states) and sweeps it
 * in reverse.  Device memory: the checkpoints plus two segments of stage
 * states (2 checkpoint_every steps x (3 states + their Derived planes)). */
int rh3t_adjoint_create(rh3t_ctx* ctx, int32_t checkpoint_every,
                        rh3t_adjoint** out);
void rh3t_adjoint_destroy(rh3t_adjoint* adj);
/* nsteps rk3_ssp steps from the context's state with dt[n]; Field.time
 * advances by each dt. */
int rh3t_adjoint_forward(rh3t_adjoint* adj, int32_t nsteps, const double* dt);
/* lam_initial = (d u^N / d u^0)^T lam_final and the parameter cotangent,
 * for the last forward.  The context's state is not touched (it stays at
 * u^N unless the caller changed it).  The stage states of the forward's
 * last two checkpoint segments are kept, so the backward recomputes only
 * the segments before them, each on a second stream while the reverse
 * sweep of the segment after it runs. */
int rh3t_adjoint_backward(rh3t_adjoint* adj, const double* lam_final,
                          double* lam_initial, double* pgrad);
/* Device time in ms of the last forward and the last backward (HIP events
 * on the context's stream, both streams joined): from after the host inputs
 * are on the device to before the results are copied back, so the
 * backward's cotangent upload and gradient download are not included.  No
 * counterpart in the reference (which has no reverse mode); for measuring
 * the adjoint's cost against the forward's with device-resident data.
 * RH3T_ERR_ARG before one forward and one backward have run. */
int rh3t_adjoint_device_ms(const rh3t_adjoint* adj, double* forward_ms,
                           double* backward_ms);

#ifdef __cplusplus
}
#endif

#endif /* RH3T_H */
