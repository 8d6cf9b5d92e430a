/*
 * rh3t_oracle.h -- CPU restatement of the reference's l() / rk3_ssp() path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation and the reported CPU baseline; it is never linked into or
 * called by the product library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.
 *
 * It follows colinhu123/3T_RH in operation order (plain C99, compiled with
 * -ffp-contract=off, no fast-math), with OpenMP loops at the reference's
 * Rayon sites and the data model of the reference: AoS states at
 * l = i*ny + j, a ghost value array indexed by sorted ghost id, and the
 * 17-slot per-fluid-cell access table of src/ghost.rs:68-107.
 */
#ifndef RH3T_ORACLE_H
#define RH3T_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double gamma_e, gamma_i, gamma_r;
  double kappa_e, kappa_i, kappa_r;
  double cve, cvi, a_rad;
  double omega_ei, omega_er;
  double lambda, weno_q, eps;
} ro_physics;

typedef struct {
  double u, v, e_e, e_i, e_r, pe, pi, pr, te, ti, tr;
} ro_derived;

/* BC kinds use the same numbering as BCType in src/bc1.rs:16-28 as listed:
 * 0 ReflectiveWall, 1 Wall, 2 Periodic, 3 Constant, 4 TimeDependent,
 * 5 Outflow, 6 ZerothOrder, 7 FarField. */
typedef struct {
  int32_t nx, ny;
  double dx, dy;
  const uint8_t* fluid;       /* nx*ny                                        */
  const uint32_t* access;     /* nx*ny*17                                     */
  int32_t n_ghost;
  const int32_t* g_kind;      /* per ghost                                    */
  const int32_t* g_src;       /* linear source index (cheap BCs) or -1        */
  const int32_t* g_heavy;     /* heavy id or -1                               */
  const double* g_normal;     /* 2 per ghost (final projection normal)        */
  const double* g_dist;       /* per ghost (signed distance D)                */
  const double* g_bcstate;    /* 6 per ghost: Constant / FarField state       */
  const double* g_outflow;    /* 3 per ghost: p_inf, sigma, l_domain          */
  const int32_t* h_nearest;   /* per heavy: linear index of GhostBC::nearest  */
  const double* h_d;          /* 5 per heavy                                  */
  const uint32_t* h_stencil;  /* 55 per heavy                                 */
  const double* h_s;          /* 979 per heavy                                */
  const double* h_w;          /* 225 per heavy                                */
  /* TimeDependent (moving front profile) and the state's Field.time */
  double time;                /* time of the state ro_l evaluates             */
  const double* g_p0x;        /* per ghost: x of the projection point P0       */
  const double* g_bcstate2;   /* 6 per ghost: TimeDependent state ahead        */
  const double* g_front;      /* 2 per ghost: TimeDependent speed, front0      */
} ro_problem;

typedef struct {
  int32_t kind;   /* 0 ok, 3 non-physical state, 4 non-finite ghost, 8 singular */
  int32_t stage;  /* 1..3 */
  int32_t i, j;
  int32_t ghost_id;
  double state[6];
} ro_error;

/* --- pointwise kernels (exposed for the known-answer tests) ----------- */
void ro_derived_from_state(const ro_physics* ph, const double s[6], ro_derived* d);
double ro_weno5(const double st[5], double eps);
/* dir 0 = X, 1 = Y.  L (6x6 row-major) of weno.rs:28-459 */
void ro_build_l(const ro_physics* ph, const double s3[6], const double s2[6], const ro_derived* d3,
                const ro_derived* d2, int dir, double L[36]);
void ro_build_r(const ro_physics* ph, const double s3[6], const double s2[6], const ro_derived* d3,
                const ro_derived* d2, int dir, double lambda[6], double R[36]);
void ro_reconstruction_fast(const ro_physics* ph, const double pts[36], const ro_derived d[6],
                            int dir, int recon_type, double out[6]);
void ro_noncon(const ro_derived d[9], double ds, int dir, double out[6]);
void ro_diffusion(const ro_physics* ph, const ro_derived d[6], double out[6]);
void ro_source(const ro_physics* ph, const double s[6], double out[6]);
double ro_local_dt(const ro_physics* ph, const double s[6], double dx, double dy);

/* Heavy ghost g (Wall / Outflow / FarField): L, R, lambda (6x6 row-major)
 * of the nearest state in the (n, t) frame, V^(k) (v[k*6+c]) and the
 * normal derivatives u^(k) (uk[k*6+f]) before the Taylor expansion.
 * Returns 0, 1 (not a heavy ghost) or 8 (singular ILW system). */
int ro_heavy_terms(const ro_problem* pb, const ro_physics* ph, const double* u, int g, double L[36],
                   double R[36], double lam[6], double v[30], double uk[30]);

/* --- operator level -------------------------------------------------- */
/* ghosts.update_values_parallel: ghost_vals (n_ghost*6). Returns 0 or 4/8. */
int ro_update_ghosts(const ro_problem* pb, const ro_physics* ph, const double* u,
                     double* ghost_vals, int nthreads, ro_error* err);
/* l(): rhs (nx*ny*6). ghost_vals may be NULL. */
int ro_l(const ro_problem* pb, const ro_physics* ph, const double* u, double* rhs,
         double* ghost_vals, int nthreads, ro_error* err);
/* rk3_ssp in place on u (nx*ny*6); *time advanced by dt. */
int ro_rk3_step(const ro_problem* pb, const ro_physics* ph, double* u, double* time, double dt,
                int nthreads, ro_error* err);
/* calc_global_dt (serial, as main.rs:387-405) */
double ro_global_dt(const ro_problem* pb, const ro_physics* ph, const double* u);

#ifdef __cplusplus
}
#endif
#endif
