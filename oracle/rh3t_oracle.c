/*
 * rh3t_oracle.c -- CPU restatement of colinhu123/3T_RH's l() / rk3_ssp().
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the reported CPU baseline; never
 * part of the product).  Every function cites the reference lines it
 * restates and keeps their operation order: left-associative sums, literal
 * divisions, x.powi(n) as repeated squaring, f64::min/max as fmin/fmax.
 * Compile with -ffp-contract=off and without fast-math.
 *
 * Parity status of each function (how it is pinned) is listed in the
 * header comment of tests/test_oracle.py.
 */
#include "rh3t_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* x.powi(n), n >= 0: repeated squaring as LLVM lowers constant exponents
 * and compiler-builtins' __powidf2 evaluates runtime ones. */
static inline double powi(double x, int n) {
  double mul = 1.0, a = x;
  for (;;) {
    if (n & 1) mul *= a;
    n >>= 1;
    if (n == 0) break;
    a *= a;
  }
  return mul;
}

/* (eps + beta)^q of the boundary WENO weights (powf(WENO_Q), src/bc1.rs:241).
 * For q = 10 the power is evaluated in double-double arithmetic and rounded
 * once, i.e. correctly rounded -- the value glibc's pow returns except for
 * inputs within ~2^-60 of a rounding midpoint.  The GPU kernel evaluates it
 * the same way, so the two agree bit for bit. */
static inline void dd_mul(double ah, double al, double bh, double bl, double* rh, double* rl) {
  double p = ah * bh;
  double e = fma(ah, bh, -p);
  e += ah * bl + al * bh;
  double s = p + e;
  *rl = e - (s - p);
  *rh = s;
}
static inline double weno_pow(double x, double q) {
  if (q != 10.0) return pow(x, q);
  double x2h = x * x, x2l = fma(x, x, -x2h);
  double x4h, x4l, x8h, x8l, rh, rl;
  dd_mul(x2h, x2l, x2h, x2l, &x4h, &x4l);
  dd_mul(x4h, x4l, x4h, x4l, &x8h, &x8l);
  dd_mul(x8h, x8l, x2h, x2l, &rh, &rl);
  return rh + rl;
}

/* ------------------------------------------------------------ state.rs */

/* Derived::from_state, src/state.rs:241-275 */
void ro_derived_from_state(const ro_physics* ph, const double s[6], ro_derived* d) {
  double rho = s[0];
  double u = s[1] / rho;
  double v = s[2] / rho;
  double w2 = powi(u, 2) + powi(v, 2);
  d->u = u;
  d->v = v;
  d->e_e = s[3] / rho - w2 / 6.0;
  d->e_i = s[4] / rho - w2 / 6.0;
  d->e_r = s[5] / rho - w2 / 6.0;
  d->pe = (ph->gamma_e - 1.0) * (s[3] - rho * w2 / 6.0);
  d->pi = (ph->gamma_i - 1.0) * (s[4] - rho * w2 / 6.0);
  d->pr = (ph->gamma_r - 1.0) * (s[5] - rho * w2 / 6.0);
  d->te = s[3] / (rho * ph->cve) - w2 / (6.0 * ph->cve);
  d->ti = s[4] / (rho * ph->cvi) - w2 / (6.0 * ph->cvi);
  double er_term = s[5] / rho - w2 / 6.0;
  d->tr = pow(er_term * rho / ph->a_rad, 0.25);
}

/* State::flux_from_derived, src/state.rs:56-85 */
static void flux_from_derived(const double s[6], const ro_derived* d, int dir, double f[6]) {
  double u = d->u, v = d->v;
  double p = d->pe + d->pi + d->pr;
  double rho = s[0];
  if (dir == 0) {
    f[0] = rho * u;
    f[1] = rho * u * u + p;
    f[2] = rho * u * v;
    f[3] = (s[3] + d->pe) * u;
    f[4] = (s[4] + d->pi) * u;
    f[5] = (s[5] + d->pr) * u;
  } else {
    f[0] = rho * v;
    f[1] = rho * u * v;
    f[2] = rho * v * v + p;
    f[3] = (s[3] + d->pe) * v;
    f[4] = (s[4] + d->pi) * v;
    f[5] = (s[5] + d->pr) * v;
  }
}

/* ------------------------------------------------------------- weno.rs */

/* Roe-averaged quantities shared by build_l_plain / build_r_plain,
 * src/weno.rs:34-69 and 491-512 (identical expressions in both). */
typedef struct {
  double u, v, w2, ee, ei, er, gi, ge, gr, gt, cs2, cs;
} roe_t;

static void roe_average(const ro_physics* ph, const double s1[6], const double s2[6],
                        const ro_derived* d3, const ro_derived* d2, roe_t* a) {
  double rho1 = sqrt(s1[0]);
  double rho2 = sqrt(s2[0]);
  double rho_sum = rho1 + rho2;
  a->u = (d3->u * rho1 + d2->u * rho2) / rho_sum;
  a->v = (d3->v * rho1 + d2->v * rho2) / rho_sum;
  a->w2 = powi(a->u, 2) + powi(a->v, 2);
  a->ee = (d3->e_e * rho1 + d2->e_e * rho2) / rho_sum;
  a->ei = (d3->e_i * rho1 + d2->e_i * rho2) / rho_sum;
  a->er = (d3->e_r * rho1 + d2->e_r * rho2) / rho_sum;
  a->gi = ph->gamma_i - 1.0;
  a->ge = ph->gamma_e - 1.0;
  a->gr = ph->gamma_r - 1.0;
  a->gt = a->gi + a->ge + a->gr;
  a->cs2 = ph->gamma_e * a->ge * a->ee + ph->gamma_i * a->gi * a->ei + ph->gamma_r * a->gr * a->er;
  a->cs = sqrt(a->cs2);
}

/* Rows L_A^(3..5) / L_B^(3..5) (energy modes), src/weno.rs:110-233:
 * gk = gamma_k - 1 of the mode, hk its h, and (g1, g2) the other two. */
static void energy_row(double gk, double hk, double u, double v, double w2, double cs2, double gt,
                       double b, double ge, double gi, double gr, int mode, double row[6]) {
  row[0] = gk / gt - gk * w2 / (6.0 * cs2) - gk * gt * w2 / b * (hk - cs2 * w2);
  row[1] = gk * u / (3.0 * cs2) + 2.0 * gk * gt * u / b * (hk - cs2 * w2);
  row[2] = gk * v / (3.0 * cs2) + 2.0 * gk * gt * v / b * (hk - cs2 * w2);
  if (mode == 0) { /* electron: src/weno.rs:120-160 */
    row[3] = -powi(ge, 2) / (gt * cs2) - 6.0 * ge / b * (ge * hk - gt * cs2 * w2);
    row[4] = -ge * gi / (gt * cs2) - 6.0 * ge * gi / b * (hk - 6.0 * gr * cs2);
    row[5] = -ge * gr / (gt * cs2) - 6.0 * ge * gr / b * (hk + 6.0 * gi * cs2);
  } else if (mode == 1) { /* ion: src/weno.rs:162-198 */
    row[3] = -gi * ge / (gt * cs2) - 6.0 * gi * ge / b * (hk + 6.0 * gr * cs2);
    row[4] = -powi(gi, 2) / (gt * cs2) - 6.0 * gi / b * (gi * hk - gt * cs2 * w2);
    row[5] = -gi * gr / (gt * cs2) - 6.0 * gi * gr / b * (hk - 6.0 * ge * cs2);
  } else { /* radiation: src/weno.rs:200-233 */
    row[3] = -gr * ge / (gt * cs2) - 6.0 * gr * ge / b * (hk - 6.0 * gi * cs2);
    row[4] = -gr * gi / (gt * cs2) - 6.0 * gr * gi / b * (hk + 6.0 * ge * cs2);
    row[5] = -powi(gr, 2) / (gt * cs2) - 6.0 * gr / b * (gr * hk - gt * cs2 * w2);
  }
  (void)gk;
}

/* Stencil6::build_l_plain, src/weno.rs:28-459 */
void ro_build_l(const ro_physics* ph, const double s3[6], const double s2[6], const ro_derived* d3,
                const ro_derived* d2, int dir, double L[36]) {
  roe_t a;
  roe_average(ph, s3, s2, d3, d2, &a);
  double u = a.u, v = a.v, w2 = a.w2, cs = a.cs, cs2 = a.cs2, gt = a.gt;
  double ge = a.ge, gi = a.gi, gr = a.gr;
  double he = 6.0 * gi * gr * (ph->gamma_i * a.ei - ph->gamma_r * a.er) + ph->gamma_e * gt * a.ee * w2;
  double hi = 6.0 * ge * gr * (ph->gamma_r * a.er - ph->gamma_e * a.ee) + ph->gamma_i * gt * a.ei * w2;
  double hr = 6.0 * ge * gi * (ph->gamma_e * a.ee - ph->gamma_i * a.ei) + ph->gamma_r * gt * a.er * w2;
  double b = gt * (36.0 * ge * gi * gr + gt * powi(w2, 2)) * cs2;
  double ad = 12.0 * cs2;
  double* r0 = L + 0;
  double* r5 = L + 30;
  int ie, ii, ir, it; /* row slots of the energy and transverse modes */
  if (dir == 0) {
    r0[0] = (gt * w2 + 6.0 * u * cs) / ad;
    r0[1] = (-2.0 * gt * u - 6.0 * cs) / ad;
    r0[2] = (-2.0 * gt * v) / ad;
    r5[0] = (gt * w2 - 6.0 * u * cs) / ad;
    r5[1] = (-2.0 * gt * u + 6.0 * cs) / ad;
    r5[2] = (-2.0 * gt * v) / ad;
    ie = 1; ii = 2; ir = 3; it = 4;
    L[it * 6 + 0] = -v; L[it * 6 + 1] = 0.0; L[it * 6 + 2] = 1.0;
  } else {
    r0[0] = (gt * w2 + 6.0 * v * cs) / ad;
    r0[1] = (-2.0 * gt * u) / ad;
    r0[2] = (-2.0 * gt * v - 6.0 * cs) / ad;
    r5[0] = (gt * w2 - 6.0 * v * cs) / ad;
    r5[1] = (-2.0 * gt * u) / ad;
    r5[2] = (-2.0 * gt * v + 6.0 * cs) / ad;
    it = 1; ie = 2; ii = 3; ir = 4;
    L[it * 6 + 0] = -u; L[it * 6 + 1] = 1.0; L[it * 6 + 2] = 0.0;
  }
  L[it * 6 + 3] = 0.0; L[it * 6 + 4] = 0.0; L[it * 6 + 5] = 0.0;
  r0[3] = 6.0 * ge / ad; r0[4] = 6.0 * gi / ad; r0[5] = 6.0 * gr / ad;
  r5[3] = 6.0 * ge / ad; r5[4] = 6.0 * gi / ad; r5[5] = 6.0 * gr / ad;
  energy_row(ge, he, u, v, w2, cs2, gt, b, ge, gi, gr, 0, L + ie * 6);
  energy_row(gi, hi, u, v, w2, cs2, gt, b, ge, gi, gr, 1, L + ii * 6);
  energy_row(gr, hr, u, v, w2, cs2, gt, b, ge, gi, gr, 2, L + ir * 6);
}

/* Stencil6::build_r_plain, src/weno.rs:484-548 */
void ro_build_r(const ro_physics* ph, const double s3[6], const double s2[6], const ro_derived* d3,
                const ro_derived* d2, int dir, double lambda[6], double R[36]) {
  roe_t a;
  roe_average(ph, s3, s2, d3, d2, &a);
  double u = a.u, v = a.v, w2 = a.w2, cs = a.cs, gt = a.gt;
  double ge = a.ge, gi = a.gi, gr = a.gr;
  double ge_ = ph->gamma_e * a.ee + w2 / 6.0;
  double gi_ = ph->gamma_i * a.ei + w2 / 6.0;
  double gr_ = ph->gamma_r * a.er + w2 / 6.0;
  if (dir == 0) {
    double ucs = u * cs / 3.0;
    const double r[36] = {1.0, 1.0, 1.0, 1.0, 0.0, 1.0,
                          u - cs, u, u, u, 0.0, u + cs,
                          v, v, v, v, 1.0, v,
                          ge_ - ucs, gt * w2 / (6.0 * ge), -gr, gi, v / 3.0, ge_ + ucs,
                          gi_ - ucs, gr, gt * w2 / (6.0 * gi), -ge, v / 3.0, gi_ + ucs,
                          gr_ - ucs, -gi, ge, gt * w2 / (6.0 * gr), v / 3.0, gr_ + ucs};
    memcpy(R, r, sizeof(r));
    lambda[0] = u - cs; lambda[1] = u; lambda[2] = u; lambda[3] = u; lambda[4] = u; lambda[5] = u + cs;
  } else {
    double vcs = v * cs / 3.0;
    const double r[36] = {1.0, 0.0, 1.0, 1.0, 1.0, 1.0,
                          u, 1.0, u, u, u, u,
                          v - cs, 0.0, v, v, v, v + cs,
                          ge_ - vcs, u / 3.0, gt * w2 / (6.0 * ge), -gr, gi, ge_ + vcs,
                          gi_ - vcs, u / 3.0, gr, gt * w2 / (6.0 * gi), -ge, gi_ + vcs,
                          gr_ - vcs, u / 3.0, -gi, ge, gt * w2 / (6.0 * gr), gr_ + vcs};
    memcpy(R, r, sizeof(r));
    lambda[0] = v - cs; lambda[1] = v; lambda[2] = v; lambda[3] = v; lambda[4] = v; lambda[5] = v + cs;
  }
}

/* weno5, src/weno.rs:846-886 */
double ro_weno5(const double st[5], double eps) {
  double u0 = st[0], u1 = st[1], u2 = st[2], u3 = st[3], u4 = st[4];
  double beta2 = 13.0 / 12.0 * powi(u2 - 2.0 * u3 + u4, 2) + 0.25 * powi(3.0 * u2 - 4.0 * u3 + u4, 2);
  double beta1 = 13.0 / 12.0 * powi(u1 - 2.0 * u2 + u3, 2) + 0.25 * powi(u1 - u3, 2);
  double beta0 = 13.0 / 12.0 * powi(u0 - 2.0 * u1 + u2, 2) + 0.25 * powi(u0 - 4.0 * u1 + 3.0 * u2, 2);
  double a0 = 0.1 / powi(eps + beta0, 2);
  double a1 = 0.6 / powi(eps + beta1, 2);
  double a2 = 0.3 / powi(eps + beta2, 2);
  double sum_o = a0 + a1 + a2;
  double w0 = a0 / sum_o, w1 = a1 / sum_o, w2 = a2 / sum_o;
  double p0 = u0 / 3.0 - 7.0 / 6.0 * u1 + 11.0 / 6.0 * u2;
  double p1 = -u1 / 6.0 + 5.0 / 6.0 * u2 + u3 / 3.0;
  double p2 = u2 / 3.0 + 5.0 / 6.0 * u3 - u4 / 6.0;
  return w0 * p0 + w1 * p1 + w2 * p2;
}

/* l_dot / r_dot, src/weno.rs:823-843 */
static inline void mat6_dot(const double* m, const double* x, double* out) {
  for (int i = 0; i < 6; ++i) {
    const double* row = m + 6 * i;
    out[i] = row[0] * x[0] + row[1] * x[1] + row[2] * x[2] + row[3] * x[3] + row[4] * x[4] +
             row[5] * x[5];
  }
}

/* Stencil6::reconstruction_fast, src/weno.rs:685-802 */
void ro_reconstruction_fast(const ro_physics* ph, const double pts[36], const ro_derived d[6],
                            int dir, int recon_type, double out[6]) {
  double L[36], R[36], lam[6];
  ro_build_l(ph, pts + 18, pts + 12, &d[3], &d[2], dir, L);
  ro_build_r(ph, pts + 18, pts + 12, &d[3], &d[2], dir, lam, R);
  double cf[6][6], cs[6][6]; /* [point][component] */
  for (int i = 0; i < 6; ++i) {
    double fl[6];
    flux_from_derived(pts + 6 * i, &d[i], dir, fl);
    if (recon_type) {
      mat6_dot(L, fl, cf[i]);
      mat6_dot(L, pts + 6 * i, cs[i]);
    } else {
      memcpy(cf[i], fl, sizeof(fl));
      memcpy(cs[i], pts + 6 * i, sizeof(double) * 6);
    }
  }
  double a[6];
  for (int m = 0; m < 6; ++m) a[m] = fabs(lam[m]);
  double fp[6][6], fm[6][6]; /* [component][point] (stencil_arr layout) */
  for (int i = 0; i < 6; ++i)
    for (int m = 0; m < 6; ++m) {
      fp[m][i] = 0.5 * (cf[i][m] + a[m] * cs[i][m]);
      fm[m][i] = 0.5 * (cf[i][m] - a[m] * cs[i][m]);
    }
  double flux[6];
  for (int m = 0; m < 6; ++m) {
    double sp[5] = {fp[m][0], fp[m][1], fp[m][2], fp[m][3], fp[m][4]};
    double sm[5] = {fm[m][5], fm[m][4], fm[m][3], fm[m][2], fm[m][1]};
    flux[m] = ro_weno5(sp, ph->eps) + ro_weno5(sm, ph->eps);
  }
  mat6_dot(R, flux, out);
}

/* ----------------------------------------------------------- noncon.rs */

/* sixth_order_derivative, src/noncon.rs:26-44 */
static inline double sixth(double qm3, double qm2, double qm1, double qp1, double qp2, double qp3,
                           double dx) {
  return (qp3 - 9.0 * qp2 + 45.0 * qp1 - 45.0 * qm1 + 9.0 * qm2 - qm3) / (60.0 * dx);
}

/* q_pulse, src/noncon.rs:49-57 (literal decimal coefficients) */
static inline double q_pulse(const double b[7]) {
  return -0.00714285714285723 * b[0] + 0.0595238095238118 * b[1] - 0.240476190476197 * b[2] +
         0.759523809523819 * b[3] + 0.509523809523803 * b[4] - 0.0904761904761891 * b[5] +
         0.00952380952380977 * b[6];
}

/* n_jump_pre's jump_component, src/noncon.rs:259-285 */
static inline double jump_component(const double arr[8]) {
  double b2[7] = {arr[0], arr[1], arr[2], arr[3], arr[4], arr[5], arr[6]};
  double b1[7] = {arr[7], arr[6], arr[5], arr[4], arr[3], arr[2], arr[1]};
  return q_pulse(b1) - q_pulse(b2);
}

/* nonconservative_direction_pre, src/noncon.rs:288-332, with dN_pre
 * (235-257), n_jump_pre (259-285), n_vector_from_pressures (14-24). */
void ro_noncon(const ro_derived d[9], double ds, int dir, double out[6]) {
  double v = dir == 0 ? d[4].u : d[4].v;
  /* dN_pre on d[1..8] */
  double pe[7], pi[7], pr[7];
  for (int i = 0; i < 7; ++i) {
    pe[i] = d[1 + i].pe;
    pi[i] = d[1 + i].pi;
    pr[i] = d[1 + i].pr;
  }
  double dpe = sixth(pe[0], pe[1], pe[2], pe[4], pe[5], pe[6], ds);
  double dpi = sixth(pi[0], pi[1], pi[2], pi[4], pi[5], pi[6], ds);
  double dpr = sixth(pr[0], pr[1], pr[2], pr[4], pr[5], pr[6], ds);
  double der_ee = 2.0 * dpe - dpi - dpr;
  double der_ei = 2.0 * dpi - dpe - dpr;
  double der_er = 2.0 * dpr - dpe - dpi;
  /* N vectors on the 9 points */
  double nee[9], nei[9], ner[9];
  for (int k = 0; k < 9; ++k) {
    nee[k] = 2.0 * d[k].pe - d[k].pi - d[k].pr;
    nei[k] = 2.0 * d[k].pi - d[k].pe - d[k].pr;
    ner[k] = 2.0 * d[k].pr - d[k].pe - d[k].pi;
  }
  double jm_ee = jump_component(nee), jp_ee = jump_component(nee + 1);
  double jm_ei = jump_component(nei), jp_ei = jump_component(nei + 1);
  double jm_er = jump_component(ner), jp_er = jump_component(ner + 1);
  double vm = dir == 0 ? d[3].u : d[3].v;
  double vp = dir == 0 ? d[5].u : d[5].v;
  double coef_left = fmax(fmax(v, vm), 0.0) / 3.0;
  double coef_right = fmin(fmin(v, vp), 0.0) / 3.0;
  out[0] = 0.0;
  out[1] = 0.0;
  out[2] = 0.0;
  out[3] = v / 3.0 * der_ee + (coef_left * jm_ee + coef_right * jp_ee) / ds;
  out[4] = v / 3.0 * der_ei + (coef_left * jm_ei + coef_right * jp_ei) / ds;
  out[5] = v / 3.0 * der_er + (coef_left * jm_er + coef_right * jp_er) / ds;
}

/* build_diffusion_from_derived, src/diffusion.rs:27-50 */
void ro_diffusion(const ro_physics* ph, const ro_derived d[6], double out[6]) {
  const ro_derived *p3 = &d[5], *p2 = &d[4], *p1 = &d[3], *s0 = &d[2], *m1 = &d[1], *m2 = &d[0];
  double ee = ph->kappa_e * (2.0 * p3->te - 25.0 * p2->te + 245.0 * p1->te - 245.0 * s0->te +
                             25.0 * m1->te - 2.0 * m2->te);
  double ei = ph->kappa_i * (2.0 * p3->ti - 25.0 * p2->ti + 245.0 * p1->ti - 245.0 * s0->ti +
                             25.0 * m1->ti - 2.0 * m2->ti);
  double er = ph->kappa_r * (2.0 * powi(p3->tr, 4) - 25.0 * powi(p2->tr, 4) + 245.0 * powi(p1->tr, 4) -
                             245.0 * powi(s0->tr, 4) + 25.0 * powi(m1->tr, 4) - 2.0 * powi(m2->tr, 4));
  out[0] = 0.0;
  out[1] = 0.0;
  out[2] = 0.0;
  out[3] = ee / 180.0;
  out[4] = ei / 180.0;
  out[5] = er / 180.0;
}

/* source, src/source.rs:7-21 with State::te/ti/tr (state.rs:118-135) */
void ro_source(const ro_physics* ph, const double s[6], double out[6]) {
  double rho = s[0];
  double u = s[1] / rho, v = s[2] / rho;
  double te = s[3] / (rho * ph->cve) - (powi(u, 2) + powi(v, 2)) / (6.0 * ph->cve);
  double er = s[5] / rho - (powi(u, 2) + powi(v, 2)) / 6.0;
  double tr = pow(er * rho / ph->a_rad, 0.25);
  double ti = s[4] / (rho * ph->cvi) - (powi(u, 2) + powi(v, 2)) / (6.0 * ph->cvi);
  out[0] = 0.0;
  out[1] = 0.0;
  out[2] = 0.0;
  out[3] = -ph->omega_ei * (te - ti) - ph->omega_er * (powi(te, 4) - powi(tr, 4));
  out[4] = ph->omega_ei * (te - ti);
  out[5] = ph->omega_er * (powi(te, 4) - powi(tr, 4));
}

/* get_local_dt, src/dt.rs:5-34 */
double ro_local_dt(const ro_physics* ph, const double s[6], double dx, double dy) {
  const double A = ph->a_rad, OEI = ph->omega_ei, OER = ph->omega_er, CVE = ph->cve, CVI = ph->cvi;
  double rho = s[0];
  double u = s[1] / rho, v = s[2] / rho;
  double ee = s[3] / rho - (powi(u, 2) + powi(v, 2)) / 6.0;
  double ei = s[4] / rho - (powi(u, 2) + powi(v, 2)) / 6.0;
  double er = s[5] / rho - (powi(u, 2) + powi(v, 2)) / 6.0;
  double gi = ph->gamma_i - 1.0, ge = ph->gamma_e - 1.0, gr = ph->gamma_r - 1.0;
  double cs = sqrt(ph->gamma_e * ge * ee + ph->gamma_i * gi * ei + ph->gamma_r * gr * er);
  double t = A * OEI * powi(CVE, 4) + A * OEI * powi(CVE, 3) * CVI + 4.0 * A * OER * CVI * powi(ee, 3) +
             OER * powi(CVE, 4) * CVI * rho;
  double s1 = sqrt(powi(t, 2) - 4.0 * A * OEI * OER * powi(CVE, 4) * CVI *
                                    (4.0 * A * powi(ee, 3) + powi(CVE, 4) * rho + powi(CVE, 3) * CVI * rho));
  double s2 = -(A * OEI * powi(CVE, 4) + A * OEI * powi(CVE, 3) * CVI + 4.0 * A * OER * CVI * powi(ee, 3) +
                OER * powi(CVE, 4) * CVI * rho);
  double sd = 2.0 * A * powi(CVE, 4) * CVI * rho;
  double alpha1 = fabs((-s1 + s2) / sd);
  double alpha2 = fabs((s1 + s2) / sd);
  double sj = fmax(alpha1, alpha2);
  double dj = fmax(fmax(ph->kappa_e / (CVE * rho), ph->kappa_i / (CVI * rho)), ph->kappa_r / A);
  double vj = (fabs(u) + cs) / dx + 2.0 * dj / (powi(dx, 2) + powi(dy, 2)) + sj + (fabs(v) + cs) / dy;
  return ph->lambda / vj;
}

/* -------------------------------------------------------------- bc1.rs */

/* build_local_plain, src/bc1.rs:294-312 */
static void build_local_plain(const ro_physics* ph, const double n[2], const double* val, double L[36],
                              double R[36], double lam[6]) {
  double mom_n = val[1] * n[0] + val[2] * n[1];
  double mom_t = -val[1] * n[1] + val[2] * n[0];
  double s1[6] = {val[0], mom_n, mom_t, val[3], val[4], val[5]};
  ro_derived d1;
  ro_derived_from_state(ph, s1, &d1);
  ro_build_l(ph, s1, s1, &d1, &d1, 0, L);
  ro_build_r(ph, s1, s1, &d1, &d1, 0, lam, R);
}

static const int kStOff[6] = {0, 1, 5, 14, 30, 55};
static const int kSOff[6] = {0, 1, 17, 98, 354, 979};
static const int kWOff[6] = {0, 1, 9, 36, 100, 225};

/* weno_extrapolation_pre, src/bc1.rs:193-255 */
static void weno_extrapolation_pre(const ro_physics* ph, const double L[36], const double n[2],
                                   const double* d5, const uint32_t* stencil, const double* S,
                                   const double* W, const double* u, int k, double result[6]) {
  double alpha[6][5], vks[6][5], alpha_sum[6] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r <= 4; ++r) {
    int m = (r + 1) * (r + 1);
    double data[6][25];
    for (int a = 0; a < m; ++a) {
      const double* s = u + 6 * (size_t)stencil[kStOff[r] + a];
      double mn = s[1] * n[0] + s[2] * n[1];
      double mt = -s[1] * n[1] + s[2] * n[0];
      double u6[6] = {s[0], mn, mt, s[3], s[4], s[5]};
      for (int c = 0; c < 6; ++c) {
        const double* row = L + 6 * c;
        data[c][a] = row[0] * u6[0] + row[1] * u6[1] + row[2] * u6[2] + row[3] * u6[3] + row[4] * u6[4] +
                     row[5] * u6[5];
      }
    }
    for (int c = 0; c < 6; ++c) {
      int soff = kSOff[r];
      double beta = 0.0;
      for (int a = 0; a < m; ++a) {
        int base = soff + a * m;
        double acc = 0.0;
        for (int b = 0; b < m; ++b) acc += S[base + b] * data[c][b];
        beta += data[c][a] * acc;
      }
      double vk = 0.0;
      if (k <= r) {
        int woff = kWOff[r] + k * m;
        for (int a = 0; a < m; ++a) vk += W[woff + a] * data[c][a];
      }
      alpha[c][r] = d5[r] / weno_pow(ph->eps + beta, ph->weno_q);
      vks[c][r] = vk;
      alpha_sum[c] += alpha[c][r];
    }
  }
  for (int c = 0; c < 6; ++c) {
    result[c] = 0.0;
    for (int r = 0; r <= 4; ++r) result[c] += (alpha[c][r] / alpha_sum[c]) * vks[c][r];
  }
}

/* Array2 . Array1 of the BC branches (ndarray -> BLAS dgemv), evaluated
 * left to right. */
static inline void matvec6(const double* M, const double* x, double* y) { mat6_dot(M, x, y); }

/* solve6 (LAPACK dgesv): LU with partial pivoting (first maximal pivot). */
static int solve6(const double* A0, const double* b0, double* x) {
  double A[36], b[6];
  memcpy(A, A0, sizeof(A));
  memcpy(b, b0, sizeof(b));
  for (int c = 0; c < 6; ++c) {
    int p = c;
    double best = fabs(A[c * 6 + c]);
    for (int r = c + 1; r < 6; ++r)
      if (fabs(A[r * 6 + c]) > best) {
        best = fabs(A[r * 6 + c]);
        p = r;
      }
    if (best == 0.0) return 8;
    if (p != c) {
      for (int q = 0; q < 6; ++q) {
        double t = A[c * 6 + q];
        A[c * 6 + q] = A[p * 6 + q];
        A[p * 6 + q] = t;
      }
      double t = b[c];
      b[c] = b[p];
      b[p] = t;
    }
    for (int r = c + 1; r < 6; ++r) {
      double f = A[r * 6 + c] / A[c * 6 + c];
      A[r * 6 + c] = f;
      for (int q = c + 1; q < 6; ++q) A[r * 6 + q] -= f * A[c * 6 + q];
      b[r] -= f * b[c];
    }
  }
  for (int r = 5; r >= 0; --r) {
    double s = b[r];
    for (int q = r + 1; q < 6; ++q) s -= A[r * 6 + q] * x[q];
    x[r] = s / A[r * 6 + r];
  }
  return 0;
}

/* Taylor expansion to the ghost point and rotation back, shared tail of the
 * Wall / Outflow / FarField branches (src/bc1.rs:728-755, 809-827, 896-923). */
static void taylor_rotate(const double* uk, double dist, const double n[2], double out[6]) {
  double ug[6];
  memcpy(ug, uk, sizeof(ug));
  double coef = 1.0;
  for (int k = 1; k <= 4; ++k) {
    coef *= dist / (double)k;
    for (int c = 0; c < 6; ++c) ug[c] = ug[c] + coef * uk[6 * k + c];
  }
  double mom_n = ug[1], mom_t = ug[2];
  out[0] = ug[0];
  out[1] = mom_n * n[0] - mom_t * n[1];
  out[2] = mom_n * n[1] + mom_t * n[0];
  out[3] = ug[3];
  out[4] = ug[4];
  out[5] = ug[5];
}

/* The heavy branches of set_ghost_point_value (src/bc1.rs:663-928):
 * L, R, lambda of the nearest state in the (n, t) frame, V^(k) of the
 * characteristic WENO extrapolation and the normal derivatives u^(k) of the
 * conservative state at the boundary point, before the Taylor expansion. */
int ro_heavy_terms(const ro_problem* pb, const ro_physics* ph, const double* u, int g, double L[36],
                   double R[36], double lam[6], double v[30], double uk[30]) {
  int kind = pb->g_kind[g];
  int h = pb->g_heavy[g];
  if (h < 0 || (kind != 1 && kind != 5 && kind != 7)) return 1;
  const double* n = pb->g_normal + 2 * g;
  build_local_plain(ph, n, u + 6 * (size_t)pb->h_nearest[h], L, R, lam);
  for (int k = 0; k < 5; ++k)
    weno_extrapolation_pre(ph, L, n, pb->h_d + 5 * (size_t)h, pb->h_stencil + 55 * (size_t)h,
                           pb->h_s + 979 * (size_t)h, pb->h_w + 225 * (size_t)h, u, k, v + 6 * k);
  if (kind == 7) { /* FarField, src/bc1.rs:829-924 */
    const double* ui = pb->g_bcstate + 6 * g;
    double mn = ui[1] * n[0] + ui[2] * n[1];
    double mt = -ui[1] * n[1] + ui[2] * n[0];
    double uloc[6] = {ui[0], mn, mt, ui[3], ui[4], ui[5]};
    double vinf[6];
    matvec6(L, uloc, vinf);
    double rhs[6];
    for (int m = 0; m < 6; ++m) rhs[m] = lam[m] < 0.0 ? vinf[m] : v[m];
    matvec6(R, rhs, uk);
    for (int k = 1; k <= 4; ++k) {
      for (int m = 0; m < 6; ++m) rhs[m] = lam[m] < 0.0 ? 0.0 : v[6 * k + m];
      matvec6(R, rhs, uk + 6 * k);
    }
  } else if (kind == 5) { /* Outflow-LODI, src/bc1.rs:760-828 */
    const double* of = pb->g_outflow + 3 * g;
    double p_inf = of[0], sigma = of[1], l_domain = of[2];
    matvec6(R, v, uk);
    double rho0 = uk[0];
    double un0 = uk[1] / rho0, ut0 = uk[2] / rho0;
    /* local_pressure_and_a2, src/bc1.rs:424-440 */
    double w2 = un0 * un0 + ut0 * ut0;
    double ee1 = uk[3] / rho0 - w2 / 6.0;
    double ei1 = uk[4] / rho0 - w2 / 6.0;
    double er1 = uk[5] / rho0 - w2 / 6.0;
    double ge = ph->gamma_e - 1.0, gi = ph->gamma_i - 1.0, gr = ph->gamma_r - 1.0;
    double p0 = rho0 * (ge * ee1 + gi * ei1 + gr * er1);
    double a2 = ph->gamma_e * ge * ee1 + ph->gamma_i * gi * ei1 + ph->gamma_r * gr * er1;
    double a = sqrt(a2);
    double lambda1 = un0 - a;
    double rhs1[6];
    memcpy(rhs1, v + 6, sizeof(rhs1));
    if (lambda1 < 0.0) {
      double m_local = fabs(un0 / a);
      double l1 = sigma * (1.0 - powi(m_local, 2)) * a / l_domain * (p0 - p_inf);
      rhs1[0] = l1 / lambda1;
    }
    matvec6(R, rhs1, uk + 6);
    for (int k = 2; k <= 4; ++k) matvec6(R, v + 6 * k, uk + 6 * k);
  } else { /* Wall-ILW, src/bc1.rs:691-756 */
    double L0[36], rhs0[6];
    memcpy(L0, L, sizeof(L0));
    for (int c = 0; c < 6; ++c) L0[c] = c == 1 ? 1.0 : 0.0;
    memcpy(rhs0, v, sizeof(rhs0));
    rhs0[0] = 0.0;
    if (solve6(L0, rhs0, uk)) return 8;
    /* build_ilw_row1, src/bc1.rs:448-465 */
    double u1 = uk[0], u3 = uk[2];
    double gi = ph->gamma_i - 1.0, ge = ph->gamma_e - 1.0, gr = ph->gamma_r - 1.0;
    double gt = gi + ge + gr;
    double L1[36], rhs1[6];
    memcpy(L1, L, sizeof(L1));
    L1[0] = -gt * powi(u3, 2) / (6.0 * powi(u1, 2));
    L1[1] = 0.0;
    L1[2] = gt * u3 / (6.0 * u1);
    L1[3] = -ge;
    L1[4] = -gi;
    L1[5] = -gr;
    memcpy(rhs1, v + 6, sizeof(rhs1));
    rhs1[0] = 0.0;
    if (solve6(L1, rhs1, uk + 6)) return 8;
    for (int k = 2; k <= 4; ++k) matvec6(R, v + 6 * k, uk + 6 * k);
  }
  return 0;
}

/* set_ghost_point_value, src/bc1.rs:609-928 (precomputed fast path) */
static int ghost_value(const ro_problem* pb, const ro_physics* ph, const double* u, int g, double out[6]) {
  int kind = pb->g_kind[g];
  const double* n = pb->g_normal + 2 * g;
  if (kind == 3) { /* Constant */
    memcpy(out, pb->g_bcstate + 6 * g, sizeof(double) * 6);
    return 0;
  }
  if (kind == 0) { /* ReflectiveWall, src/bc1.rs:930-1037 */
    const double* s = u + 6 * (size_t)pb->g_src[g];
    double mom_n = s[1] * n[0] + s[2] * n[1];
    out[0] = s[0];
    out[1] = s[1] - 2.0 * mom_n * n[0];
    out[2] = s[2] - 2.0 * mom_n * n[1];
    out[3] = s[3];
    out[4] = s[4];
    out[5] = s[5];
    return 0;
  }
  if (kind == 6 || kind == 2) { /* ZerothOrder / Periodic: copy */
    memcpy(out, u + 6 * (size_t)pb->g_src[g], sizeof(double) * 6);
    return 0;
  }
  if (kind == 4) { /* TimeDependent, src/bc1.rs:634-636: the moving front
                      closure of src/init.rs:61-65 at the state's time */
    const double* fr = pb->g_front + 2 * g;
    double x_front = fr[1] + fr[0] * pb->time;
    const double* st = pb->g_p0x[g] <= x_front ? pb->g_bcstate + 6 * g : pb->g_bcstate2 + 6 * g;
    memcpy(out, st, sizeof(double) * 6);
    return 0;
  }
  if (kind != 1 && kind != 5 && kind != 7) {
    memset(out, 0, sizeof(double) * 6);
    return 0;
  }
  double L[36], R[36], lam[6], v[30], uk[30];
  int st = ro_heavy_terms(pb, ph, u, g, L, R, lam, v, uk);
  if (st) return st;
  taylor_rotate(uk, pb->g_dist[g], n, out);
  return 0;
}

int ro_update_ghosts(const ro_problem* pb, const ro_physics* ph, const double* u, double* gv,
                     int nthreads, ro_error* err) {
  int status = 0;
  int first = -1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int g = 0; g < pb->n_ghost; ++g) {
    int st = ghost_value(pb, ph, u, g, gv + 6 * (size_t)g);
    int bad = st != 0;
    if (!bad)
      for (int c = 0; c < 6; ++c)
        if (!isfinite(gv[6 * (size_t)g + c])) bad = 1;
    if (bad) {
#ifdef _OPENMP
#pragma omp critical
#endif
      {
        if (first < 0 || g < first) {
          first = g;
          status = st ? st : 4;
        }
      }
    }
  }
  if (first >= 0 && err) {
    err->kind = status;
    err->stage = 0;
    err->ghost_id = first;
    err->i = err->j = -1;
    memcpy(err->state, gv + 6 * (size_t)first, sizeof(err->state));
  }
  return status;
}

/* ------------------------------------------------------------- main.rs */

/* GhostGrid::k_for with the sorted radius-4 cross (src/ghost.rs:141-146) */
static inline int k_for_x(int d) { return d < 0 ? d + 4 : (d == 0 ? 8 : 12 + d); }
static inline int k_for_y(int d) { return 8 + d; }

/* l(), src/main.rs:135-305 */
int ro_l(const ro_problem* pb, const ro_physics* ph, const double* u, double* rhs, double* ghost_out,
         int nthreads, ro_error* err) {
  const int nx = pb->nx, ny = pb->ny;
  const size_t ncell = (size_t)nx * ny;
  const double dx = pb->dx, dy = pb->dy;
  int nt = nthreads > 0 ? nthreads : 1;
  (void)nt;
  double* gv = ghost_out ? ghost_out : (double*)malloc(sizeof(double) * 6 * (size_t)(pb->n_ghost > 0 ? pb->n_ghost : 1));
  ro_derived* gd = (ro_derived*)malloc(sizeof(ro_derived) * (size_t)(pb->n_ghost > 0 ? pb->n_ghost : 1));
  ro_derived* der = (ro_derived*)calloc(ncell, sizeof(ro_derived));
  double* fx = (double*)calloc((size_t)(nx + 1) * ny * 6, sizeof(double));
  double* fy = (double*)calloc((size_t)nx * (ny + 1) * 6, sizeof(double));
  int diffusion = ph->kappa_e != 0.0 || ph->kappa_i != 0.0 || ph->kappa_r != 0.0;
  int source = ph->omega_ei != 0.0 || ph->omega_er != 0.0;
  double* dfx = diffusion ? (double*)calloc((size_t)(nx + 1) * ny * 6, sizeof(double)) : NULL;
  double* dfy = diffusion ? (double*)calloc((size_t)nx * (ny + 1) * 6, sizeof(double)) : NULL;

  int st = ro_update_ghosts(pb, ph, u, gv, nthreads, err);
  if (st) goto done;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt)
#endif
  for (int g = 0; g < pb->n_ghost; ++g) ro_derived_from_state(ph, gv + 6 * (size_t)g, &gd[g]);

  /* derived, main.rs:149-153 */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt)
#endif
  for (long l = 0; l < (long)ncell; ++l)
    if (pb->fluid[l]) ro_derived_from_state(ph, u + 6 * (size_t)l, &der[l]);

#define VAL(t) ((size_t)(t) < ncell ? u + 6 * (size_t)(t) : gv + 6 * ((size_t)(t) - ncell))
#define DER(t) ((size_t)(t) < ncell ? &der[(t)] : &gd[(size_t)(t) - ncell])

  /* x fluxes, main.rs:165-181 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (long lin = 0; lin < (long)(nx + 1) * ny; ++lin) {
    long i = lin / ny;
    long anchor;
    int d0;
    if (i < nx && pb->fluid[lin]) {
      anchor = lin;
      d0 = -3;
    } else if (i >= 1 && pb->fluid[lin - ny]) {
      anchor = lin - ny;
      d0 = -2;
    } else {
      continue; /* zero (calloc) */
    }
    double pts[36];
    ro_derived dd[6];
    for (int q = 0; q < 6; ++q) {
      uint32_t t = pb->access[(size_t)anchor * 17 + k_for_x(d0 + q)];
      memcpy(pts + 6 * q, VAL(t), sizeof(double) * 6);
      dd[q] = *DER(t);
    }
    ro_reconstruction_fast(ph, pts, dd, 0, 1, fx + 6 * (size_t)lin);
    if (diffusion) ro_diffusion(ph, dd, dfx + 6 * (size_t)lin);
  }
  /* y fluxes, main.rs:186-204 (layout j*nx + i) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (long lin = 0; lin < (long)nx * (ny + 1); ++lin) {
    long j = lin / nx, i = lin % nx;
    long cell = i * ny + j;
    long anchor;
    int d0;
    if (j < ny && pb->fluid[cell]) {
      anchor = cell;
      d0 = -3;
    } else if (j >= 1 && pb->fluid[cell - 1]) {
      anchor = cell - 1;
      d0 = -2;
    } else {
      continue;
    }
    double pts[36];
    ro_derived dd[6];
    for (int q = 0; q < 6; ++q) {
      uint32_t t = pb->access[(size_t)anchor * 17 + k_for_y(d0 + q)];
      memcpy(pts + 6 * q, VAL(t), sizeof(double) * 6);
      dd[q] = *DER(t);
    }
    ro_reconstruction_fast(ph, pts, dd, 1, 1, fy + 6 * (size_t)lin);
    if (diffusion) ro_diffusion(ph, dd, dfy + 6 * (size_t)lin);
  }

  /* rhs, main.rs:251-304 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (long lin = 0; lin < (long)ncell; ++lin) {
    double* out = rhs + 6 * (size_t)lin;
    if (!pb->fluid[lin]) {
      for (int c = 0; c < 6; ++c) out[c] = 0.0;
      continue;
    }
    long i = lin / ny, j = lin % ny;
    const double* fl = fx + 6 * (size_t)lin;
    const double* fr = fx + 6 * (size_t)(lin + ny);
    const double* fb = fy + 6 * (size_t)(j * nx + i);
    const double* ft = fy + 6 * (size_t)((j + 1) * nx + i);
    double inv_dx = 1.0 / dx, inv_dy = 1.0 / dy;
    double fxt[6], fyt[6];
    for (int c = 0; c < 6; ++c) {
      fxt[c] = -(fr[c] - fl[c]) * inv_dx;
      fyt[c] = -(ft[c] - fb[c]) * inv_dy;
    }
    ro_derived d9x[9], d9y[9];
    for (int q = 0; q < 9; ++q) {
      d9x[q] = *DER(pb->access[(size_t)lin * 17 + k_for_x(q - 4)]);
      d9y[q] = *DER(pb->access[(size_t)lin * 17 + k_for_y(q - 4)]);
    }
    double ncx[6], ncy[6];
    ro_noncon(d9x, dx, 0, ncx);
    ro_noncon(d9y, dy, 1, ncy);
    double src[6] = {0, 0, 0, 0, 0, 0};
    if (source) ro_source(ph, u + 6 * (size_t)lin, src);
    double dif[6] = {0, 0, 0, 0, 0, 0};
    if (diffusion) {
      double cx = -1.0 / (dx * dx), cy = -1.0 / (dy * dy);
      const double* dl = dfx + 6 * (size_t)lin;
      const double* dr = dfx + 6 * (size_t)(lin + ny);
      const double* db = dfy + 6 * (size_t)(j * nx + i);
      const double* dt_ = dfy + 6 * (size_t)((j + 1) * nx + i);
      for (int c = 0; c < 6; ++c) dif[c] = (-(dr[c] - dl[c]) * cx) + (-(dt_[c] - db[c]) * cy);
    }
    for (int c = 0; c < 6; ++c) out[c] = fxt[c] + fyt[c] + ncx[c] + ncy[c] + src[c] + dif[c];
  }
#undef VAL
#undef DER
done:
  if (!ghost_out) free(gv);
  free(gd);
  free(der);
  free(fx);
  free(fy);
  free(dfx);
  free(dfy);
  return st;
}

/* assert_admissible, src/main.rs:22-33 */
static int admissible(const double s[6]) {
  for (int c = 0; c < 6; ++c)
    if (!isfinite(s[c])) return 0;
  if (s[0] <= 0.0) return 0;
  double ux = s[1] / s[0], uy = s[2] / s[0];
  double k = (ux * ux + uy * uy) / 6.0;
  if (s[3] / s[0] - k <= 0.0 || s[4] / s[0] - k <= 0.0 || s[5] / s[0] - k <= 0.0) return 0;
  return 1;
}

/* stage_update_rhs / stage_update_comb, src/main.rs:307-349.  The first
 * failing cell is the smallest linear index. */
static int stage(const ro_problem* pb, const double* base, const double* add, double* dst, const double* rhs,
                 double coef, double wb, double wa, int label, int nt, ro_error* err) {
  const size_t ncell = (size_t)pb->nx * pb->ny;
  long first = -1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (long l = 0; l < (long)ncell; ++l) {
    if (!pb->fluid[l]) continue;
    double v[6];
    for (int c = 0; c < 6; ++c) {
      if (!add)
        v[c] = base[6 * l + c] + rhs[6 * l + c] * coef;
      else
        v[c] = base[6 * l + c] * wb + add[6 * l + c] * wa + rhs[6 * l + c] * coef;
    }
    if (!admissible(v)) {
#ifdef _OPENMP
#pragma omp critical
#endif
      {
        if (first < 0 || l < first) first = l;
      }
    }
    memcpy(dst + 6 * l, v, sizeof(v));
  }
  if (first >= 0) {
    if (err) {
      err->kind = 3;
      err->stage = label;
      err->i = (int)(first / pb->ny);
      err->j = (int)(first % pb->ny);
      err->ghost_id = -1;
      memcpy(err->state, dst + 6 * first, sizeof(err->state));
    }
    return 3;
  }
  return 0;
}

/* rk3_ssp, src/main.rs:351-373 */
int ro_rk3_step(const ro_problem* pb, const ro_physics* ph, double* u, double* time, double dt, int nthreads,
                ro_error* err) {
  const size_t n = (size_t)pb->nx * pb->ny * 6;
  int nt = nthreads > 0 ? nthreads : 1;
  double* u1 = (double*)calloc(n, sizeof(double));
  double* u2 = (double*)calloc(n, sizeof(double));
  double* r = (double*)malloc(n * sizeof(double));
  /* the three l() calls see the stage states' Field.time: u.time,
   * u1.time = u.time + dt, u2.time = u.time + 0.5 * dt (src/main.rs:356-366) */
  const double t0 = time ? *time : pb->time;
  ro_problem ps = *pb;
  ps.time = t0;
  int st = ro_l(&ps, ph, u, r, NULL, nt, err);
  if (!st) st = stage(pb, u, NULL, u1, r, dt, 0, 0, 1, nt, err);
  ps.time = t0 + dt;
  if (!st) st = ro_l(&ps, ph, u1, r, NULL, nt, err);
  if (!st) st = stage(pb, u, u1, u2, r, dt / 4.0, 0.75, 0.25, 2, nt, err);
  ps.time = t0 + 0.5 * dt;
  if (!st) st = ro_l(&ps, ph, u2, r, NULL, nt, err);
  if (!st) st = stage(pb, u, u2, u1, r, 2.0 * dt / 3.0, 1.0 / 3.0, 2.0 / 3.0, 3, nt, err);
  if (!st) {
    const size_t ncell = (size_t)pb->nx * pb->ny;
    for (size_t l = 0; l < ncell; ++l)
      if (pb->fluid[l]) memcpy(u + 6 * l, u1 + 6 * l, sizeof(double) * 6);
    if (time) *time = *time + dt;
  }
  free(u1);
  free(u2);
  free(r);
  return st;
}

/* calc_global_dt, src/main.rs:375-408 */
double ro_global_dt(const ro_problem* pb, const ro_physics* ph, const double* u) {
  double g = INFINITY;
  const size_t ncell = (size_t)pb->nx * pb->ny;
  for (size_t l = 0; l < ncell; ++l) {
    if (!pb->fluid[l]) continue;
    g = fmin(g, ro_local_dt(ph, u + 6 * l, pb->dx, pb->dy));
  }
  return g;
}
