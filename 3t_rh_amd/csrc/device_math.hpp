// Pointwise FP64 numerics of the l() path for gfx950, written against the
// reference's formulas (colinhu123/3T_RH src/state.rs, src/weno.rs,
// src/noncon.rs, src/diffusion.rs, src/source.rs, src/dt.rs).  Expressions
// keep the reference's association and literal divisions; the translation
// unit is compiled with -ffp-contract=off so products and sums round as the
// reference's do.  Structural zeros and ones of the 6x6 eigenvector
// matrices are skipped (x*0 + y == y and x*1 == x exactly, signed zeros
// aside), everything else is evaluated as written.
#pragma once

#include <hip/hip_runtime.h>

namespace rh3t_dev {

// A pointer read from a descriptor in device memory is generic to the
// compiler, and loads through it are flat loads (a 64-bit VGPR address per
// lane, counted on both the vector and the scalar/LDS wait counters).  The
// planes all live in global memory: passing the pointer through an opaque
// step as a global-address-space pointer (a plain cast round trip would be
// folded away) lets the loads and stores through it be global ones, with
// the base in SGPRs.  p must be uniform over the wavefront (it is held in
// SGPRs).
template <class T>
__device__ __forceinline__ T* glb(T* p) {
  using GT = __attribute__((address_space(1))) T;
  GT* q = (GT*)p;
  __asm__("" : "+s"(q));
  return (T*)q;
}


struct Phys {
  double gamma_e, gamma_i, gamma_r;
  double ge, gi, gr;  // gamma_k - 1.0
  double gt;          // gi + ge + gr (reference order)
  double kappa_e, kappa_i, kappa_r;
  double cve, cvi, a_rad;
  double omega_ei, omega_er;
  double lambda, weno_q, eps;
  int diffusion, source;
};

// Derived planes of the padded derived array.
enum { D_U = 0, D_V, D_EE, D_EI, D_ER, D_PE, D_PI, D_PR, D_TE, D_TI, D_TR, D_N };

__device__ __forceinline__ double sq(double x) { return x * x; }

// a / b rounded to nearest, from y = RN(1/b) (a reciprocal shared by several
// quotients, or a constant): q0 = RN(a y) lies within 1.5 ulp of a / b; the
// FMA correction q' = RN(q + (a - b q) y), whose residual a - b q is exact,
// brings q0 within one ulp, and applied once more returns RN(a / b)
// (Markstein's theorem: y within half an ulp of 1/b and q within one ulp of
// a / b give a correctly rounded q').  It therefore equals the IEEE quotient
// bit for bit whenever a, b and a / b are finite and the residuals stay in
// the normal range (|a| >= 2^-969 or a = 0; a = -0 gives +0): 5 FP64 FMA /
// multiply instructions instead of the 11 of a division with its
// quarter-rate reciprocal.
__host__ __device__ __forceinline__ double div_rn(double a, double b, double y) {
  const double q0 = a * y;
  const double q1 = fma(fma(-b, q0, a), y, q0);
  return fma(fma(-b, q1, a), y, q1);
}
constexpr double kRcp3 = 1.0 / 3.0;  // RN(1/3)
constexpr double kRcp6 = 1.0 / 6.0;  // RN(1/6)
// x.powi(3) = x * (x*x), x.powi(4) = (x*x)*(x*x)
__device__ __forceinline__ double cube(double x) { return x * (x * x); }
__device__ __forceinline__ double pow4(double x) { double x2 = x * x; return x2 * x2; }

// Derived::from_state, src/state.rs:241-275.  te/ti/tr only when asked.
__device__ __forceinline__ void derived_from_state(const Phys& ph, const double s[6], double d[D_N],
                                                   bool temps) {
  double rho = s[0];
  double u = s[1] / rho;
  double v = s[2] / rho;
  double w2 = sq(u) + sq(v);
  d[D_U] = u;
  d[D_V] = v;
  d[D_EE] = s[3] / rho - w2 / 6.0;
  d[D_EI] = s[4] / rho - w2 / 6.0;
  d[D_ER] = s[5] / rho - w2 / 6.0;
  d[D_PE] = ph.ge * (s[3] - rho * w2 / 6.0);
  d[D_PI] = ph.gi * (s[4] - rho * w2 / 6.0);
  d[D_PR] = ph.gr * (s[5] - rho * w2 / 6.0);
  if (temps) {
    d[D_TE] = s[3] / (rho * ph.cve) - w2 / (6.0 * ph.cve);
    d[D_TI] = s[4] / (rho * ph.cvi) - w2 / (6.0 * ph.cvi);
    double er_term = s[5] / rho - w2 / 6.0;
    d[D_TR] = pow(er_term * rho / ph.a_rad, 0.25);
  }
}

// weno5, src/weno.rs:846-886
__device__ __forceinline__ double weno5(double u0, double u1, double u2, double u3, double u4,
                                        double eps) {
  double beta2 = 13.0 / 12.0 * sq(u2 - 2.0 * u3 + u4) + 0.25 * sq(3.0 * u2 - 4.0 * u3 + u4);
  double beta1 = 13.0 / 12.0 * sq(u1 - 2.0 * u2 + u3) + 0.25 * sq(u1 - u3);
  double beta0 = 13.0 / 12.0 * sq(u0 - 2.0 * u1 + u2) + 0.25 * sq(u0 - 4.0 * u1 + 3.0 * u2);
  double a0 = 0.1 / sq(eps + beta0);
  double a1 = 0.6 / sq(eps + beta1);
  double a2 = 0.3 / sq(eps + beta2);
  double sum_o = a0 + a1 + a2;
  // the quotients by sum_o, 3 and 6 through their reciprocals (div_rn:
  // the same roundings as the literal divisions)
  const double ys = 1.0 / sum_o;
  double w0 = div_rn(a0, sum_o, ys), w1 = div_rn(a1, sum_o, ys), w2 = div_rn(a2, sum_o, ys);
  double p0 = div_rn(u0, 3.0, kRcp3) - 7.0 / 6.0 * u1 + 11.0 / 6.0 * u2;
  double p1 = div_rn(-u1, 6.0, kRcp6) + 5.0 / 6.0 * u2 + div_rn(u3, 3.0, kRcp3);
  double p2 = div_rn(u2, 3.0, kRcp3) + 5.0 / 6.0 * u3 - div_rn(u4, 6.0, kRcp6);
  return w0 * p0 + w1 * p1 + w2 * p2;
}

// Energy-mode row of L, src/weno.rs:110-233 (identical in X and Y).
// The quotients by 6 cs2, 3 cs2, gt cs2 and b go through div_rn with the
// reciprocals y6 = RN(1/(6 cs2)) (RN(1/(3 cs2)) = 2 y6 exactly), ygt =
// RN(1/(gt cs2)) and yb = RN(1/b): the same roundings as the divisions.
__device__ __forceinline__ void energy_row(int mode, double hk, double u, double v, double w2,
                                           double cs2, double gt, double b, double ge, double gi,
                                           double gr, double y6, double ygt, double yb, double row[6]) {
  double gk = mode == 0 ? ge : (mode == 1 ? gi : gr);
  const double c6 = 6.0 * cs2, c3 = 3.0 * cs2, y3 = 2.0 * y6, gtc = gt * cs2;
  row[0] = gk / gt - div_rn(gk * w2, c6, y6) - div_rn(gk * gt * w2, b, yb) * (hk - cs2 * w2);
  row[1] = div_rn(gk * u, c3, y3) + div_rn(2.0 * gk * gt * u, b, yb) * (hk - cs2 * w2);
  row[2] = div_rn(gk * v, c3, y3) + div_rn(2.0 * gk * gt * v, b, yb) * (hk - cs2 * w2);
  if (mode == 0) {
    row[3] = div_rn(-sq(ge), gtc, ygt) - div_rn(6.0 * ge, b, yb) * (ge * hk - gt * cs2 * w2);
    row[4] = div_rn(-ge * gi, gtc, ygt) - div_rn(6.0 * ge * gi, b, yb) * (hk - 6.0 * gr * cs2);
    row[5] = div_rn(-ge * gr, gtc, ygt) - div_rn(6.0 * ge * gr, b, yb) * (hk + 6.0 * gi * cs2);
  } else if (mode == 1) {
    row[3] = div_rn(-gi * ge, gtc, ygt) - div_rn(6.0 * gi * ge, b, yb) * (hk + 6.0 * gr * cs2);
    row[4] = div_rn(-sq(gi), gtc, ygt) - div_rn(6.0 * gi, b, yb) * (gi * hk - gt * cs2 * w2);
    row[5] = div_rn(-gi * gr, gtc, ygt) - div_rn(6.0 * gi * gr, b, yb) * (hk - 6.0 * ge * cs2);
  } else {
    row[3] = div_rn(-gr * ge, gtc, ygt) - div_rn(6.0 * gr * ge, b, yb) * (hk - 6.0 * gi * cs2);
    row[4] = div_rn(-gr * gi, gtc, ygt) - div_rn(6.0 * gr * gi, b, yb) * (hk + 6.0 * ge * cs2);
    row[5] = div_rn(-sq(gr), gtc, ygt) - div_rn(6.0 * gr, b, yb) * (gr * hk - gt * cs2 * w2);
  }
}

// Roe state and the closed-form left/right eigenvectors of the X direction
// of Stencil6::build_l_plain / build_r_plain (src/weno.rs:28-548), entry by
// entry, in the reference's expressions: roe_x forms the Roe state of points
// [3] and [2] (densities rho3 / rho2, their u, v, e_e, e_i, e_r), lx_row a
// row of L (u-c, electron, ion, radiation, transverse, u+c), rx_entry an
// entry of R, lamx an eigenvalue.
struct RoeX {
  double u, v, w2, ee, ei, er, cs2, cs, h[3], b, ad;
  double y6, yb, ygt;  // RN(1/(6 cs2)), RN(1/b), RN(1/(gt cs2)) for div_rn
};

__device__ __forceinline__ void roe_x(const Phys& ph, double rho3, double rho2, double u3, double v3, double ee3,
                                      double ei3, double er3, double u2, double v2, double ee2, double ei2,
                                      double er2, RoeX& Q) {
  double r1 = sqrt(rho3);
  double r2 = sqrt(rho2);
  double rs = r1 + r2;
  const double yrs = 1.0 / rs;  // the five divisions by rs through div_rn
  Q.u = div_rn(u3 * r1 + u2 * r2, rs, yrs);
  Q.v = div_rn(v3 * r1 + v2 * r2, rs, yrs);
  Q.w2 = sq(Q.u) + sq(Q.v);
  Q.ee = div_rn(ee3 * r1 + ee2 * r2, rs, yrs);
  Q.ei = div_rn(ei3 * r1 + ei2 * r2, rs, yrs);
  Q.er = div_rn(er3 * r1 + er2 * r2, rs, yrs);
  const double ge = ph.ge, gi = ph.gi, gr = ph.gr, gt = ph.gt;
  Q.cs2 = ph.gamma_e * ge * Q.ee + ph.gamma_i * gi * Q.ei + ph.gamma_r * gr * Q.er;
  Q.cs = sqrt(Q.cs2);
  Q.h[0] = 6.0 * gi * gr * (ph.gamma_i * Q.ei - ph.gamma_r * Q.er) + ph.gamma_e * gt * Q.ee * Q.w2;
  Q.h[1] = 6.0 * ge * gr * (ph.gamma_r * Q.er - ph.gamma_e * Q.ee) + ph.gamma_i * gt * Q.ei * Q.w2;
  Q.h[2] = 6.0 * ge * gi * (ph.gamma_e * Q.ee - ph.gamma_i * Q.ei) + ph.gamma_r * gt * Q.er * Q.w2;
  Q.b = gt * (36.0 * ge * gi * gr + gt * sq(Q.w2)) * Q.cs2;
  Q.ad = 12.0 * Q.cs2;
  Q.y6 = 1.0 / (6.0 * Q.cs2);
  Q.yb = 1.0 / Q.b;
  Q.ygt = 1.0 / (gt * Q.cs2);
}

// row c of L (X rows: u-c, electron, ion, radiation, transverse, u+c)
__device__ __forceinline__ void lx_row(const Phys& ph, const RoeX& Q, int c, double row[6]) {
  const double gt = ph.gt;
  if (c == 0 || c == 5) {
    // ad = 12 cs2 = 2 (6 cs2) exactly, so RN(1/ad) = y6 / 2
    const double yad = 0.5 * Q.y6;
    if (c == 0) {
      row[0] = div_rn(gt * Q.w2 + 6.0 * Q.u * Q.cs, Q.ad, yad);
      row[1] = div_rn(-2.0 * gt * Q.u - 6.0 * Q.cs, Q.ad, yad);
    } else {
      row[0] = div_rn(gt * Q.w2 - 6.0 * Q.u * Q.cs, Q.ad, yad);
      row[1] = div_rn(-2.0 * gt * Q.u + 6.0 * Q.cs, Q.ad, yad);
    }
    row[2] = div_rn(-2.0 * gt * Q.v, Q.ad, yad);
    row[3] = div_rn(6.0 * ph.ge, Q.ad, yad);
    row[4] = div_rn(6.0 * ph.gi, Q.ad, yad);
    row[5] = div_rn(6.0 * ph.gr, Q.ad, yad);
  } else if (c == 4) {
    row[0] = -Q.v;
    row[1] = 0.0;
    row[2] = 1.0;
    row[3] = row[4] = row[5] = 0.0;
  } else {
    energy_row(c - 1, Q.h[c - 1], Q.u, Q.v, Q.w2, Q.cs2, gt, Q.b, ph.ge, ph.gi, ph.gr, Q.y6, Q.ygt, Q.yb, row);
  }
}

// entry (a, c) of R (X)
__device__ __forceinline__ double rx_entry(const Phys& ph, const RoeX& Q, int a, int c) {
  const double u = Q.u, v = Q.v, cs = Q.cs, w2 = Q.w2, gt = ph.gt;
  if (a == 0) return c == 4 ? 0.0 : 1.0;
  if (a == 1) return c == 0 ? u - cs : (c == 5 ? u + cs : (c == 4 ? 0.0 : u));
  if (a == 2) return c == 4 ? 1.0 : v;
  const int k = a - 3;  // energy row e, i, r
  if (c == 4) return div_rn(v, 3.0, kRcp3);
  if (c == 0 || c == 5) {
    const double gam = k == 0 ? ph.gamma_e : (k == 1 ? ph.gamma_i : ph.gamma_r);
    const double ek = k == 0 ? Q.ee : (k == 1 ? Q.ei : Q.er);
    const double g_ = gam * ek + div_rn(w2, 6.0, kRcp6);
    const double ucs = div_rn(u * cs, 3.0, kRcp3);
    return c == 0 ? g_ - ucs : g_ + ucs;
  }
  const int j = c - 1;  // energy column
  const double g[3] = {ph.ge, ph.gi, ph.gr};
  if (j == k) return gt * w2 / (6.0 * g[k]);
  // off-diagonal: rows (e, i, r) x columns (e, i, r) = [., -gr, gi; gr, ., -ge; -gi, ge, .]
  if (k == 0) return j == 1 ? -ph.gr : ph.gi;
  if (k == 1) return j == 0 ? ph.gr : -ph.ge;
  return j == 0 ? -ph.gi : ph.ge;
}

// Row i of R and row c of L for an index that differs between lanes: one
// compile-time case per index (a select between constant fields would
// otherwise become an indexed load and force the constants into scratch).
__device__ __forceinline__ void rx_row_rt(const Phys& ph, const RoeX& Q, int i, double r[6]) {
  switch (i) {
#define RH3T_RX_ROW(a) \
  case a:              \
    for (int c = 0; c < 6; ++c) r[c] = rx_entry(ph, Q, a, c); \
    break;
    RH3T_RX_ROW(0)
    RH3T_RX_ROW(1)
    RH3T_RX_ROW(2)
    RH3T_RX_ROW(3)
    RH3T_RX_ROW(4)
    default:
      for (int c = 0; c < 6; ++c) r[c] = rx_entry(ph, Q, 5, c);
#undef RH3T_RX_ROW
  }
}

__device__ __forceinline__ void lx_row_rt(const Phys& ph, const RoeX& Q, int c, double row[6]) {
  switch (c) {
    case 0: lx_row(ph, Q, 0, row); break;
    case 1: lx_row(ph, Q, 1, row); break;
    case 2: lx_row(ph, Q, 2, row); break;
    case 3: lx_row(ph, Q, 3, row); break;
    case 4: lx_row(ph, Q, 4, row); break;
    default: lx_row(ph, Q, 5, row);
  }
}

__device__ __forceinline__ double lamx(const RoeX& Q, int c) {
  return c == 0 ? Q.u - Q.cs : (c == 5 ? Q.u + Q.cs : Q.u);
}

// l_dot / r_dot row, src/weno.rs:823-843
__device__ __forceinline__ double dot6(const double r[6], const double x[6]) {
  return r[0] * x[0] + r[1] * x[1] + r[2] * x[2] + r[3] * x[3] + r[4] * x[4] + r[5] * x[5];
}

// q_pulse, src/noncon.rs:49-57
__device__ __forceinline__ double q_pulse(double b0, double b1, double b2, double b3, double b4,
                                          double b5, double b6) {
  return -0.00714285714285723 * b0 + 0.0595238095238118 * b1 - 0.240476190476197 * b2 +
         0.759523809523819 * b3 + 0.509523809523803 * b4 - 0.0904761904761891 * b5 +
         0.00952380952380977 * b6;
}

// jump_component of n_jump_pre, src/noncon.rs:259-285, on a[k..k+8)
__device__ __forceinline__ double jump(const double* a) {
  return q_pulse(a[7], a[6], a[5], a[4], a[3], a[2], a[1]) -
         q_pulse(a[0], a[1], a[2], a[3], a[4], a[5], a[6]);
}

// sixth_order_derivative, src/noncon.rs:26-44; d60 = 60 dx, y60 = RN(1/d60)
__device__ __forceinline__ double sixth(double qm3, double qm2, double qm1, double qp1, double qp2,
                                        double qp3, double d60, double y60) {
  return div_rn(qp3 - 9.0 * qp2 + 45.0 * qp1 - 45.0 * qm1 + 9.0 * qm2 - qm3, d60, y60);
}

// nonconservative_direction_pre, src/noncon.rs:288-332.  vel[k], pe/pi/pr[k]
// at the 9 points d-4..d+4; writes the ee, ei, er components.  yds =
// RN(1/ds) and y60 = RN(1/(60 ds)) (the divisions by ds and 60 ds through
// div_rn, which rounds as the literal divisions do).
__device__ __forceinline__ void noncon(const double vel[9], const double pe[9], const double pi[9],
                                       const double pr[9], double ds, double yds, double y60, double out[3]) {
  double v = vel[4];
  const double d60 = 60.0 * ds;
  double dpe = sixth(pe[1], pe[2], pe[3], pe[5], pe[6], pe[7], d60, y60);
  double dpi = sixth(pi[1], pi[2], pi[3], pi[5], pi[6], pi[7], d60, y60);
  double dpr = sixth(pr[1], pr[2], pr[3], pr[5], pr[6], pr[7], d60, y60);
  double der_ee = 2.0 * dpe - dpi - dpr;
  double der_ei = 2.0 * dpi - dpe - dpr;
  double der_er = 2.0 * dpr - dpe - dpi;
  double nee[9], nei[9], ner[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    nee[k] = 2.0 * pe[k] - pi[k] - pr[k];
    nei[k] = 2.0 * pi[k] - pe[k] - pr[k];
    ner[k] = 2.0 * pr[k] - pe[k] - pi[k];
  }
  double cl = div_rn(fmax(fmax(v, vel[3]), 0.0), 3.0, kRcp3);
  double cr = div_rn(fmin(fmin(v, vel[5]), 0.0), 3.0, kRcp3);
  const double v3 = div_rn(v, 3.0, kRcp3);
  out[0] = v3 * der_ee + div_rn(cl * jump(nee) + cr * jump(nee + 1), ds, yds);
  out[1] = v3 * der_ei + div_rn(cl * jump(nei) + cr * jump(nei + 1), ds, yds);
  out[2] = v3 * der_er + div_rn(cl * jump(ner) + cr * jump(ner + 1), ds, yds);
}

// noncon in FAST arithmetic, on the linear structure of its terms (the same
// algebra; the divisions by 3, ds and 60 ds as products with 1/3 and the
// caller's ids = 1/ds).
// * jump(a) = q_pulse(a7..a1) - q_pulse(a0..a6) = sum_k J_k a_k with
//   antisymmetric weights J_{7-k} = -J_k: J_0 = -Q_0, J_1 = Q_6 - Q_1,
//   J_2 = Q_5 - Q_2, J_3 = Q_4 - Q_3 (Q the q_pulse coefficients), so the
//   upwind term cl jump(x) + cr jump(x + 1) of a field x is one combination
//   sum_m W_m x_m over the 9 points, W_m = cl J_m + cr J_{m-1} (J_{-1} =
//   J_8 = 0), formed once per cell and line.
// * jump and the sixth-order derivative are linear, so the combinations
//   2 p_e - p_i - p_r (and its rotations) are taken after them:
//   2 y_e - y_i - y_r = 3 y_e - (y_e + y_i + y_r).
// * the sixth-order derivative (q7 - q1) - 9 (q6 - q2) + 45 (q5 - q3) keeps
//   its factor 1 / (60 ds) in the velocity factor.
__device__ __forceinline__ void noncon_fast(const double vel[9], const double pe[9], const double pi[9],
                                            const double pr[9], double ids, double out[3]) {
  constexpr double Q0 = -0.00714285714285723, Q1 = 0.0595238095238118, Q2 = -0.240476190476197,
                   Q3 = 0.759523809523819, Q4 = 0.509523809523803, Q5 = -0.0904761904761891,
                   Q6 = 0.00952380952380977;
  constexpr double J0 = -Q0, J1 = Q6 - Q1, J2 = Q5 - Q2, J3 = Q4 - Q3;
  const double v = vel[4];
  const double i3 = ids * (1.0 / 3.0);
  const double cl = fmax(fmax(v, vel[3]), 0.0) * i3;  // cl / ds
  const double cr = fmin(fmin(v, vel[5]), 0.0) * i3;  // cr / ds
  const double vs = v * (ids * (1.0 / 180.0));         // (v / 3) / (60 ds)
  double W[9];
  W[0] = cl * J0;
  W[1] = fma(cl, J1, cr * J0);
  W[2] = fma(cl, J2, cr * J1);
  W[3] = fma(cl, J3, cr * J2);
  W[4] = (cr - cl) * J3;
  W[5] = fma(-cl, J2, -cr * J3);
  W[6] = fma(-cl, J1, -cr * J2);
  W[7] = fma(-cl, J0, -cr * J1);
  W[8] = -cr * J0;
  auto upwind = [&](const double x[9]) {
    double acc = W[0] * x[0];
#pragma unroll
    for (int m = 1; m < 9; ++m) acc = fma(W[m], x[m], acc);
    return acc;
  };
  auto sixth_u = [](const double q[9]) {
    return fma(45.0, q[5] - q[3], fma(-9.0, q[6] - q[2], q[7] - q[1]));
  };
  const double ye = upwind(pe), yi = upwind(pi), yr = upwind(pr);
  const double se = sixth_u(pe), si = sixth_u(pi), sr = sixth_u(pr);
  const double S = ye + yi + yr, D = se + si + sr;
  out[0] = fma(vs, fma(3.0, se, -D), fma(3.0, ye, -S));
  out[1] = fma(vs, fma(3.0, si, -D), fma(3.0, yi, -S));
  out[2] = fma(vs, fma(3.0, sr, -D), fma(3.0, yr, -S));
}

// build_diffusion_from_derived, src/diffusion.rs:27-50: t[0..6) are the
// temperatures at points -2..+3 around the interface.  Returns the flux of
// one energy equation (kappa * .../180); tr4 selects Tr^4.
__device__ __forceinline__ double diffusion_flux(double kappa, const double t[6]) {
  return kappa * (2.0 * t[5] - 25.0 * t[4] + 245.0 * t[3] - 245.0 * t[2] + 25.0 * t[1] -
                  2.0 * t[0]) /
         180.0;
}

// get_local_dt, src/dt.rs:5-34
__device__ __forceinline__ double local_dt(const Phys& ph, const double s[6], double dx, double dy) {
  const double A = ph.a_rad, OEI = ph.omega_ei, OER = ph.omega_er, CVE = ph.cve, CVI = ph.cvi;
  double rho = s[0];
  double u = s[1] / rho, v = s[2] / rho;
  double ee = s[3] / rho - (sq(u) + sq(v)) / 6.0;
  double ei = s[4] / rho - (sq(u) + sq(v)) / 6.0;
  double er = s[5] / rho - (sq(u) + sq(v)) / 6.0;
  double cs = sqrt(ph.gamma_e * ph.ge * ee + ph.gamma_i * ph.gi * ei + ph.gamma_r * ph.gr * er);
  double cve4 = pow4(CVE), cve3 = cube(CVE);
  double t = A * OEI * cve4 + A * OEI * cve3 * CVI + 4.0 * A * OER * CVI * cube(ee) +
             OER * cve4 * CVI * rho;
  double s1 = sqrt(sq(t) - 4.0 * A * OEI * OER * cve4 * CVI *
                               (4.0 * A * cube(ee) + cve4 * rho + cve3 * CVI * rho));
  double s2 = -(A * OEI * cve4 + A * OEI * cve3 * CVI + 4.0 * A * OER * CVI * cube(ee) +
                OER * cve4 * CVI * rho);
  double sd = 2.0 * A * cve4 * CVI * rho;
  double alpha1 = fabs((-s1 + s2) / sd);
  double alpha2 = fabs((s1 + s2) / sd);
  double sj = fmax(alpha1, alpha2);
  double dj = fmax(fmax(ph.kappa_e / (CVE * rho), ph.kappa_i / (CVI * rho)), ph.kappa_r / A);
  double vj = (fabs(u) + cs) / dx + 2.0 * dj / (sq(dx) + sq(dy)) + sj + (fabs(v) + cs) / dy;
  return ph.lambda / vj;
}

// (eps + beta)^q of the boundary WENO weights (powf(WENO_Q), src/bc1.rs:241).
// For q = 10: double-double product chain rounded once (correctly rounded,
// the value glibc's pow returns except within ~2^-60 of a rounding
// midpoint); the oracle evaluates it identically.
__device__ __forceinline__ void dd_mul(double ah, double al, double bh, double bl, double* rh, double* rl) {
  double p = ah * bh;
  double e = fma(ah, bh, -p);
  e += ah * bl + al * bh;
  double s = p + e;
  *rl = e - (s - p);
  *rh = s;
}
__device__ __forceinline__ double weno_pow(double x, double q) {
  if (q != 10.0) return pow(x, q);
  double x2h = x * x, x2l = fma(x, x, -x2h);
  double x4h, x4l, x8h, x8l, rh, rl;
  dd_mul(x2h, x2l, x2h, x2l, &x4h, &x4l);
  dd_mul(x4h, x4l, x4h, x4l, &x8h, &x8l);
  dd_mul(x8h, x8l, x2h, x2l, &rh, &rl);
  return rh + rl;
}

// solve6 (LAPACK dgesv): LU with partial pivoting, first maximal pivot.
// Fully unrolled: the row interchange is a predicated exchange with
// compile-time row indices, so the matrix stays in registers.
__device__ __forceinline__ bool solve6(const double A0[6][6], const double b0[6], double x[6]) {
  double A[6][6], b[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    b[i] = b0[i];
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i][j] = A0[i][j];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    int p = c;
    double best = fabs(A[c][c]);
#pragma unroll
    for (int r = c + 1; r < 6; ++r)
      if (fabs(A[r][c]) > best) {
        best = fabs(A[r][c]);
        p = r;
      }
    if (best == 0.0) return false;
#pragma unroll
    for (int r = c + 1; r < 6; ++r)
      if (r == p) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          double t = A[c][q];
          A[c][q] = A[r][q];
          A[r][q] = t;
        }
        double t = b[c];
        b[c] = b[r];
        b[r] = t;
      }
#pragma unroll
    for (int r = c + 1; r < 6; ++r) {
      double f = A[r][c] / A[c][c];
      A[r][c] = f;
#pragma unroll
      for (int q = c + 1; q < 6; ++q) A[r][q] -= f * A[c][q];
      b[r] -= f * b[c];
    }
  }
#pragma unroll
  for (int r = 5; r >= 0; --r) {
    double s = b[r];
#pragma unroll
    for (int q = r + 1; q < 6; ++q) s -= A[r][q] * x[q];
    x[r] = s / A[r][r];
  }
  return true;
}

// assert_admissible, src/main.rs:22-33
__device__ __forceinline__ bool admissible(const double s[6]) {
#pragma unroll
  for (int c = 0; c < 6; ++c)
    if (!isfinite(s[c])) return false;
  if (s[0] <= 0.0) return false;
  double ux = s[1] / s[0], uy = s[2] / s[0];
  double k = (ux * ux + uy * uy) / 6.0;
  return !(s[3] / s[0] - k <= 0.0 || s[4] / s[0] - k <= 0.0 || s[5] / s[0] - k <= 0.0);
}

// The value of the partner lane (lane ^ 1) of a lane pair: DPP quad
// permutation [1, 0, 3, 2] of each 32-bit half.
__device__ __forceinline__ double pair_swap(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0xB1, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

}  // namespace rh3t_dev
