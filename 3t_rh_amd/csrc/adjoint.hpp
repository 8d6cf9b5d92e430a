// Reverse-mode (VJP) building blocks of l(): pullbacks of the pointwise
// pieces of colinhu123/3T_RH's operator, derived by hand.
//
//   derived_vjp      Derived::from_state            src/state.rs:241-275
//   weno5_vjp        weno5                          src/weno.rs:846-886
//   interface_vjp    Stencil6::reconstruction_fast  src/weno.rs:685-802
//                    with build_l_plain / build_r_plain (src/weno.rs:28-548)
//                    and flux_from_derived (src/state.rs:56-85)
//
// They differentiate the mathematical function the forward kernels evaluate
// (both arithmetic variants compute the same function); at the kinks of
// |lambda| and of max/min the derivative of the branch the forward value
// takes is used.  The adjoint is not held to the forward's rounding
// sequence, so products and sums may contract to FMA here.
#pragma once

#include "device_math.hpp"
#include "flux.hpp"
#include "kernels.hpp"

#pragma clang fp contract(fast)

namespace rh3t_dev {

// ----------------------------------------------------------- WENO5

// Value of weno5(u) (src/weno.rs:846-886: alpha_k = d_k / (eps + beta_k)^2)
// and g = s * d weno5 / d u; the four reciprocals from v_rcp_f64 and a
// Newton step (within 2.3e-15 relative).
__device__ __forceinline__ double weno5_vjp(const double u[5], double eps, double s, double g[5]) {
  const double A0 = u[0] - 2.0 * u[1] + u[2], B0 = u[0] - 4.0 * u[1] + 3.0 * u[2];
  const double A1 = u[1] - 2.0 * u[2] + u[3], B1 = u[1] - u[3];
  const double A2 = u[2] - 2.0 * u[3] + u[4], B2 = 3.0 * u[2] - 4.0 * u[3] + u[4];
  const double beta[3] = {13.0 / 12.0 * A0 * A0 + 0.25 * B0 * B0, 13.0 / 12.0 * A1 * A1 + 0.25 * B1 * B1,
                          13.0 / 12.0 * A2 * A2 + 0.25 * B2 * B2};
  const double dk[3] = {0.1, 0.6, 0.3};
  double a[3], ie[3], S = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ie[k] = rcp_refined(eps + beta[k]);
    a[k] = dk[k] * ie[k] * ie[k];
    S += a[k];
  }
  const double iS = rcp_refined(S);
  const double p[3] = {u[0] * (1.0 / 3.0) - 7.0 / 6.0 * u[1] + 11.0 / 6.0 * u[2],
                       -u[1] * (1.0 / 6.0) + 5.0 / 6.0 * u[2] + u[3] * (1.0 / 3.0),
                       u[2] * (1.0 / 3.0) + 5.0 / 6.0 * u[3] - u[4] * (1.0 / 6.0)};
  const double W = (a[0] * p[0] + a[1] * p[1] + a[2] * p[2]) * iS;
  // candidate polynomials
  const double sS = s * iS;
  const double pb[3] = {sS * a[0], sS * a[1], sS * a[2]};
  g[0] = pb[0] * (1.0 / 3.0);
  g[1] = -7.0 / 6.0 * pb[0] - pb[1] * (1.0 / 6.0);
  g[2] = 11.0 / 6.0 * pb[0] + 5.0 / 6.0 * pb[1] + pb[2] * (1.0 / 3.0);
  g[3] = pb[1] * (1.0 / 3.0) + 5.0 / 6.0 * pb[2];
  g[4] = -pb[2] * (1.0 / 6.0);
  // weights: dW/dbeta_k = (da_k/dbeta_k) (p_k - W) / S, da_k/dbeta_k = -2 a_k / (eps + beta_k)
  double bb[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) bb[k] = -2.0 * pb[k] * ie[k] * (p[k] - W);
  // beta = 13/12 A^2 + 1/4 B^2
  const double a0 = bb[0] * (13.0 / 6.0) * A0, b0 = bb[0] * 0.5 * B0;
  const double a1 = bb[1] * (13.0 / 6.0) * A1, b1 = bb[1] * 0.5 * B1;
  const double a2 = bb[2] * (13.0 / 6.0) * A2, b2 = bb[2] * 0.5 * B2;
  g[0] += a0 + b0;
  g[1] += -2.0 * a0 - 4.0 * b0 + a1 + b1;
  g[2] += a0 + 3.0 * b0 - 2.0 * a1 + a2 + 3.0 * b2;
  g[3] += a1 - b1 - 2.0 * a2 - 4.0 * b2;
  g[4] += a2 + b2;
  return W;
}

// ------------------------------------------------------- Derived VJP

// db: cotangents of (u, v, e_e, e_i, e_r, p_e, p_i, p_r, T_e, T_i, T_r) of
// one point; accumulates the state cotangent into sb.
__device__ __forceinline__ void derived_vjp(const Phys& ph, const double s[6], const double db[D_N], double sb[6]) {
  const double rho = s[0], ir = 1.0 / rho, u = s[1] * ir, v = s[2] * ir;
  const double w2 = u * u + v * v;
  const double g[3] = {ph.ge, ph.gi, ph.gr};
  // T_r = (X / A)^(1/4), X = E_r - rho w2 / 6
  const double X = s[5] - rho * w2 * (1.0 / 6.0);
  double Xb = 0.0;
  if (db[D_TR] != 0.0 && X > 0.0) Xb = db[D_TR] * 0.25 * pow(X / ph.a_rad, 0.25) / X;
  const double icve = 1.0 / ph.cve, icvi = 1.0 / ph.cvi;
  double w2b = -(db[D_EE] + db[D_EI] + db[D_ER]) * (1.0 / 6.0) -
               (db[D_PE] * g[0] + db[D_PI] * g[1] + db[D_PR] * g[2]) * rho * (1.0 / 6.0) -
               (db[D_TE] * icve + db[D_TI] * icvi) * (1.0 / 6.0) - Xb * rho * (1.0 / 6.0);
  const double ub = db[D_U] + 2.0 * u * w2b, vb = db[D_V] + 2.0 * v * w2b;
  double rb = -(u * ub + v * vb) * ir;
  const double pk[3] = {db[D_PE], db[D_PI], db[D_PR]};
  const double ek[3] = {db[D_EE], db[D_EI], db[D_ER]};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    rb += -ek[k] * s[3 + k] * ir * ir - pk[k] * g[k] * w2 * (1.0 / 6.0);
    sb[3 + k] += ek[k] * ir + pk[k] * g[k];
  }
  rb += -(db[D_TE] * s[3] * icve + db[D_TI] * s[4] * icvi) * ir * ir - Xb * w2 * (1.0 / 6.0);
  sb[0] += rb;
  sb[1] += ub * ir;
  sb[2] += vb * ir;
  sb[3] += db[D_TE] * ir * icve;
  sb[4] += db[D_TI] * ir * icvi;
  sb[5] += Xb;
}

// ------------------------------------------------ Roe eigensystem VJP

// Roe-state intermediates of build_l_plain / build_r_plain
// (src/weno.rs:28-548) as functions of (u, v, e_e, e_i, e_r).
struct RoeV {
  double u, v, w, w2, cs2, cs, b, ics2, ib, igtcs2, gtcs2, cs2w2, pcs;
  double e[3], h[3], gk[3];
};

// cotangents of those intermediates
struct RoeBar {
  double u, v, w2, cs, cs2, b, ics2, ib, igtcs2, pcs;
  double h[3], gk[3];
};

template <int DIR>
__device__ __forceinline__ void roe_values(const EigConst& K, double u, double v, const double e[3], RoeV& R) {
  R.u = u;
  R.v = v;
  R.w = DIR == 0 ? u : v;
  R.w2 = u * u + v * v;
  R.cs2 = K.gamg[0] * e[0] + K.gamg[1] * e[1] + K.gamg[2] * e[2];
  R.cs = sqrt(R.cs2);
  const double ge = K.gam[0] * e[0], gi = K.gam[1] * e[1], gr = K.gam[2] * e[2];
  R.h[0] = K.h6[0] * (gi - gr) + K.gam_gt[0] * e[0] * R.w2;
  R.h[1] = K.h6[1] * (gr - ge) + K.gam_gt[1] * e[1] * R.w2;
  R.h[2] = K.h6[2] * (ge - gi) + K.gam_gt[2] * e[2] * R.w2;
  R.b = K.gt * (K.c36 + K.gt * R.w2 * R.w2) * R.cs2;
  R.ics2 = 1.0 / R.cs2;
  R.ib = 1.0 / R.b;
  R.gtcs2 = K.gt * R.cs2;
  R.igtcs2 = 1.0 / R.gtcs2;
  R.cs2w2 = R.cs2 * R.w2;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    R.e[k] = e[k];
    R.gk[k] = K.gam[k] * e[k] + R.w2 * (1.0 / 6.0);
  }
  R.pcs = R.w * R.cs * (1.0 / 3.0);
}

__device__ __forceinline__ double pick3(int k, double a0, double a1, double a2) {
  return k == 0 ? a0 : (k == 1 ? a1 : a2);
}

// the bracket of entry (k, j) of the energy rows (src/weno.rs:110-233):
// g_k h_k - gt cs2 w2 on the diagonal, h_k -/+ 6 g_m cs2 off it
__device__ __forceinline__ double energy_inner(const EigConst& K, const RoeV& R, int k, int j, double gk,
                                               double hk) {
  if (j == k) return gk * hk - R.gtcs2 * R.w2;
  const int m = 3 - k - j;
  const double t6 = pick3(m, K.six_g[0], K.six_g[1], K.six_g[2]) * R.cs2;
  return (j == (k + 1) % 3) ? hk - t6 : hk + t6;
}

// Row c of L, column c of R and lambda_c.  X rows: u-c, e, i, r,
// transverse, u+c; Y rows: v-c, transverse, e, i, r, v+c.
template <int DIR>
__device__ __forceinline__ void row_col(const EigConst& K, const RoeV& R, int c, double row[6], double col[6],
                                        double* lam) {
  const int tr = DIR == 0 ? 4 : 1;
  if (c == 0 || c == 5) {
    const double sg = c == 0 ? 1.0 : -1.0;
    const double iad = R.ics2 * (1.0 / 12.0);
    row[0] = (K.gt * R.w2 + sg * 6.0 * R.w * R.cs) * iad;
    row[1] = (K.m2gt * R.u - (DIR == 0 ? sg * 6.0 * R.cs : 0.0)) * iad;
    row[2] = (K.m2gt * R.v - (DIR == 1 ? sg * 6.0 * R.cs : 0.0)) * iad;
#pragma unroll
    for (int k = 0; k < 3; ++k) row[3 + k] = K.six_g[k] * iad;
    col[0] = 1.0;
    col[1] = DIR == 0 ? R.u - sg * R.cs : R.u;
    col[2] = DIR == 1 ? R.v - sg * R.cs : R.v;
#pragma unroll
    for (int k = 0; k < 3; ++k) col[3 + k] = R.gk[k] - sg * R.pcs;
    *lam = R.w - sg * R.cs;
    return;
  }
  if (c == tr) {
    const double t = DIR == 0 ? R.v : R.u;
    row[0] = -t;
    row[1] = DIR == 0 ? 0.0 : 1.0;
    row[2] = DIR == 0 ? 1.0 : 0.0;
    row[3] = row[4] = row[5] = 0.0;
    col[0] = 0.0;
    col[1] = DIR == 0 ? 0.0 : 1.0;
    col[2] = DIR == 0 ? 1.0 : 0.0;
    col[3] = col[4] = col[5] = t * (1.0 / 3.0);
    *lam = R.w;
    return;
  }
  // energy row k (runtime k: values are picked, never indexed)
  const int k = DIR == 0 ? c - 1 : c - 2;
  const double gk = pick3(k, K.g[0], K.g[1], K.g[2]), hk = pick3(k, R.h[0], R.h[1], R.h[2]);
  const double t = R.ib * (hk - R.cs2w2);
  row[0] = pick3(k, K.g_over_gt[0], K.g_over_gt[1], K.g_over_gt[2]) - gk * (1.0 / 6.0) * R.w2 * R.ics2 -
           pick3(k, K.g_gt[0], K.g_gt[1], K.g_gt[2]) * R.w2 * t;
  const double cuv = gk * (1.0 / 3.0) * R.ics2 + pick3(k, K.two_g_gt[0], K.two_g_gt[1], K.two_g_gt[2]) * t;
  row[1] = cuv * R.u;
  row[2] = cuv * R.v;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double inner = energy_inner(K, R, k, j, gk, hk);
    row[3 + j] = pick3(k, K.num[0][j], K.num[1][j], K.num[2][j]) * R.igtcs2 -
                 pick3(k, K.coef[0][j], K.coef[1][j], K.coef[2][j]) * R.ib * inner;
  }
  col[0] = 1.0;
  col[1] = R.u;
  col[2] = R.v;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int m = 3 - k - j;
    const double gm = pick3(m, K.g[0], K.g[1], K.g[2]);
    col[3 + j] = j == k ? K.gt * pick3(k, K.inv6g[0], K.inv6g[1], K.inv6g[2]) * R.w2
                        : ((j == (k + 1) % 3) ? gm : -gm);
  }
  *lam = R.w;
}

// Accumulates into B the pullback of (row c, column c, lambda_c) with
// cotangents Lb, Rb, lamb.
template <int DIR>
__device__ __forceinline__ void row_col_bar(const EigConst& K, const RoeV& R, int c, const double Lb[6],
                                            const double Rb[6], double lamb, RoeBar& B) {
  const int tr = DIR == 0 ? 4 : 1;
  double wb = 0.0;
  if (c == 0 || c == 5) {
    const double sg = c == 0 ? 1.0 : -1.0;
    const double iad = R.ics2 * (1.0 / 12.0);
    const double A0 = K.gt * R.w2 + sg * 6.0 * R.w * R.cs;
    const double A1 = K.m2gt * R.u - (DIR == 0 ? sg * 6.0 * R.cs : 0.0);
    const double A2 = K.m2gt * R.v - (DIR == 1 ? sg * 6.0 * R.cs : 0.0);
    const double iadb = Lb[0] * A0 + Lb[1] * A1 + Lb[2] * A2 + Lb[3] * K.six_g[0] + Lb[4] * K.six_g[1] +
                        Lb[5] * K.six_g[2];
    const double Ab0 = Lb[0] * iad, Ab1 = Lb[1] * iad, Ab2 = Lb[2] * iad;
    B.w2 += K.gt * Ab0;
    wb += sg * 6.0 * R.cs * Ab0;
    B.cs += sg * 6.0 * R.w * Ab0;
    B.u += K.m2gt * Ab1;
    B.v += K.m2gt * Ab2;
    B.cs -= sg * 6.0 * (DIR == 0 ? Ab1 : Ab2);
    B.ics2 += iadb * (1.0 / 12.0);
    // column (1, u -/+ c, v, gk - pcs) or its Y form
    B.u += Rb[1];
    B.v += Rb[2];
    B.cs -= sg * (DIR == 0 ? Rb[1] : Rb[2]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      B.gk[r] += Rb[3 + r];
      B.pcs -= sg * Rb[3 + r];
    }
    wb += lamb;
    B.cs -= sg * lamb;
  } else if (c == tr) {
    const double tb = -Lb[0] + (Rb[3] + Rb[4] + Rb[5]) * (1.0 / 3.0);
    if (DIR == 0)
      B.v += tb;
    else
      B.u += tb;
    wb += lamb;
  } else {
    const int k = DIR == 0 ? c - 1 : c - 2;
    const double gk = pick3(k, K.g[0], K.g[1], K.g[2]), hk = pick3(k, R.h[0], R.h[1], R.h[2]);
    const double g_gt = pick3(k, K.g_gt[0], K.g_gt[1], K.g_gt[2]);
    const double two_g_gt = pick3(k, K.two_g_gt[0], K.two_g_gt[1], K.two_g_gt[2]);
    const double A = hk - R.cs2w2;
    const double t = R.ib * A;
    const double ulv = R.u * Lb[1] + R.v * Lb[2];
    B.w2 += -(gk * (1.0 / 6.0) * R.ics2 + g_gt * t) * Lb[0];
    B.ics2 += -gk * (1.0 / 6.0) * R.w2 * Lb[0] + gk * (1.0 / 3.0) * ulv;
    const double tb = -g_gt * R.w2 * Lb[0] + two_g_gt * ulv;
    const double cuv = gk * (1.0 / 3.0) * R.ics2 + two_g_gt * t;
    B.u += cuv * Lb[1];
    B.v += cuv * Lb[2];
    double hb = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double l = Lb[3 + j];
      const double coef = pick3(k, K.coef[0][j], K.coef[1][j], K.coef[2][j]);
      const double inner = energy_inner(K, R, k, j, gk, hk);
      const double innerb = -coef * R.ib * l;
      if (j == k) {
        hb += gk * innerb;
        B.cs2 -= K.gt * R.w2 * innerb;
        B.w2 -= R.gtcs2 * innerb;
      } else {
        const int m = 3 - k - j;
        const double sgn = (j == (k + 1) % 3) ? -1.0 : 1.0;
        hb += innerb;
        B.cs2 += sgn * pick3(m, K.six_g[0], K.six_g[1], K.six_g[2]) * innerb;
      }
      B.igtcs2 += pick3(k, K.num[0][j], K.num[1][j], K.num[2][j]) * l;
      B.ib -= coef * inner * l;
    }
    B.ib += tb * A;
    const double Ab = tb * R.ib;
    hb += Ab;
    B.cs2 -= R.w2 * Ab;
    B.w2 -= R.cs2 * Ab;
    B.h[0] += k == 0 ? hb : 0.0;
    B.h[1] += k == 1 ? hb : 0.0;
    B.h[2] += k == 2 ? hb : 0.0;
    // column (1, u, v, .., gt w2 / (6 g_k), ..)
    B.u += Rb[1];
    B.v += Rb[2];
    B.w2 += K.gt * pick3(k, K.inv6g[0], K.inv6g[1], K.inv6g[2]) * pick3(k, Rb[3], Rb[4], Rb[5]);
    wb += lamb;
  }
  if (DIR == 0)
    B.u += wb;
  else
    B.v += wb;
}

// Chains the intermediate cotangents back to (u, v, e_e, e_i, e_r).
template <int DIR>
__device__ __forceinline__ void roe_bar_final(const EigConst& K, const RoeV& R, RoeBar B, double out[5]) {
  B.b -= B.ib * R.ib * R.ib;
  B.cs2 -= B.ics2 * R.ics2 * R.ics2 + B.igtcs2 * K.gt * R.igtcs2 * R.igtcs2;
  // b = gt (c36 + gt w2^2) cs2
  B.w2 += B.b * K.gt * K.gt * 2.0 * R.w2 * R.cs2;
  B.cs2 += B.b * K.gt * (K.c36 + K.gt * R.w2 * R.w2);
  // pcs = w cs / 3
  const double wb = B.pcs * R.cs * (1.0 / 3.0);
  if (DIR == 0)
    B.u += wb;
  else
    B.v += wb;
  B.cs += B.pcs * R.w * (1.0 / 3.0);
  double eb[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    eb[r] = K.gam[r] * B.gk[r];
    B.w2 += B.gk[r] * (1.0 / 6.0);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    eb[k1] += K.h6[k] * K.gam[k1] * B.h[k];
    eb[k2] -= K.h6[k] * K.gam[k2] * B.h[k];
    eb[k] += K.gam_gt[k] * R.w2 * B.h[k];
    B.w2 += K.gam_gt[k] * R.e[k] * B.h[k];
  }
  B.cs2 += B.cs * 0.5 / R.cs;
#pragma unroll
  for (int k = 0; k < 3; ++k) eb[k] += K.gamg[k] * B.cs2;
  out[0] = B.u + 2.0 * R.u * B.w2;
  out[1] = B.v + 2.0 * R.v * B.w2;
  out[2] = eb[0];
  out[3] = eb[1];
  out[4] = eb[2];
}

// Keeps the compiler from carrying loaded stencil data across the phases of
// interface_vjp in registers (it is re-read from L1/L2 instead).
__device__ __forceinline__ void reload_barrier() { __asm__ volatile("" ::: "memory"); }

// -------------------------------------------------- interface flux VJP

// State and physical flux (flux_from_derived, src/state.rs:56-85) of the
// stencil point at byte offset boff of the padded planes.
template <int DIR>
__device__ __forceinline__ void load_point(const double* S, const double* D, long P, unsigned boff, double U[6],
                                           double F[6]) {
#pragma unroll
  for (int f = 0; f < 6; ++f) U[f] = plane_at(S + f * P, boff);
  const double u = plane_at(D + D_U * P, boff), v = plane_at(D + D_V * P, boff);
  const double pe = plane_at(D + D_PE * P, boff), pi = plane_at(D + D_PI * P, boff);
  const double pr = plane_at(D + D_PR * P, boff);
  const double w = DIR == 0 ? u : v;
  F[0] = U[0] * w;
  F[1] = DIR == 0 ? U[0] * u * u + (pe + pi + pr) : U[0] * u * v;
  F[2] = DIR == 0 ? U[0] * u * v : U[0] * v * v + (pe + pi + pr);
  F[3] = (U[3] + pe) * w;
  F[4] = (U[4] + pi) * w;
  F[5] = (U[5] + pr) * w;
}

// Characteristic c of the pullback of one interface flux (interface between
// stencil points 2 and 3), split over a lane pair: lane `half` holds points
// 3 half .. 3 half + 2 (U, F: their states and physical fluxes) and
// evaluates one of the two WENO5 pullbacks (half 0: the reconstruction from
// fp[0..5), half 1: the one from fm[5], fm[4], .., fm[1]); the values each
// needs from the other half cross by pair_swap.  ob: flux cotangent.  cb6:
// the cotangents of the split fluxes fp then fm at this lane's three points
// (their pullback to the stencil states is linear and is formed by the
// point phase).  Both lanes accumulate the same cotangents of row c of L,
// column c of R and lambda_c into B (partial sums over the points are
// combined lane 0 first).
template <int DIR>
__device__ __forceinline__ void characteristic_step_pair(int c, int half, const EigConst& K, double eps,
                                                         const RoeV& R, const double U[3][6], const double F[3][6],
                                                         const double ob[6], double cb6[6], RoeBar& B) {
  double row[6], col[6], lam;
  row_col<DIR>(K, R, c, row, col, &lam);
  const double a = fabs(lam);
  double cu[3], fp[3], fm[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double cf = 0.0, u_ = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      cf += row[f] * F[k][f];
      u_ += row[f] * U[k][f];
    }
    cu[k] = u_;
    fp[k] = 0.5 * (cf + a * u_);
    fm[k] = 0.5 * (cf - a * u_);
  }
  double fcb = 0.0;
#pragma unroll
  for (int r = 0; r < 6; ++r) fcb += col[r] * ob[r];
  // half 0 reconstructs from fp[0..5), half 1 from fm[5], fm[4], .., fm[1]
  const double y0 = pair_swap(half ? fp[0] : fm[1]);
  const double y1 = pair_swap(half ? fp[1] : fm[2]);
  const double in[5] = {half ? fm[2] : fp[0], half ? fm[1] : fp[1], half ? fm[0] : fp[2], half ? y1 : y0,
                        half ? y0 : y1};
  double gw[5];
  const double W = weno5_vjp(in, eps, fcb, gw);
  const double Wo = pair_swap(W);
  const double fc = half ? Wo + W : W + Wo;
  // fpb[q] = gp[q] (q < 5), fmb[q] = gm[5 - q] (q > 0), for this lane's q
  const double w0 = pair_swap(half ? gw[4] : gw[3]);
  const double w1 = pair_swap(half ? gw[3] : gw[4]);
  const double fpb[3] = {half ? w0 : gw[0], half ? w1 : gw[1], half ? 0.0 : gw[2]};
  const double fmb[3] = {half ? gw[2] : 0.0, half ? gw[1] : w0, half ? gw[0] : w1};
  double abp = 0.0;
  double Lbp[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    cb6[k] = fpb[k];
    cb6[3 + k] = fmb[k];
    const double cfb = 0.5 * (fpb[k] + fmb[k]);
    const double d = 0.5 * (fpb[k] - fmb[k]);
    const double cub = a * d;
    abp += cu[k] * d;
#pragma unroll
    for (int f = 0; f < 6; ++f) Lbp[f] += cfb * F[k][f] + cub * U[k][f];
  }
  const double abo = pair_swap(abp);
  const double ab = half ? abo + abp : abp + abo;
  double Lb[6], Rb[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const double o = pair_swap(Lbp[f]);
    Lb[f] = half ? o + Lbp[f] : Lbp[f] + o;
    Rb[f] = fc * ob[f];
  }
  const double lamb = lam >= 0.0 ? ab : -ab;
  row_col_bar<DIR>(K, R, c, Lb, Rb, lamb, B);
}

__device__ __forceinline__ void roe_bar_zero(RoeBar& B) {
  B.u = B.v = B.w2 = B.cs = B.cs2 = B.b = B.ics2 = B.ib = B.igtcs2 = B.pcs = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) B.h[k] = B.gk[k] = 0.0;
}

}  // namespace rh3t_dev

#pragma clang fp contract(off)
