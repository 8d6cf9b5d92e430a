// Reverse-mode (VJP) building blocks of l(): the pullbacks of the pointwise
// pieces of colinhu123/3T_RH's operator, written by hand.
//
//   derived_vjp      Derived::from_state            src/state.rs:241-275
//   weno5_vjp        weno5                          src/weno.rs:846-886
//   interface_vjp    Stencil6::reconstruction_fast  src/weno.rs:685-802
//                    (L rows / R columns / lambda as functions of the Roe
//                    state are differentiated with 5-direction forward-mode
//                    duals, everything else in reverse mode)
//   noncon_vjp       nonconservative_direction_pre  src/noncon.rs:288-332
//
// The derivatives are those of the mathematical function the reference
// evaluates (kinks of |lambda|, max/min and the WENO weights' branches are
// taken on the side the forward value selects).
#pragma once

#include "device_math.hpp"
#include "flux.hpp"

namespace rh3t_dev {

// ----------------------------------------------------------- Dual5

struct D5 {
  double v, d[5];
};
__device__ __forceinline__ D5 dconst(double x) {
  D5 r;
  r.v = x;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = 0.0;
  return r;
}
__device__ __forceinline__ D5 dvar(double x, int k) {
  D5 r = dconst(x);
  r.d[k] = 1.0;
  return r;
}
__device__ __forceinline__ D5 operator+(const D5& a, const D5& b) {
  D5 r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ D5 operator-(const D5& a, const D5& b) {
  D5 r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ __forceinline__ D5 operator-(const D5& a) {
  D5 r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = -a.d[k];
  return r;
}
__device__ __forceinline__ D5 operator*(const D5& a, const D5& b) {
  D5 r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ D5 operator*(double a, const D5& b) {
  D5 r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a * b.d[k];
  return r;
}
__device__ __forceinline__ D5 operator/(const D5& a, const D5& b) {
  D5 r;
  const double ib = 1.0 / b.v;
  r.v = a.v * ib;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
__device__ __forceinline__ D5 operator/(double a, const D5& b) { return dconst(a) / b; }
__device__ __forceinline__ D5 dsqrt(const D5& a) {
  D5 r;
  r.v = sqrt(a.v);
  const double h = 0.5 / r.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * h;
  return r;
}

// Roe-state functions of (u, v, e_e, e_i, e_r) with the closed forms of
// src/weno.rs:28-548 (flux.hpp keeps the production rounding; these carry
// derivatives).
template <int DIR>
struct RoeD {
  D5 u, v, w2, cs, cs2, b, h[3], gk_[3], pcs, cs2w2, gtcs2;
};

template <int DIR>
__device__ __forceinline__ void roe_dual(const EigConst& K, double u, double v, const double e[3], RoeD<DIR>& R) {
  R.u = dvar(u, 0);
  R.v = dvar(v, 1);
  D5 E[3] = {dvar(e[0], 2), dvar(e[1], 3), dvar(e[2], 4)};
  R.w2 = R.u * R.u + R.v * R.v;
  R.cs2 = K.gamg[0] * E[0] + K.gamg[1] * E[1] + K.gamg[2] * E[2];
  R.cs = dsqrt(R.cs2);
  D5 ge_e = K.gam[0] * E[0], gi_e = K.gam[1] * E[1], gr_e = K.gam[2] * E[2];
  R.h[0] = K.h6[0] * (gi_e - gr_e) + K.gam_gt[0] * E[0] * R.w2;
  R.h[1] = K.h6[1] * (gr_e - ge_e) + K.gam_gt[1] * E[1] * R.w2;
  R.h[2] = K.h6[2] * (ge_e - gi_e) + K.gam_gt[2] * E[2] * R.w2;
  R.b = K.gt * (dconst(K.c36) + K.gt * (R.w2 * R.w2)) * R.cs2;
  R.gtcs2 = K.gt * R.cs2;
  R.cs2w2 = R.cs2 * R.w2;
  D5 w6 = (1.0 / 6.0) * R.w2;
#pragma unroll
  for (int k = 0; k < 3; ++k) R.gk_[k] = K.gam[k] * E[k] + w6;
  R.pcs = (1.0 / 3.0) * ((DIR == 0 ? R.u : R.v) * R.cs);
}

template <int DIR>
__device__ __forceinline__ void l_row_dual(const EigConst& K, const RoeD<DIR>& R, int c, D5 row[6]) {
  const int tr = DIR == 0 ? 4 : 1;
  if (c == 0 || c == 5) {
    D5 iad = 1.0 / (12.0 * R.cs2);
    D5 wv = DIR == 0 ? R.u : R.v;
    D5 six_w_cs = 6.0 * (wv * R.cs);
    D5 a0 = c == 0 ? K.gt * R.w2 + six_w_cs : K.gt * R.w2 - six_w_cs;
    D5 six_cs = 6.0 * R.cs;
    D5 a1, a2;
    if (DIR == 0) {
      a1 = c == 0 ? K.m2gt * R.u - six_cs : K.m2gt * R.u + six_cs;
      a2 = K.m2gt * R.v;
    } else {
      a1 = K.m2gt * R.u;
      a2 = c == 0 ? K.m2gt * R.v - six_cs : K.m2gt * R.v + six_cs;
    }
    row[0] = a0 * iad;
    row[1] = a1 * iad;
    row[2] = a2 * iad;
    row[3] = K.six_g[0] * iad;
    row[4] = K.six_g[1] * iad;
    row[5] = K.six_g[2] * iad;
    return;
  }
  if (c == tr) {
    row[0] = DIR == 0 ? -R.v : -R.u;
    row[1] = dconst(DIR == 0 ? 0.0 : 1.0);
    row[2] = dconst(DIR == 0 ? 1.0 : 0.0);
    row[3] = row[4] = row[5] = dconst(0.0);
    return;
  }
  const int k = DIR == 0 ? c - 1 : c - 2;
  const double gk = K.g[k];
  D5 hk = R.h[k];
  D5 A = hk - R.cs2w2;
  D5 ib = 1.0 / R.b;
  D5 ics2 = 1.0 / R.cs2;
  D5 t = ib * A;
  row[0] = dconst(K.g_over_gt[k]) - (gk / 6.0) * (R.w2 * ics2) - K.g_gt[k] * (R.w2 * t);
  row[1] = (gk / 3.0) * (R.u * ics2) + K.two_g_gt[k] * (R.u * t);
  row[2] = (gk / 3.0) * (R.v * ics2) + K.two_g_gt[k] * (R.v * t);
  D5 igtcs2 = 1.0 / R.gtcs2;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    D5 inner;
    if (j == k) {
      inner = gk * hk - R.gtcs2 * R.w2;
    } else {
      const int m = 3 - k - j;
      D5 t6 = K.six_g[m] * R.cs2;
      inner = (j == (k + 1) % 3) ? hk - t6 : hk + t6;
    }
    row[3 + j] = K.num[k][j] * igtcs2 - K.coef[k][j] * (ib * inner);
  }
}

// column c of R
template <int DIR>
__device__ __forceinline__ void r_col_dual(const EigConst& K, const RoeD<DIR>& R, int c, D5 col[6]) {
  const int tr = DIR == 0 ? 4 : 1;
  if (c == tr) {
    D5 w3 = (1.0 / 3.0) * (DIR == 0 ? R.v : R.u);
    col[0] = dconst(0.0);
    col[1] = dconst(DIR == 0 ? 0.0 : 1.0);
    col[2] = dconst(DIR == 0 ? 1.0 : 0.0);
    col[3] = col[4] = col[5] = w3;
    return;
  }
  col[0] = dconst(1.0);
  if (c == 0 || c == 5) {
    col[1] = DIR == 0 ? (c == 0 ? R.u - R.cs : R.u + R.cs) : R.u;
    col[2] = DIR == 1 ? (c == 0 ? R.v - R.cs : R.v + R.cs) : R.v;
#pragma unroll
    for (int r = 0; r < 3; ++r) col[3 + r] = c == 0 ? R.gk_[r] - R.pcs : R.gk_[r] + R.pcs;
    return;
  }
  col[1] = R.u;
  col[2] = R.v;
  const int k = DIR == 0 ? c - 1 : c - 2;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j == k) {
      col[3 + j] = (K.gt * K.inv6g[k]) * R.w2;
    } else {
      const int m = 3 - k - j;
      col[3 + j] = dconst((j == (k + 1) % 3) ? K.g[m] : -K.g[m]);
    }
  }
}

template <int DIR>
__device__ __forceinline__ D5 lambda_dual(const RoeD<DIR>& R, int c) {
  D5 w = DIR == 0 ? R.u : R.v;
  return c == 0 ? w - R.cs : (c == 5 ? w + R.cs : w);
}

// ----------------------------------------------------------- WENO5

// value of weno5 and the gradient of s * weno5 w.r.t. its 5 inputs
__device__ __forceinline__ double weno5_vjp(const double u[5], double eps, double s, double g[5]) {
  const double A0 = u[0] - 2.0 * u[1] + u[2], B0 = u[0] - 4.0 * u[1] + 3.0 * u[2];
  const double A1 = u[1] - 2.0 * u[2] + u[3], B1 = u[1] - u[3];
  const double A2 = u[2] - 2.0 * u[3] + u[4], B2 = 3.0 * u[2] - 4.0 * u[3] + u[4];
  const double beta[3] = {13.0 / 12.0 * A0 * A0 + 0.25 * B0 * B0, 13.0 / 12.0 * A1 * A1 + 0.25 * B1 * B1,
                          13.0 / 12.0 * A2 * A2 + 0.25 * B2 * B2};
  const double dk[3] = {0.1, 0.6, 0.3};
  double a[3], S = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double e = eps + beta[k];
    a[k] = dk[k] / (e * e);
    S += a[k];
  }
  const double p[3] = {u[0] / 3.0 - 7.0 / 6.0 * u[1] + 11.0 / 6.0 * u[2],
                       -u[1] / 6.0 + 5.0 / 6.0 * u[2] + u[3] / 3.0,
                       u[2] / 3.0 + 5.0 / 6.0 * u[3] - u[4] / 6.0};
  double W = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) W += a[k] / S * p[k];
  // p_k pullback
  const double pb[3] = {s * a[0] / S, s * a[1] / S, s * a[2] / S};
  g[0] = pb[0] / 3.0;
  g[1] = -7.0 / 6.0 * pb[0] - pb[1] / 6.0;
  g[2] = 11.0 / 6.0 * pb[0] + 5.0 / 6.0 * pb[1] + pb[2] / 3.0;
  g[3] = pb[1] / 3.0 + 5.0 / 6.0 * pb[2];
  g[4] = -pb[2] / 6.0;
  // beta_k pullback: dW/dbeta_k = (da_k/dbeta_k) (p_k - W) / S,
  // da_k/dbeta_k = -2 a_k / (eps + beta_k)
  double bb[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) bb[k] = s * (-2.0 * a[k] / (eps + beta[k])) * (p[k] - W) / S;
  // beta = 13/12 A^2 + 1/4 B^2
  const double a0 = bb[0] * 13.0 / 6.0 * A0, b0 = bb[0] * 0.5 * B0;
  const double a1 = bb[1] * 13.0 / 6.0 * A1, b1 = bb[1] * 0.5 * B1;
  const double a2 = bb[2] * 13.0 / 6.0 * A2, b2 = bb[2] * 0.5 * B2;
  g[0] += a0 + b0;
  g[1] += -2.0 * a0 - 4.0 * b0 + a1 + b1;
  g[2] += a0 + 3.0 * b0 - 2.0 * a1 + a2 + 3.0 * b2;
  g[3] += a1 - b1 - 2.0 * a2 - 4.0 * b2;
  g[4] += a2 + b2;
  return W;
}

// ------------------------------------------------------- Derived VJP

// db: (u, v, e_e, e_i, e_r, p_e, p_i, p_r, T_e, T_i, T_r) cotangents;
// accumulates the state cotangent into sb.
__device__ __forceinline__ void derived_vjp(const Phys& ph, const double s[6], const double db[11], double sb[6]) {
  const double rho = s[0], u = s[1] / rho, v = s[2] / rho;
  const double w2 = u * u + v * v;
  const double g[3] = {ph.ge, ph.gi, ph.gr};
  const double X = s[5] - rho * w2 / 6.0;  // rho e_r, Tr = (X/A)^(1/4)
  double Xb = 0.0;
  if (db[10] != 0.0 && X > 0.0) {
    const double tr = pow(X / ph.a_rad, 0.25);
    Xb = db[10] * 0.25 * tr / X;
  }
  double w2b = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) w2b += -db[2 + k] / 6.0 - db[5 + k] * g[k] * rho / 6.0;
  w2b += -db[8] / (6.0 * ph.cve) - db[9] / (6.0 * ph.cvi) - Xb * rho / 6.0;
  const double ub = db[0] + 2.0 * u * w2b, vb = db[1] + 2.0 * v * w2b;
  sb[1] += ub / rho;
  sb[2] += vb / rho;
  double rb = -(u * ub + v * vb) / rho;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    rb += -db[2 + k] * s[3 + k] / (rho * rho) - db[5 + k] * g[k] * w2 / 6.0;
    sb[3 + k] += db[2 + k] / rho + db[5 + k] * g[k];
  }
  rb += -db[8] * s[3] / (rho * rho * ph.cve) - db[9] * s[4] / (rho * rho * ph.cvi) - Xb * w2 / 6.0;
  sb[3] += db[8] / (rho * ph.cve);
  sb[4] += db[9] / (rho * ph.cvi);
  sb[5] += Xb;
  sb[0] += rb;
}

// -------------------------------------------------- interface flux VJP

// Pullback of one interface flux (interface between stencil points 2 and
// 3).  Inputs as for interface_flux; ob: flux cotangent.  Outputs: Ub[q][6]
// state cotangents and dbq[q][5] cotangents of (u, v, p_e, p_i, p_r) of
// every point, eb2/eb3 cotangents of (e_e, e_i, e_r) at points 2 and 3.
template <int DIR>
__device__ __forceinline__ void interface_vjp(const EigConst& K, double eps, const double U[6][6],
                                              const double dq[6][5], const double e3[3], const double e2[3],
                                              const double ob[6], double Ub[6][6], double dbq[6][5], double eb3[3],
                                              double eb2[3]) {
  // physical fluxes
  double F[6][6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double u = dq[q][0], v = dq[q][1], ptot = dq[q][2] + dq[q][3] + dq[q][4];
    const double w = DIR == 0 ? u : v;
    F[q][0] = U[q][0] * w;
    F[q][1] = DIR == 0 ? U[q][0] * u * u + ptot : U[q][0] * u * v;
    F[q][2] = DIR == 0 ? U[q][0] * u * v : U[q][0] * v * v + ptot;
    F[q][3] = (U[q][3] + dq[q][2]) * w;
    F[q][4] = (U[q][4] + dq[q][3]) * w;
    F[q][5] = (U[q][5] + dq[q][4]) * w;
  }
  // Roe state (points 3 and 2)
  const double r1 = sqrt(U[3][0]), r2 = sqrt(U[2][0]), rs = r1 + r2;
  const double ru = (dq[3][0] * r1 + dq[2][0] * r2) / rs;
  const double rv = (dq[3][1] * r1 + dq[2][1] * r2) / rs;
  double re[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) re[k] = (e3[k] * r1 + e2[k] * r2) / rs;
  RoeD<DIR> R;
  roe_dual<DIR>(K, ru, rv, re, R);
  double roeb[5] = {0, 0, 0, 0, 0};
  double Fb[6][6];
#pragma unroll
  for (int q = 0; q < 6; ++q)
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      Fb[q][f] = 0.0;
      Ub[q][f] = 0.0;
    }
  for (int c = 0; c < 6; ++c) {
    D5 row[6], col[6];
    l_row_dual<DIR>(K, R, c, row);
    r_col_dual<DIR>(K, R, c, col);
    const D5 lam = lambda_dual<DIR>(R, c);
    const double a = fabs(lam.v);
    double cf[6], cu[6], fp[6], fm[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      cf[q] = 0.0;
      cu[q] = 0.0;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        cf[q] += row[f].v * F[q][f];
        cu[q] += row[f].v * U[q][f];
      }
      fp[q] = 0.5 * (cf[q] + a * cu[q]);
      fm[q] = 0.5 * (cf[q] - a * cu[q]);
    }
    // forward value of this characteristic flux, its cotangent
    double fcb = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r) fcb += col[r].v * ob[r];
    double gp[5], gm[5];
    const double sp[5] = {fp[0], fp[1], fp[2], fp[3], fp[4]};
    const double sm[5] = {fm[5], fm[4], fm[3], fm[2], fm[1]};
    const double fc = weno5_vjp(sp, eps, fcb, gp) + weno5_vjp(sm, eps, fcb, gm);
    double fpb[6] = {gp[0], gp[1], gp[2], gp[3], gp[4], 0.0};
    double fmb[6] = {0.0, gm[4], gm[3], gm[2], gm[1], gm[0]};
    double ab = 0.0;
    double Lb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double cfb = 0.5 * (fpb[q] + fmb[q]);
      const double cub = 0.5 * a * (fpb[q] - fmb[q]);
      ab += 0.5 * cu[q] * (fpb[q] - fmb[q]);
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        Lb[f] += cfb * F[q][f] + cub * U[q][f];
        Fb[q][f] += cfb * row[f].v;
        Ub[q][f] += cub * row[f].v;
      }
    }
    const double lamb = lam.v >= 0.0 ? ab : -ab;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      double acc = lamb * lam.d[k];
#pragma unroll
      for (int f = 0; f < 6; ++f) acc += Lb[f] * row[f].d[k] + ob[f] * fc * col[f].d[k];
      roeb[k] += acc;
    }
  }
  // physical flux pullback -> U and (u, v, p_k) of each point
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double u = dq[q][0], v = dq[q][1];
    const double rho = U[q][0];
    const double w = DIR == 0 ? u : v;
    double ub = 0.0, vb = 0.0, wb = 0.0;
    // F0 = rho w
    Ub[q][0] += Fb[q][0] * w;
    wb += Fb[q][0] * rho;
    if (DIR == 0) {
      // F1 = rho u u + P, F2 = rho u v
      Ub[q][0] += Fb[q][1] * u * u + Fb[q][2] * u * v;
      ub += Fb[q][1] * 2.0 * rho * u + Fb[q][2] * rho * v;
      vb += Fb[q][2] * rho * u;
    } else {
      Ub[q][0] += Fb[q][1] * u * v + Fb[q][2] * v * v;
      ub += Fb[q][1] * rho * v;
      vb += Fb[q][1] * rho * u + Fb[q][2] * 2.0 * rho * v;
    }
    const double Pb = DIR == 0 ? Fb[q][1] : Fb[q][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Ub[q][3 + k] += Fb[q][3 + k] * w;
      wb += Fb[q][3 + k] * (U[q][3 + k] + dq[q][2 + k]);
      dbq[q][2 + k] = Pb + Fb[q][3 + k] * w;
    }
    if (DIR == 0)
      ub += wb;
    else
      vb += wb;
    dbq[q][0] = ub;
    dbq[q][1] = vb;
  }
  // Roe average pullback
  const double qv3[5] = {dq[3][0], dq[3][1], e3[0], e3[1], e3[2]};
  const double qv2[5] = {dq[2][0], dq[2][1], e2[0], e2[1], e2[2]};
  const double qr[5] = {ru, rv, re[0], re[1], re[2]};
  double r1b = 0.0, r2b = 0.0;
  double b3[5], b2[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    b3[k] = roeb[k] * r1 / rs;
    b2[k] = roeb[k] * r2 / rs;
    r1b += roeb[k] * (qv3[k] - qr[k]) / rs;
    r2b += roeb[k] * (qv2[k] - qr[k]) / rs;
  }
  dbq[3][0] += b3[0];
  dbq[3][1] += b3[1];
  dbq[2][0] += b2[0];
  dbq[2][1] += b2[1];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    eb3[k] = b3[2 + k];
    eb2[k] = b2[2 + k];
  }
  Ub[3][0] += r1b * 0.5 / r1;
  Ub[2][0] += r2b * 0.5 / r2;
}

// ----------------------------------------------------------- noncon VJP

// Pullback of nonconservative_direction_pre's (ee, ei, er) output nb[3]
// w.r.t. vel[k] (k = 3, 4, 5 only are used) and p_e/p_i/p_r at the 9 points.
__device__ __forceinline__ void noncon_vjp(const double vel[9], const double pe[9], const double pi[9],
                                           const double pr[9], double ds, const double nb[3], double velb[9],
                                           double peb[9], double pib[9], double prb[9]) {
  const double* P[3] = {pe, pi, pr};
  double* PB[3] = {peb, pib, prb};
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    velb[k] = 0.0;
    peb[k] = pib[k] = prb[k] = 0.0;
  }
  const double v = vel[4];
  double dp[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) dp[m] = sixth(P[m][1], P[m][2], P[m][3], P[m][5], P[m][6], P[m][7], ds);
  const double der[3] = {2.0 * dp[0] - dp[1] - dp[2], 2.0 * dp[1] - dp[0] - dp[2], 2.0 * dp[2] - dp[0] - dp[1]};
  double N[3][9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    N[0][k] = 2.0 * pe[k] - pi[k] - pr[k];
    N[1][k] = 2.0 * pi[k] - pe[k] - pr[k];
    N[2][k] = 2.0 * pr[k] - pe[k] - pi[k];
  }
  const double m1 = fmax(v, vel[3]);
  const double cl = fmax(m1, 0.0) / 3.0;
  const double n1 = fmin(v, vel[5]);
  const double cr = fmin(n1, 0.0) / 3.0;
  double clb = 0.0, crb = 0.0, vb = 0.0;
  double derb[3];
  double Nb[3][9];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const double jm = jump(N[e]), jp = jump(N[e] + 1);
    vb += nb[e] * der[e] / 3.0;
    derb[e] = nb[e] * v / 3.0;
    clb += nb[e] * jm / ds;
    crb += nb[e] * jp / ds;
    const double jmb = nb[e] * cl / ds, jpb = nb[e] * cr / ds;
    // jump(a) = q(a7..a1) - q(a0..a6)
    const double qc[7] = {-0.00714285714285723, 0.0595238095238118, -0.240476190476197, 0.759523809523819,
                          0.509523809523803,    -0.0904761904761891, 0.00952380952380977};
#pragma unroll
    for (int k = 0; k < 9; ++k) Nb[e][k] = 0.0;
#pragma unroll
    for (int t = 0; t < 7; ++t) {
      Nb[e][7 - t] += jmb * qc[t];
      Nb[e][t] -= jmb * qc[t];
      Nb[e][8 - t] += jpb * qc[t];
      Nb[e][1 + t] -= jpb * qc[t];
    }
  }
  const double dpb[3] = {2.0 * derb[0] - derb[1] - derb[2], -derb[0] + 2.0 * derb[1] - derb[2],
                         -derb[0] - derb[1] + 2.0 * derb[2]};
  const double inv = 1.0 / (60.0 * ds);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    PB[m][1] += -dpb[m] * inv;
    PB[m][2] += 9.0 * dpb[m] * inv;
    PB[m][3] += -45.0 * dpb[m] * inv;
    PB[m][5] += 45.0 * dpb[m] * inv;
    PB[m][6] += -9.0 * dpb[m] * inv;
    PB[m][7] += dpb[m] * inv;
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    peb[k] += 2.0 * Nb[0][k] - Nb[1][k] - Nb[2][k];
    pib[k] += -Nb[0][k] + 2.0 * Nb[1][k] - Nb[2][k];
    prb[k] += -Nb[0][k] - Nb[1][k] + 2.0 * Nb[2][k];
  }
  // cl = max(max(v, v3), 0)/3, cr = min(min(v, v5), 0)/3
  if (m1 >= 0.0) {
    if (v >= vel[3])
      vb += clb / 3.0;
    else
      velb[3] += clb / 3.0;
  }
  if (n1 <= 0.0) {
    if (v <= vel[5])
      vb += crb / 3.0;
    else
      velb[5] += crb / 3.0;
  }
  velb[4] += vb;
}

}  // namespace rh3t_dev
