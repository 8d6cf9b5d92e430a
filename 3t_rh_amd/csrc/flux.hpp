// Characteristic WENO5 interface flux, Stencil6::reconstruction_fast of
// colinhu123/3T_RH (src/weno.rs:685-802), for one interface per lane.
//
// Register plan: the 6-point stencil of U and each point's physical flux F
// (EXACT) or normal velocity and pressures (FAST) stay in VGPRs (72 or 66
// doubles); the Roe state is reduced to ~16 scalars and each row
// of L / column of R (src/weno.rs:28-548) is generated when its
// characteristic field is processed, so the 6x6 matrices never exist as a
// whole.  Uniform subexpressions of the physics constants are evaluated once
// on the host with the reference's association (they round identically).
//
// Two arithmetic variants share this code:
//   Exact  (FAST = false): the reference's rounding sequence -- literal
//          divisions, left-to-right dot products, no FMA contraction.
//   Fast   (FAST = true): one reciprocal per shared denominator, FMA
//          chains, WENO5 on first differences with one refined reciprocal,
//          and the flux projections formed from the state projections.
//          Same algebra, different rounding.
#pragma once

#include "device_math.hpp"

namespace rh3t_dev {

// Host-precomputed uniform constants of the eigensystem (index k: 0 e, 1 i,
// 2 r).  Each is evaluated exactly as the reference's expression prefix.
struct EigConst {
  double g[3];         // gamma_k - 1
  double gam[3];       // gamma_k
  double gamg[3];      // gamma_k * g_k                (cs2 terms)
  double g_over_gt[3]; // g_k / gt
  double g_gt[3];      // g_k * gt
  double two_g_gt[3];  // 2.0 * g_k * gt
  double six_g[3];     // 6.0 * g_k
  double h6[3];        // 6.0*gi*gr, 6.0*ge*gr, 6.0*ge*gi (h_e, h_i, h_r)
  double gam_gt[3];    // gamma_k * gt
  double num[3][3];    // -g_k^2 (j == k) or -g_k * g_j
  double coef[3][3];   // 6.0*g_k (j == k) or 6.0*g_k*g_j
  double c36;          // 36.0*ge*gi*gr
  double m2gt;         // -2.0*gt
  double gt;
  // Fast variant only
  double inv_gt;       // 1/gt
  double inv6g[3];     // 1/(6 g_k)
  // Exact variant: RN(1 / six_g[k]), the reciprocal div_rn uses
  double rcp_six_g[3];
};

// The same pointer, opaque to the compiler: the constants read through it
// are loaded (scalar loads) where they are used instead of being hoisted to
// the kernel entry and held in spilled scalar registers throughout.
template <class T>
__device__ __forceinline__ const T* fresh(const T* p) {
  // laundered as a constant-address-space pointer, so that the loads
  // through the result are inferred as constant (scalar) loads, not flat
  using CT = const __attribute__((address_space(4))) T;
  CT* q = (CT*)p;
  __asm__ volatile("" : "+s"(q));
  return (const T*)q;
}

template <int DIR, bool FAST>
struct Roe {
  double u, v, w2, cs, cs2, b, h[3], gk_[3], pcs;  // gk_: gamma_k e_k + w2/6, pcs: (u|v)*cs/3
  double cs2w2, gtcs2;
  // Fast: reciprocals
  double ics2, ib, iad, igtcs2;
  // Exact: RN(1 / (6 cs2)), RN(1 / b), RN(1 / (gt cs2)) for div_rn (with
  // 12 cs2 = 2 (6 cs2) and 3 cs2 = (6 cs2) / 2 exactly, RN(1/(12 cs2)) and
  // RN(1/(3 cs2)) are y6 / 2 and 2 y6)
  double y6, yb, ygt;
};

// a / b: FAST as a * ib (ib ~ 1/b), EXACT correctly rounded through div_rn
// with ib = RN(1/b)
template <bool FAST>
__device__ __forceinline__ double mdiv(double a, double b, double ib) {
  return FAST ? a * ib : div_rn(a, b, ib);
}

template <int DIR, bool FAST>
__device__ __forceinline__ void roe_state(const EigConst& K, double rho3, double rho2, double u3, double v3,
                                          const double e3[3], double u2, double v2, const double e2[3],
                                          Roe<DIR, FAST>& R) {
  const double r1 = sqrt(rho3), r2 = sqrt(rho2);
  const double rs = r1 + r2;
  double e[3];
  if (FAST) {
    const double irs = 1.0 / rs;
    R.u = fma(u3, r1, u2 * r2) * irs;
    R.v = fma(v3, r1, v2 * r2) * irs;
#pragma unroll
    for (int k = 0; k < 3; ++k) e[k] = fma(e3[k], r1, e2[k] * r2) * irs;
    R.w2 = fma(R.u, R.u, R.v * R.v);
    R.cs2 = fma(K.gamg[2], e[2], fma(K.gamg[1], e[1], K.gamg[0] * e[0]));
  } else {
    const double yrs = 1.0 / rs;  // the five divisions by rs through div_rn
    R.u = div_rn(u3 * r1 + u2 * r2, rs, yrs);
    R.v = div_rn(v3 * r1 + v2 * r2, rs, yrs);
#pragma unroll
    for (int k = 0; k < 3; ++k) e[k] = div_rn(e3[k] * r1 + e2[k] * r2, rs, yrs);
    R.w2 = sq(R.u) + sq(R.v);
    R.cs2 = K.gamg[0] * e[0] + K.gamg[1] * e[1] + K.gamg[2] * e[2];
  }
  R.cs = sqrt(R.cs2);
  const double w2 = R.w2, cs2 = R.cs2;
  // h_e, h_i, h_r (src/weno.rs:70-76)
  const double ge_e = K.gam[0] * e[0], gi_e = K.gam[1] * e[1], gr_e = K.gam[2] * e[2];
  if (FAST) {
    R.h[0] = fma(K.h6[0], gi_e - gr_e, K.gam_gt[0] * e[0] * w2);
    R.h[1] = fma(K.h6[1], gr_e - ge_e, K.gam_gt[1] * e[1] * w2);
    R.h[2] = fma(K.h6[2], ge_e - gi_e, K.gam_gt[2] * e[2] * w2);
    R.b = K.gt * fma(K.gt, w2 * w2, K.c36) * cs2;
  } else {
    R.h[0] = K.h6[0] * (gi_e - gr_e) + K.gam_gt[0] * e[0] * w2;
    R.h[1] = K.h6[1] * (gr_e - ge_e) + K.gam_gt[1] * e[1] * w2;
    R.h[2] = K.h6[2] * (ge_e - gi_e) + K.gam_gt[2] * e[2] * w2;
    R.b = K.gt * (K.c36 + K.gt * sq(w2)) * cs2;
  }
  R.gtcs2 = K.gt * cs2;
  R.cs2w2 = cs2 * w2;
  const double w6 = FAST ? w2 * (1.0 / 6.0) : div_rn(w2, 6.0, kRcp6);
#pragma unroll
  for (int k = 0; k < 3; ++k) R.gk_[k] = K.gam[k] * e[k] + w6;
  const double wv = DIR == 0 ? R.u : R.v;
  R.pcs = FAST ? wv * R.cs * (1.0 / 3.0) : div_rn(wv * R.cs, 3.0, kRcp3);
  if (FAST) {
    R.ics2 = 1.0 / cs2;
    R.ib = 1.0 / R.b;
    R.iad = R.ics2 * (1.0 / 12.0);
    R.igtcs2 = R.ics2 * K.inv_gt;
  } else {
    R.y6 = 1.0 / (6.0 * cs2);
    R.yb = 1.0 / R.b;
    R.ygt = 1.0 / R.gtcs2;
  }
}

// Row c of L (src/weno.rs:28-459).  X: rows u-c, e, i, r, transverse, u+c.
// Y: rows v-c, transverse, e, i, r, v+c.
template <int DIR, bool FAST>
__device__ __forceinline__ void l_row(const EigConst& K, const Roe<DIR, FAST>& R, int c, double row[6]) {
  const double u = R.u, v = R.v, w2 = R.w2, cs = R.cs, cs2 = R.cs2;
  const int tr = DIR == 0 ? 4 : 1;
  if (c == 0 || c == 5) {
    const double ad = 12.0 * cs2;
    const double sgn = c == 0 ? 1.0 : -1.0;
    const double wv = DIR == 0 ? u : v;
    // (gt*w2 +- 6*w*cs) / ad, src/weno.rs:87-108 / 258-280
    const double six_w_cs = 6.0 * wv * cs;
    const double a0 = c == 0 ? K.gt * w2 + six_w_cs : K.gt * w2 - six_w_cs;
    const double six_cs = 6.0 * cs;
    double a1, a2;
    if (DIR == 0) {
      a1 = c == 0 ? K.m2gt * u - six_cs : K.m2gt * u + six_cs;
      a2 = K.m2gt * v;
    } else {
      a1 = K.m2gt * u;
      a2 = c == 0 ? K.m2gt * v - six_cs : K.m2gt * v + six_cs;
    }
    (void)sgn;
    const double yad = FAST ? R.iad : 0.5 * R.y6;
    row[0] = mdiv<FAST>(a0, ad, yad);
    row[1] = mdiv<FAST>(a1, ad, yad);
    row[2] = mdiv<FAST>(a2, ad, yad);
    row[3] = mdiv<FAST>(K.six_g[0], ad, yad);
    row[4] = mdiv<FAST>(K.six_g[1], ad, yad);
    row[5] = mdiv<FAST>(K.six_g[2], ad, yad);
    return;
  }
  if (c == tr) {
    row[0] = DIR == 0 ? -v : -u;
    row[1] = DIR == 0 ? 0.0 : 1.0;
    row[2] = DIR == 0 ? 1.0 : 0.0;
    row[3] = row[4] = row[5] = 0.0;
    return;
  }
  const int k = DIR == 0 ? c - 1 : c - 2;  // energy mode 0 e, 1 i, 2 r
  const double gk = K.g[k], hk = R.h[k];
  const double A = hk - R.cs2w2;
  if (FAST) {
    const double t = R.ib * A;
    const double ics6 = R.ics2 * (1.0 / 6.0), ics3 = R.ics2 * (1.0 / 3.0);
    row[0] = K.g_over_gt[k] - gk * w2 * ics6 - K.g_gt[k] * w2 * t;
    row[1] = fma(gk * u, ics3, K.two_g_gt[k] * u * t);
    row[2] = fma(gk * v, ics3, K.two_g_gt[k] * v * t);
  } else {
    const double c6 = 6.0 * cs2, c3 = 3.0 * cs2, y3 = 2.0 * R.y6;
    row[0] = K.g_over_gt[k] - div_rn(gk * w2, c6, R.y6) - div_rn(K.g_gt[k] * w2, R.b, R.yb) * A;
    row[1] = div_rn(gk * u, c3, y3) + div_rn(K.two_g_gt[k] * u, R.b, R.yb) * A;
    row[2] = div_rn(gk * v, c3, y3) + div_rn(K.two_g_gt[k] * v, R.b, R.yb) * A;
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double inner;
    if (j == k) {
      inner = gk * hk - R.gtcs2 * w2;
    } else {
      const int m = 3 - k - j;
      const double t6 = K.six_g[m] * cs2;
      inner = (j == (k + 1) % 3) ? hk - t6 : hk + t6;
    }
    if (FAST)
      row[3 + j] = fma(K.num[k][j], R.igtcs2, -(K.coef[k][j] * R.ib) * inner);
    else
      row[3 + j] = div_rn(K.num[k][j], R.gtcs2, R.ygt) - div_rn(K.coef[k][j], R.b, R.yb) * inner;
  }
}

// Entry (r, c) of R (src/weno.rs:484-548); structural zeros are reported
// through *zero, ones through *one.
template <int DIR, bool FAST>
__device__ __forceinline__ double r_entry(const EigConst& K, const Roe<DIR, FAST>& R, int r, int c, bool* zero,
                                          bool* one) {
  *zero = false;
  *one = false;
  const double u = R.u, v = R.v, cs = R.cs;
  const int tr = DIR == 0 ? 4 : 1;
  if (c == tr) {  // transverse column: X (0,0,1,v/3,..), Y (0,1,0,u/3,..)
    const double w = DIR == 0 ? v : u;
    if (r < 3) {
      const bool is_one = (DIR == 0 && r == 2) || (DIR == 1 && r == 1);
      *zero = !is_one;
      *one = is_one;
      return is_one ? 1.0 : 0.0;
    }
    return FAST ? w * (1.0 / 3.0) : div_rn(w, 3.0, kRcp3);
  }
  if (r == 0) {
    *one = true;
    return 1.0;
  }
  if (c == 0 || c == 5) {
    const double s = c == 0 ? -1.0 : 1.0;
    if (r == 1) return DIR == 0 ? (c == 0 ? u - cs : u + cs) : u;
    if (r == 2) return DIR == 1 ? (c == 0 ? v - cs : v + cs) : v;
    (void)s;
    return c == 0 ? R.gk_[r - 3] - R.pcs : R.gk_[r - 3] + R.pcs;
  }
  if (r == 1) return u;
  if (r == 2) return v;
  const int k = DIR == 0 ? c - 1 : c - 2;  // energy column
  const int j = r - 3;
  if (j == k) return FAST ? K.gt * R.w2 * K.inv6g[k] : div_rn(K.gt * R.w2, K.six_g[k], K.rcp_six_g[k]);
  const int m = 3 - k - j;
  return (j == (k + 1) % 3) ? K.g[m] : -K.g[m];
}

// 1/x from v_rcp_f64 (relative error up to 2^-24.4 on gfx950) refined by
// RH3T_RCP_NEWTON Newton steps: one step leaves at most 2.3e-15 (about
// 10 ulp), two give the correctly rounded reciprocal.  FAST only.
#ifndef RH3T_RCP_NEWTON
#define RH3T_RCP_NEWTON 1
#endif
__device__ __forceinline__ double rcp_refined(double x) {
  double r = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int k = 0; k < RH3T_RCP_NEWTON; ++k) r = fma(fma(-x, r, 1.0), r, r);
  return r;
}

// weno5, src/weno.rs:846-886.
// EXACT: the reference's expression (weno5 of device_math.hpp).
// FAST: the same weights and candidate values written on the first
// differences a = u1 - u0, b = u2 - u1, c = u3 - u2, dd = u4 - u3.  The
// smoothness indicators 13/12 d_k^2 + 1/4 e_k^2 (d_k the second difference,
// e_k the one-sided first difference of candidate k) are the quadratic forms
//   3 beta_0 = 4a^2 - 11ab + 10b^2,
//   3 beta_1 = 4b^2 - 5bc + 4c^2,
//   3 beta_2 = 10c^2 - 11c dd + 4dd^2,
// and the candidates are p_k = u2 + r_k / 6 with r = (5b - 2a, b + 2c,
// 4c - dd), so weno5 = u2 + sum_k t_k r_k / (6 sum_k t_k) with
// t_k = d_k prod_{j!=k} s_j and s_k = (eps3 + 3 beta_k)^2: the caller passes
// eps3 = 3 eps, which leaves the normalised weights as they are.
template <bool FAST>
__device__ __forceinline__ double weno5v(double u0, double u1, double u2, double u3, double u4, double eps) {
  if (!FAST) return weno5(u0, u1, u2, u3, u4, eps);
  const double a = u1 - u0, b = u2 - u1, c = u3 - u2, dd = u4 - u3;
  const double bb = b * b, cc = c * c;
  double s0 = eps + fma(a, fma(4.0, a, -11.0 * b), 10.0 * bb);
  double s1 = eps + fma(-5.0 * b, c, 4.0 * (bb + cc));
  double s2 = eps + fma(dd, fma(4.0, dd, -11.0 * c), 10.0 * cc);
  s0 *= s0;
  s1 *= s1;
  s2 *= s2;
  // w_k = (d_k / s_k) / sum_m (d_m / s_m), with all three over one quotient
  const double t0 = 0.1 * (s1 * s2), t1 = 0.6 * (s0 * s2), t2 = 0.3 * (s0 * s1);
  const double r0 = fma(5.0, b, -2.0 * a), r1 = fma(2.0, c, b), r2 = fma(4.0, c, -dd);
  const double num = fma(t0, r0, fma(t1, r1, t2 * r2));
  return fma(num, rcp_refined(6.0 * (t0 + t1 + t2)), u2);
}

template <bool FAST>
__device__ __forceinline__ double dot6v(const double r[6], const double x[6]) {
  if (!FAST) return dot6(r, x);
  double acc = r[0] * x[0];
#pragma unroll
  for (int k = 1; k < 6; ++k) acc = fma(r[k], x[k], acc);
  return acc;
}

// The 6-point stencil of one interface (points q = 0..5, the interface
// between points 2 and 3).  EXACT keeps each point's physical flux F
// (flux_from_derived, src/state.rs:56-85) and projects it as the reference
// does.  FAST keeps the normal velocity w and the pressures instead: with
// F = w U + (0, p n_x, p n_y, w p_e, w p_i, w p_r), the projection of the
// flux on a row l of L follows from that of the state,
//   l . F = w (l . U + l_3 p_e + l_4 p_i + l_5 p_r) + l_n p,
// five operations where the dot product takes six, and five values per
// point held instead of six.
template <bool FAST>
struct Stencil;
template <>
struct Stencil<false> {
  double U[6][6], F[6][6];
};
template <>
struct Stencil<true> {
  double U[6][6], w[6], pe[6], pi[6], pr[6], p[6];
};

// One interface from its stencil; e2/e3 the specific energies, (u2, v2) and
// (u3, v3) the velocities of points 2 and 3.
template <int DIR, bool FAST>
__device__ __forceinline__ void interface_flux(const EigConst* Kp, double eps, const Stencil<FAST>& P, double u3,
                                               double v3, const double e3[3], double u2, double v2,
                                               const double e2[3], double out[6]) {
  const auto& U = P.U;
  Roe<DIR, FAST> R;
  roe_state<DIR, FAST>(*fresh(Kp), U[3][0], U[2][0], u3, v3, e3, u2, v2, e2, R);
  const int tr = DIR == 0 ? 4 : 1;
  // FAST forms the split fluxes f+- = (cf +- |lam| cu) / 2 without the
  // factor 1/2 and halves the sum of the two reconstructions instead:
  // weno5(2x, 4 eps) = 2 weno5(x, eps) while no value leaves the normal
  // range (every operation commutes with scaling by a power of two), and its
  // weno5v takes 3 x that eps.  EXACT halves each split flux, as the
  // reference does: near-uniform flow gives subnormal split fluxes, where
  // the scaling is not exact.
  const double ew = FAST ? 12.0 * eps : eps;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const EigConst& K = *fresh(Kp);
    double row[6];
    l_row<DIR, FAST>(K, R, c, row);
    double lam = (c == 0) ? (DIR == 0 ? R.u - R.cs : R.v - R.cs)
                          : (c == 5 ? (DIR == 0 ? R.u + R.cs : R.v + R.cs) : (DIR == 0 ? R.u : R.v));
    const double a = fabs(lam);
    double fp[6], fm[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      double cf, cu;
      if constexpr (FAST) {
        if (c == tr) {  // row (-w_t, e_t): l . F = w (l . U)
          cu = fma(row[0], U[q][0], DIR == 0 ? U[q][2] : U[q][1]);
          cf = P.w[q] * cu;
        } else {
          cu = dot6v<FAST>(row, U[q]);
          const double lp = fma(row[5], P.pr[q], fma(row[4], P.pi[q], fma(row[3], P.pe[q], cu)));
          cf = fma(P.w[q], lp, (DIR == 0 ? row[1] : row[2]) * P.p[q]);
        }
      } else if (c == tr) {  // row (-w, e_t): two terms
        cf = row[0] * P.F[q][0] + (DIR == 0 ? P.F[q][2] : P.F[q][1]);
        cu = row[0] * U[q][0] + (DIR == 0 ? U[q][2] : U[q][1]);
      } else {
        cf = dot6v<FAST>(row, P.F[q]);
        cu = dot6v<FAST>(row, U[q]);
      }
      if (FAST) {
        fp[q] = fma(a, cu, cf);
        fm[q] = fma(-a, cu, cf);
      } else {
        fp[q] = 0.5 * (cf + a * cu);
        fm[q] = 0.5 * (cf - a * cu);
      }
    }
    const double fsum = weno5v<FAST>(fp[0], fp[1], fp[2], fp[3], fp[4], ew) +
                        weno5v<FAST>(fm[5], fm[4], fm[3], fm[2], fm[1], ew);
    const double fc = FAST ? 0.5 * fsum : fsum;
    // r_dot accumulated over c left to right
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      bool zero, one;
      const double rv = r_entry<DIR, FAST>(K, R, r, c, &zero, &one);
      if (c == 0) {
        out[r] = one ? fc : rv * fc;
      } else if (!zero) {
        if (one)
          out[r] = out[r] + fc;
        else
          out[r] = FAST ? fma(rv, fc, out[r]) : out[r] + rv * fc;
      }
    }
  }
}

}  // namespace rh3t_dev
