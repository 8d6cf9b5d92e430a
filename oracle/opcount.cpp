// Algorithmic FP64 operation count of the oracle's l() path, "as written in
// source": + - * / sqrt pow each count 1, x.powi(n) counts n-1 multiplies,
// and operations whose operands are all compile-time constants of the
// reference (src/constant.rs) count 0.  Comparisons, fabs, fmin/fmax and
// copies count 0.
//
// TEST INFRASTRUCTURE: compiles oracle/rh3t_oracle.c unchanged with its
// floating-point type replaced by a counting type, runs the pointwise
// kernels on one interior stencil and prints the per-interface and
// per-cell-stage totals that bench.py's roofline uses.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

static long long g_ops = 0;

struct OpCount {
  double v;
  bool c;  // derived only from compile-time constants
  OpCount() = default;
  OpCount(double x) : v(x), c(true) {}
  OpCount(int x) : v(x), c(true) {}
  OpCount(long x) : v((double)x), c(true) {}
  static OpCount var(double x) {
    OpCount o(x);
    o.c = false;
    return o;
  }
  explicit operator double() const { return v; }
};

static inline OpCount mk(double v, bool c) {
  OpCount o(v);
  o.c = c;
  return o;
}
static inline OpCount bin(double v, const OpCount& a, const OpCount& b) {
  bool c = a.c && b.c;
  if (!c) ++g_ops;
  return mk(v, c);
}
static inline OpCount operator+(const OpCount& a, const OpCount& b) { return bin(a.v + b.v, a, b); }
static inline OpCount operator-(const OpCount& a, const OpCount& b) { return bin(a.v - b.v, a, b); }
static inline OpCount operator*(const OpCount& a, const OpCount& b) {
  // the 1.0 that starts the oracle's powi product is not an operation of
  // the reference (LLVM emits x for the first set bit)
  if ((a.c && a.v == 1.0) || (b.c && b.v == 1.0)) return mk(a.v * b.v, a.c && b.c);
  return bin(a.v * b.v, a, b);
}
static inline OpCount operator/(const OpCount& a, const OpCount& b) { return bin(a.v / b.v, a, b); }
static inline OpCount operator-(const OpCount& a) { return mk(-a.v, a.c); }
static inline OpCount& operator+=(OpCount& a, const OpCount& b) { return a = a + b; }
static inline OpCount& operator-=(OpCount& a, const OpCount& b) { return a = a - b; }
static inline OpCount& operator*=(OpCount& a, const OpCount& b) { return a = a * b; }
static inline bool operator<(const OpCount& a, const OpCount& b) { return a.v < b.v; }
static inline bool operator>(const OpCount& a, const OpCount& b) { return a.v > b.v; }
static inline bool operator<=(const OpCount& a, const OpCount& b) { return a.v <= b.v; }
static inline bool operator>=(const OpCount& a, const OpCount& b) { return a.v >= b.v; }
static inline bool operator==(const OpCount& a, const OpCount& b) { return a.v == b.v; }
static inline bool operator!=(const OpCount& a, const OpCount& b) { return a.v != b.v; }
static inline OpCount sqrt(const OpCount& a) {
  if (!a.c) ++g_ops;
  return mk(std::sqrt(a.v), a.c);
}
static inline OpCount pow(const OpCount& a, const OpCount& b) {
  if (!(a.c && b.c)) ++g_ops;
  return mk(std::pow(a.v, b.v), a.c && b.c);
}
static inline OpCount fabs(const OpCount& a) { return mk(std::fabs(a.v), a.c); }
static inline OpCount fma(const OpCount& a, const OpCount& b, const OpCount& c) {
  if (!(a.c && b.c && c.c)) g_ops += 2;
  return mk(std::fma(a.v, b.v, c.v), a.c && b.c && c.c);
}
static inline OpCount fmin(const OpCount& a, const OpCount& b) { return mk(std::fmin(a.v, b.v), a.c && b.c); }
static inline OpCount fmax(const OpCount& a, const OpCount& b) { return mk(std::fmax(a.v, b.v), a.c && b.c); }
static inline bool isfinite(const OpCount& a) { return std::isfinite(a.v); }

#define double OpCount
#define _OPENMP_DISABLED
#include "rh3t_oracle.c"
#undef double

int main() {
  ro_physics ph;
  ph.gamma_e = 1.4;
  ph.gamma_i = 1.4;
  ph.gamma_r = 1.4;
  ph.kappa_e = ph.kappa_i = ph.kappa_r = 0.0;
  ph.cve = ph.cvi = ph.a_rad = 1.0;
  ph.omega_ei = ph.omega_er = 0.0;
  ph.lambda = 0.5;
  ph.weno_q = 10.0;
  ph.eps = 1e-12;
  // a smooth, non-degenerate 9-point stencil of states
  OpCount pts[9][6];
  for (int q = 0; q < 9; ++q) {
    const double base[6] = {1.0, 0.3, -0.2, 2.6, 2.5, 2.7};
    const double slope[6] = {0.01, 0.02, -0.01, 0.03, 0.02, -0.02};
    for (int f = 0; f < 6; ++f) pts[q][f] = OpCount::var(base[f] + slope[f] * (q - 4) + 1e-3 * q * q * (f + 1));
  }
  OpCount dx = OpCount::var(1.0 / 2048.0);

  long long c_derived, c_flux, c_noncon, c_assembly, c_stage1, c_stage23, c_adm, c_dt;
  ro_derived d[9];
  g_ops = 0;
  ro_derived_from_state(&ph, pts[0], &d[0]);
  c_derived = g_ops;
  for (int q = 1; q < 9; ++q) ro_derived_from_state(&ph, pts[q], &d[q]);

  OpCount st6[36];
  for (int q = 0; q < 6; ++q)
    for (int f = 0; f < 6; ++f) st6[6 * q + f] = pts[q + 1][f];
  OpCount out[6];
  g_ops = 0;
  ro_reconstruction_fast(&ph, st6, d + 1, 0, 1, out);
  c_flux = g_ops;

  g_ops = 0;
  ro_noncon(d, dx, 0, out);
  c_noncon = g_ops;

  // assembly of main.rs:266-302 at default physics: two flux differences
  // scaled by 1/dx, 1/dy (per-cell reciprocals as written) and the add chain
  // fx + fy + nc_x + nc_y + src + dif over 6 components
  {
    OpCount fl[6], fr[6], fb[6], ft[6], nx_[6], ny_[6], z[6], r[6];
    for (int f = 0; f < 6; ++f) {
      fl[f] = OpCount::var(1.0 + f);
      fr[f] = OpCount::var(1.1 + f);
      fb[f] = OpCount::var(0.9 + f);
      ft[f] = OpCount::var(1.2 + f);
      nx_[f] = OpCount::var(0.01 * f);
      ny_[f] = OpCount::var(0.02 * f);
      z[f] = OpCount::var(0.0);
    }
    g_ops = 0;
    OpCount ix = OpCount(1.0) / dx, iy = OpCount(1.0) / dx;
    for (int f = 0; f < 6; ++f) {
      OpCount a = -(fr[f] - fl[f]) * ix;
      OpCount b = -(ft[f] - fb[f]) * iy;
      r[f] = a + b + nx_[f] + ny_[f] + z[f] + z[f];
    }
    c_assembly = g_ops;
  }
  // stage combinations, main.rs:307-349 (coefficients are per-call scalars)
  {
    OpCount base[6], add[6], rhs[6], v[6];
    for (int f = 0; f < 6; ++f) {
      base[f] = OpCount::var(1.0 + f);
      add[f] = OpCount::var(1.5 + f);
      rhs[f] = OpCount::var(0.1 * f);
    }
    OpCount dt = OpCount::var(1e-4);
    g_ops = 0;
    for (int f = 0; f < 6; ++f) v[f] = base[f] + rhs[f] * dt;
    c_stage1 = g_ops;
    OpCount c2 = dt / OpCount(4.0), c3 = OpCount(2.0) * dt / OpCount(3.0);
    g_ops = 0;
    for (int f = 0; f < 6; ++f) v[f] = base[f] * OpCount(0.75) + add[f] * OpCount(0.25) + rhs[f] * c2;
    for (int f = 0; f < 6; ++f) v[f] = base[f] * OpCount(1.0 / 3.0) + add[f] * OpCount(2.0 / 3.0) + rhs[f] * c3;
    c_stage23 = g_ops;  // both stages
    g_ops = 0;
    (void)admissible(v);
    c_adm = g_ops;
  }
  g_ops = 0;
  (void)ro_local_dt(&ph, pts[4], dx, dx);
  c_dt = g_ops;

  // per interior cell-stage: x and y interface fluxes (one of each per
  // cell), x and y noncon, derived, assembly, the mean stage combination and
  // the admissibility check
  double stage_mean = (c_stage1 + c_stage23) / 3.0;
  double cell_stage = 2.0 * c_flux + 2.0 * c_noncon + c_derived + c_assembly + stage_mean + c_adm;
  std::printf("{\"flux\": %lld, \"noncon\": %lld, \"derived\": %lld, \"assembly\": %lld, \"stage_mean\": %.3f, "
              "\"admissible\": %lld, \"local_dt\": %lld, \"cell_stage\": %.3f}\n",
              c_flux, c_noncon, c_derived, c_assembly, stage_mean, c_adm, c_dt, cell_stage);
  return 0;
}
