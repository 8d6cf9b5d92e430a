// Forward-mode (JVP) build of the oracle: rh3t_oracle.c compiled unchanged
// with its floating-point type replaced by a dual number (value, tangent).
//
// TEST INFRASTRUCTURE ONLY: the reference operator's directional
// derivative, against which the HIP reverse-mode kernels are checked
// (<w, J v> from here equals <J^T w, v> from the adjoint).  A dual compares
// by value, except that != 0 / == 0 also look at the tangent, so that a
// physics coefficient whose value is 0 but whose tangent is not (a
// derivative with respect to kappa or omega at 0) still switches its term on
// (the forward skips zero coefficients, src/constant.rs:22-25).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

struct Dual {
  double v, d;
  Dual() = default;
  Dual(double x) : v(x), d(0.0) {}
  Dual(int x) : v(x), d(0.0) {}
  Dual(long x) : v((double)x), d(0.0) {}
  Dual(double x, double t) : v(x), d(t) {}
  explicit operator double() const { return v; }
};

static inline Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
static inline Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
static inline Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
static inline Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
static inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
static inline Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
static inline Dual& operator-=(Dual& a, const Dual& b) { return a = a - b; }
static inline Dual& operator*=(Dual& a, const Dual& b) { return a = a * b; }
static inline Dual& operator/=(Dual& a, const Dual& b) { return a = a / b; }
static inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
static inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
static inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
static inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
static inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v && a.d == b.d; }
static inline bool operator!=(const Dual& a, const Dual& b) { return !(a == b); }
static inline Dual sqrt(const Dual& a) {
  const double r = std::sqrt(a.v);
  return Dual(r, a.d * 0.5 / r);
}
static inline Dual pow(const Dual& a, const Dual& b) {
  const double p = std::pow(a.v, b.v);
  double t = a.d == 0.0 ? 0.0 : b.v * std::pow(a.v, b.v - 1.0) * a.d;
  if (b.d != 0.0) t += std::log(a.v) * p * b.d;
  return Dual(p, t);
}
static inline Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
static inline Dual fma(const Dual& a, const Dual& b, const Dual& c) {
  return Dual(std::fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d);
}
// the argument the value takes (fmin/fmax of src/noncon.rs:288-332, dt.rs)
static inline Dual fmin(const Dual& a, const Dual& b) { return b.v < a.v ? b : a; }
static inline Dual fmax(const Dual& a, const Dual& b) { return b.v > a.v ? b : a; }
static inline bool isfinite(const Dual& a) { return std::isfinite(a.v) && std::isfinite(a.d); }

#define double Dual
#include "rh3t_oracle.c"
#undef double

// ------------------------------------------------------------- bridge

namespace {

// ro_problem / ro_physics with plain doubles (the layout of rh3t_oracle.h)
struct plain_physics {
  double v[14];
};
struct plain_problem {
  int32_t nx, ny;
  double dx, dy;
  const uint8_t* fluid;
  const uint32_t* access;
  int32_t n_ghost;
  const int32_t* g_kind;
  const int32_t* g_src;
  const int32_t* g_heavy;
  const double* g_normal;
  const double* g_dist;
  const double* g_bcstate;
  const double* g_outflow;
  const int32_t* h_nearest;
  const double* h_d;
  const uint32_t* h_stencil;
  const double* h_s;
  const double* h_w;
  double time;
  const double* g_p0x;
  const double* g_bcstate2;
  const double* g_front;
};

struct DualProblem {
  ro_problem p;
  std::vector<Dual> normal, dist, bcstate, outflow, hd, hs, hw, p0x, bcstate2, front;
};

std::vector<Dual> lift(const double* x, size_t n) {
  std::vector<Dual> r(n);
  for (size_t k = 0; k < n; ++k) r[k] = Dual(x ? x[k] : 0.0);
  return r;
}

void make_problem(const plain_problem* q, DualProblem& D) {
  int n_heavy = 0;
  for (int g = 0; g < q->n_ghost; ++g)
    if (q->g_heavy[g] >= 0 && q->g_heavy[g] + 1 > n_heavy) n_heavy = q->g_heavy[g] + 1;
  const size_t ng = (size_t)(q->n_ghost > 0 ? q->n_ghost : 1);
  D.normal = lift(q->g_normal, 2 * ng);
  D.dist = lift(q->g_dist, ng);
  D.bcstate = lift(q->g_bcstate, 6 * ng);
  D.outflow = lift(q->g_outflow, 3 * ng);
  D.p0x = lift(q->g_p0x, ng);
  D.bcstate2 = lift(q->g_bcstate2, 6 * ng);
  D.front = lift(q->g_front, 2 * ng);
  const size_t nh = (size_t)(n_heavy > 0 ? n_heavy : 1);
  D.hd = lift(n_heavy ? q->h_d : nullptr, 5 * nh);
  D.hs = lift(n_heavy ? q->h_s : nullptr, 979 * nh);
  D.hw = lift(n_heavy ? q->h_w : nullptr, 225 * nh);
  ro_problem& p = D.p;
  p.nx = q->nx;
  p.ny = q->ny;
  p.dx = Dual(q->dx);
  p.dy = Dual(q->dy);
  p.fluid = q->fluid;
  p.access = q->access;
  p.n_ghost = q->n_ghost;
  p.g_kind = q->g_kind;
  p.g_src = q->g_src;
  p.g_heavy = q->g_heavy;
  p.g_normal = D.normal.data();
  p.g_dist = D.dist.data();
  p.g_bcstate = D.bcstate.data();
  p.g_outflow = D.outflow.data();
  p.h_nearest = q->h_nearest;
  p.h_d = D.hd.data();
  p.h_stencil = q->h_stencil;
  p.h_s = D.hs.data();
  p.h_w = D.hw.data();
  p.time = Dual(q->time);
  p.g_p0x = D.p0x.data();
  p.g_bcstate2 = D.bcstate2.data();
  p.g_front = D.front.data();
}

ro_physics make_physics(const plain_physics* ph, const plain_physics* dph) {
  Dual f[14];
  for (int k = 0; k < 14; ++k) f[k] = Dual(ph->v[k], dph ? dph->v[k] : 0.0);
  ro_physics r;
  static_assert(sizeof(ro_physics) == 14 * sizeof(Dual), "ro_physics layout");
  std::memcpy(&r, f, sizeof(r));
  return r;
}

}  // namespace

extern "C" {

// rhs + eps drhs = l(u + eps du; physics + eps dphysics) to first order.
int rod_l_jvp(const plain_problem* q, const plain_physics* ph, const plain_physics* dph, const double* u,
              const double* du, double* rhs, double* drhs, int nthreads) {
  DualProblem D;
  make_problem(q, D);
  const ro_physics P = make_physics(ph, dph);
  const size_t n = 6 * (size_t)q->nx * q->ny;
  std::vector<Dual> U(n), R(n);
  for (size_t k = 0; k < n; ++k) U[k] = Dual(u[k], du ? du[k] : 0.0);
  ro_error err;
  const int st = ro_l(&D.p, &P, U.data(), R.data(), nullptr, nthreads, &err);
  for (size_t k = 0; k < n; ++k) {
    rhs[k] = R[k].v;
    drhs[k] = R[k].d;
  }
  return st;
}

// nsteps rk3_ssp steps with the frozen dt[n]: u and du are advanced in place.
int rod_rk3_jvp(const plain_problem* q, const plain_physics* ph, const plain_physics* dph, double* u, double* du,
                const double* dt, int nsteps, int nthreads) {
  DualProblem D;
  make_problem(q, D);
  const ro_physics P = make_physics(ph, dph);
  const size_t n = 6 * (size_t)q->nx * q->ny;
  std::vector<Dual> U(n);
  for (size_t k = 0; k < n; ++k) U[k] = Dual(u[k], du[k]);
  Dual t(0.0);
  ro_error err;
  int st = 0;
  for (int s = 0; s < nsteps && st == 0; ++s) st = ro_rk3_step(&D.p, &P, U.data(), &t, Dual(dt[s]), nthreads, &err);
  for (size_t k = 0; k < n; ++k) {
    u[k] = U[k].v;
    du[k] = U[k].d;
  }
  return st;
}

}  // extern "C"
