// Field / GhostGrid construction on the host.  See problem.hpp.
//
// Reference anchors (colinhu123/3T_RH):
//   geometry      src/geometry.rs:1-673
//   Field         src/field1.rs:5-377
//   GhostGrid     src/ghost.rs:50-118, 284-493
//   GhostBC       src/bc1.rs:45-188, 257-281, 930-1037, 1070-1549
//   init_cylinder src/init.rs:7-24, 116-510
#include "problem.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <set>

namespace rh3t {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kArcAngleTol = 1e-12;  // src/geometry.rs:430
constexpr double kDefaultEps = 1e-12;   // src/constant.rs:1

// x.powi(n) for n >= 0 as LLVM / compiler-builtins evaluate it: repeated
// squaring, multiplying the running product by the current square for every
// set bit (1.0 * x is exact, so both lowerings agree).
inline double powi(double x, int n) {
  double mul = 1.0;
  double a = x;
  for (;;) {
    if (n & 1) mul *= a;
    n >>= 1;
    if (n == 0) break;
    a *= a;
  }
  return mul;
}

// f64::rem_euclid
inline double rem_euclid(double a, double b) {
  double r = std::fmod(a, b);
  return r < 0.0 ? r + std::fabs(b) : r;
}

inline int64_t rem_euclid_i(int64_t a, int64_t b) {
  int64_t r = a % b;
  return r < 0 ? r + (b < 0 ? -b : b) : r;
}

inline double normalize_angle(double theta) { return rem_euclid(theta, 2.0 * kPi); }

inline double cross(Point a, Point b) { return a.x * b.y - a.y * b.x; }

bool point_on_segment(Point p, Point a, Point b, double eps) {
  Point ab{b.x - a.x, b.y - a.y};
  Point ap{p.x - a.x, p.y - a.y};
  if (std::fabs(cross(ab, ap)) > eps) return false;
  return p.x >= std::fmin(a.x, b.x) - eps && p.x <= std::fmax(a.x, b.x) + eps &&
         p.y >= std::fmin(a.y, b.y) - eps && p.y <= std::fmax(a.y, b.y) + eps;
}

Point closest_point_on_segment(Point p, Point a, Point b) {
  double abx = b.x - a.x, aby = b.y - a.y;
  double apx = p.x - a.x, apy = p.y - a.y;
  double ab2 = abx * abx + aby * aby;
  if (ab2 <= 1e-30) return a;
  double t = (apx * abx + apy * aby) / ab2;
  // f64::clamp(0.0, 1.0)
  if (t < 0.0) t = 0.0;
  if (t > 1.0) t = 1.0;
  return Point{a.x + t * abx, a.y + t * aby};
}

inline double distance2(Point a, Point b) {
  double dx = a.x - b.x, dy = a.y - b.y;
  return dx * dx + dy * dy;
}

[[noreturn]] void fail(int status, const std::string& msg) {
  throw BuildError{status, msg};
}

}  // namespace

// ---------------------------------------------------------------- Polygon

double Polygon::signed_area2() const {  // src/geometry.rs:160-175
  size_t n = points.size();
  double area2 = 0.0;
  for (size_t i = 0; i < n; ++i) {
    Point a = points[i], b = points[(i + 1) % n];
    area2 += a.x * b.y - b.x * a.y;
  }
  return area2;
}

bool Polygon::is_fluid(Point p) const {  // src/geometry.rs:254-278
  size_t n = points.size();
  if (n < 3) return false;
  bool inside = false;
  for (size_t i = 0; i < n; ++i) {
    Point a = points[i], b = points[(i + 1) % n];
    bool crosses = (a.y > p.y) != (b.y > p.y);
    if (crosses) {
      double x_intersect = a.x + (p.y - a.y) * (b.x - a.x) / (b.y - a.y);
      if (p.x < x_intersect) inside = !inside;
    }
  }
  return fluid == FluidSide::Inside ? inside : !inside;
}

Point Polygon::closest_point(Point p) const {  // src/geometry.rs:280-302
  size_t n = points.size();
  if (n < 2) fail(RH3T_ERR_GEOMETRY, "Polygon must have at least two points");
  Point closest = points[0];
  double min_dist2 = std::numeric_limits<double>::infinity();
  for (size_t i = 0; i < n; ++i) {
    Point c = closest_point_on_segment(p, points[i], points[(i + 1) % n]);
    double d2 = distance2(p, c);
    if (d2 < min_dist2) {
      min_dist2 = d2;
      closest = c;
    }
  }
  return closest;
}

Vec2 Polygon::normal(Point p0) const {  // src/geometry.rs:304-420
  size_t n = points.size();
  if (n < 2) fail(RH3T_ERR_GEOMETRY, "Polygon must have at least two points");
  double area2 = signed_area2();
  if (!(std::fabs(area2) > 1e-14)) fail(RH3T_ERR_GEOMETRY, "Degenerate polygon has zero area");
  bool ccw = area2 > 0.0;
  size_t best_edge = 0;
  double min_dist2 = std::numeric_limits<double>::infinity();
  for (size_t i = 0; i < n; ++i) {
    Point c = closest_point_on_segment(p0, points[i], points[(i + 1) % n]);
    double d2 = distance2(p0, c);
    if (d2 < min_dist2) {
      min_dist2 = d2;
      best_edge = i;
    }
  }
  Point a = points[best_edge], b = points[(best_edge + 1) % n];
  double tx = b.x - a.x, ty = b.y - a.y;
  double tn = std::hypot(tx, ty);
  if (!(tn > 1e-14)) fail(RH3T_ERR_GEOMETRY, "Degenerate polygon edge");
  Vec2 left{-ty / tn, tx / tn};
  Vec2 right{ty / tn, -tx / tn};
  if (fluid == FluidSide::Inside) return ccw ? right : left;
  return ccw ? left : right;
}

Vec2 Polygon::outward_normal_of_side(size_t side) const {  // src/geometry.rs:212-250
  size_t n = points.size();
  Point a = points[side], b = points[(side + 1) % n];
  double tx = b.x - a.x, ty = b.y - a.y;
  double len = std::hypot(tx, ty);
  if (!(len > 1.0e-14)) fail(RH3T_ERR_GEOMETRY, "degenerate polygon side");
  Vec2 left{-ty / len, tx / len};
  Vec2 right{ty / len, -tx / len};
  bool ccw = signed_area2() > 0.0;
  if (fluid == FluidSide::Inside) return ccw ? right : left;
  return ccw ? left : right;
}

Projection project(const Polygon& poly, Point p) {  // src/geometry.rs:121-138
  Point p0 = poly.closest_point(p);
  Vec2 n = poly.normal(p0);
  double dx = p.x - p0.x, dy = p.y - p0.y;
  return Projection{p0, n, dx * n.x + dy * n.y};
}

std::vector<size_t> find_boundary_sides(Point p, const Polygon& poly, double eps) {
  // src/geometry.rs:35-55
  std::vector<size_t> sides;
  size_t n = poly.points.size();
  for (size_t i = 0; i < n; ++i)
    if (point_on_segment(p, poly.points[i], poly.points[(i + 1) % n], eps)) sides.push_back(i);
  return sides;
}

// ------------------------------------------------------------ CircularArc

CircularArc CircularArc::from_three_points(Point s, Point m, Point e, FluidSide f) {
  // src/geometry.rs:493-551
  double x1 = s.x, y1 = s.y, x2 = m.x, y2 = m.y, x3 = e.x, y3 = e.y;
  double d = 2.0 * (x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
  double chord_scale = 0.0;
  chord_scale = std::fmax(chord_scale, std::hypot(x1 - x2, y1 - y2));
  chord_scale = std::fmax(chord_scale, std::hypot(x2 - x3, y2 - y3));
  chord_scale = std::fmax(chord_scale, std::hypot(x1 - x3, y1 - y3));
  chord_scale = std::fmax(chord_scale, 1.0);
  if (!(std::fabs(d) > 1e-12 * chord_scale * chord_scale))
    fail(RH3T_ERR_GEOMETRY, "CircularArc::from_three_points: points are (nearly) collinear");
  double r1 = x1 * x1 + y1 * y1;
  double r2 = x2 * x2 + y2 * y2;
  double r3 = x3 * x3 + y3 * y3;
  double ux = (r1 * (y2 - y3) + r2 * (y3 - y1) + r3 * (y1 - y2)) / d;
  double uy = (r1 * (x3 - x2) + r2 * (x1 - x3) + r3 * (x2 - x1)) / d;
  CircularArc a;
  a.center = Point{ux, uy};
  a.radius = std::hypot(x1 - ux, y1 - uy);
  if (!(a.radius > 0.0)) fail(RH3T_ERR_GEOMETRY, "zero-radius arc");
  double theta_start = std::atan2(y1 - uy, x1 - ux);
  double theta_mid = std::atan2(y2 - uy, x2 - ux);
  double theta_end = std::atan2(y3 - uy, x3 - ux);
  double ccw_sweep = normalize_angle(theta_end - theta_start);
  double ccw_to_mid = normalize_angle(theta_mid - theta_start);
  double sweep = (ccw_to_mid <= ccw_sweep + kArcAngleTol) ? ccw_sweep
                                                          : -normalize_angle(theta_start - theta_end);
  if (!(std::fabs(sweep) > kArcAngleTol)) fail(RH3T_ERR_GEOMETRY, "CircularArc sweep must be nonzero");
  if (!(std::fabs(sweep) <= 2.0 * kPi + kArcAngleTol))
    fail(RH3T_ERR_GEOMETRY, "CircularArc sweep must not exceed 2*pi");
  a.theta_start = theta_start;
  a.sweep = sweep;
  a.fluid = f;
  return a;
}

bool CircularArc::contains_angle(double theta) const {  // src/geometry.rs:554-562
  if (sweep > 0.0) return normalize_angle(theta - theta_start) <= sweep + kArcAngleTol;
  return normalize_angle(theta_start - theta) <= -sweep + kArcAngleTol;
}

Point CircularArc::start_point() const {
  return Point{center.x + radius * std::cos(theta_start), center.y + radius * std::sin(theta_start)};
}

Point CircularArc::end_point() const {
  double te = theta_start + sweep;
  return Point{center.x + radius * std::cos(te), center.y + radius * std::sin(te)};
}

Point CircularArc::closest_point(Point p) const {  // src/geometry.rs:581-607
  double rx = p.x - center.x, ry = p.y - center.y;
  double rnorm = std::hypot(rx, ry);
  if (rnorm > kArcAngleTol) {
    double theta = std::atan2(ry, rx);
    if (contains_angle(theta)) {
      double s = radius / rnorm;
      return Point{center.x + s * rx, center.y + s * ry};
    }
  }
  Point st = start_point(), en = end_point();
  double d2s = powi(p.x - st.x, 2) + powi(p.y - st.y, 2);
  double d2e = powi(p.x - en.x, 2) + powi(p.y - en.y, 2);
  return d2s <= d2e ? st : en;
}

Vec2 CircularArc::outward_normal(Point p0) const {  // src/geometry.rs:613-635
  double dx = p0.x - center.x, dy = p0.y - center.y;
  double nrm = std::hypot(dx, dy);
  if (!(nrm > 1e-14)) fail(RH3T_ERR_GEOMETRY, "point coincides with arc center");
  Vec2 radial{dx / nrm, dy / nrm};
  if (fluid == FluidSide::Inside) return radial;
  return Vec2{-radial.x, -radial.y};
}

Projection CircularArc::project(Point p) const {  // src/geometry.rs:639-653
  Point p0 = closest_point(p);
  Vec2 n = outward_normal(p0);
  double dx = p.x - p0.x, dy = p.y - p0.y;
  return Projection{p0, n, dx * n.x + dy * n.y};
}

// ----------------------------------------------------------------- Field

void Problem::coord2idx(Point p, int64_t* i, int64_t* j) const {  // src/field1.rs:68-72
  *i = (int64_t)std::round((p.x - grid.x0) / grid.dx);
  *j = (int64_t)std::round((p.y - grid.y0) / grid.dy);
}

bool Problem::is_in_domain(int64_t i, int64_t j) const {  // src/field1.rs:179-192
  if (!grid_in_domain(i, j)) {
    Point p{x(i), y(j)};
    return outer.is_fluid(p) && inner.is_fluid(p);
  }
  return fluid[(size_t)i * grid.ny + j] != 0;
}

const CircularArc* Problem::arc_for_side(int boundary, size_t side) const {
  const auto& arcs = boundary == 0 ? outer_arcs : inner_arcs;
  for (const auto& ov : arcs)
    if (ov.contains_side(side)) return &ov.arc;
  return nullptr;
}

void Problem::build_fluid_mask() {  // src/field1.rs:112-122
  size_t nx = grid.nx, ny = grid.ny;
  fluid.assign(nx * ny, 0);
  for (size_t l = 0; l < nx * ny; ++l) {
    size_t i = l / ny, j = l % ny;
    Point p{x((int64_t)i), y((int64_t)j)};
    fluid[l] = (outer.is_fluid(p) && inner.is_fluid(p)) ? 1 : 0;
  }
}

// ------------------------------------------------------------- GhostGrid

namespace {

int bc_priority(int kind) {  // src/ghost.rs:284-309
  switch (kind) {
    case RH3T_BC_WALL: return 100;
    case RH3T_BC_REFLECTIVE_WALL: return 90;
    case RH3T_BC_CONSTANT: return 70;
    case RH3T_BC_TIME_DEPENDENT: return 70;
    case RH3T_BC_FARFIELD: return 20;
    case RH3T_BC_OUTFLOW: return 10;
    case RH3T_BC_ZEROTH_ORDER: return 5;
    case RH3T_BC_PERIODIC: return 0;
    default: return 0;
  }
}

size_t select_boundary_side(Point p0, const Polygon& poly, const std::vector<rh3t_bc>& bcs) {
  // src/ghost.rs:311-348
  auto cand = find_boundary_sides(p0, poly, kDefaultEps);
  if (cand.empty()) fail(RH3T_ERR_GEOMETRY, "projection point does not lie on any polygon side");
  if (poly.points.size() != bcs.size()) fail(RH3T_ERR_GEOMETRY, "polygon side count and BC count disagree");
  size_t best = cand[0];
  int best_pr = bc_priority(bcs[best].kind);
  for (size_t k = 1; k < cand.size(); ++k) {
    int pr = bc_priority(bcs[cand[k]].kind);
    if (pr > best_pr) {
      best = cand[k];
      best_pr = pr;
    }
  }
  return best;
}

// find_nearest_grid_point, src/bc1.rs:257-281
void find_nearest_grid_point(const Problem& pb, const Projection& pr, int32_t* oi, int32_t* oj) {
  int64_t i, j;
  pb.coord2idx(pr.point, &i, &j);
  const int64_t cand[9][2] = {{i, j},         {i - 1, j},     {i + 1, j},
                              {i, j - 1},     {i, j + 1},     {i + 1, j + 1},
                              {i + 1, j - 1}, {i - 1, j + 1}, {i - 1, j - 1}};
  int64_t ti = 0, tj = 0;
  double td = 1e50;
  for (int k = 0; k < 9; ++k) {
    if (pb.is_in_domain(cand[k][0], cand[k][1])) {
      double dx = pb.x(cand[k][0]) - pr.point.x;
      double dy = pb.y(cand[k][1]) - pr.point.y;
      double dis = std::sqrt(powi(dx, 2) + powi(dy, 2));
      if (dis < td) {
        ti = cand[k][0];
        tj = cand[k][1];
        td = dis;
      }
    }
  }
  *oi = (int32_t)ti;
  *oj = (int32_t)tj;
}

// Mirror-point source of reflective_wall_value / zeroth_order_value,
// src/bc1.rs:930-1003 and 471-533 (identical candidate order).
bool mirror_source(const Problem& pb, int64_t gi, int64_t gj, const Projection& pr, int32_t* si,
                   int32_t* sj) {
  Point pg{pb.x(gi), pb.y(gj)};
  Point pi{2.0 * pr.point.x - pg.x, 2.0 * pr.point.y - pg.y};
  int64_t ic, jc;
  pb.coord2idx(pi, &ic, &jc);
  const int64_t cand[9][2] = {{ic, jc},         {ic - 1, jc},     {ic + 1, jc},
                              {ic, jc - 1},     {ic, jc + 1},     {ic - 1, jc - 1},
                              {ic - 1, jc + 1}, {ic + 1, jc - 1}, {ic + 1, jc + 1}};
  bool found = false;
  double best = std::numeric_limits<double>::infinity();
  for (int k = 0; k < 9; ++k) {
    int64_t qi = cand[k][0], qj = cand[k][1];
    if (!pb.is_in_domain(qi, qj)) continue;
    double d2 = powi(pb.x(qi) - pi.x, 2) + powi(pb.y(qj) - pi.y, 2);
    if (d2 < best) {
      best = d2;
      *si = (int32_t)qi;
      *sj = (int32_t)qj;
      found = true;
    }
  }
  return found;
}

// ---- least-squares tables (Tan et al. 2012, Sec. 2.4) ------------------

double factorial(int n) {
  double p = 1.0;
  for (int k = 1; k <= n; ++k) p *= (double)k;
  return p;
}

int xk_coefficient_column(int k) {  // src/bc1.rs:89-101
  int col = 0;
  for (int t = 0; t <= k; ++t)
    for (int i2 = 0; i2 <= t; ++i2) {
      int j2 = t - i2;
      if (i2 == k && j2 == 0) return col;
      ++col;
    }
  return -1;
}

struct Mono {
  int i, j;
  double c;
};

// derivative_coeffs, src/bc1.rs:1119-1160
std::vector<Mono> derivative_coeffs(const std::vector<double>& coef, int degree, int ax, int ay) {
  std::vector<Mono> out;
  int k = 0;
  for (int total = 0; total <= degree; ++total)
    for (int i = 0; i <= total; ++i) {
      int j = total - i;
      double aij = coef[k++];
      if (i < ax || j < ay) continue;
      double factor = 1.0;
      for (int m = 0; m < ax; ++m) factor *= (double)(i - m);
      for (int m = 0; m < ay; ++m) factor *= (double)(j - m);
      out.push_back(Mono{i - ax, j - ay, aij * factor});
    }
  return out;
}

double integrate_1d_monomial(int power, double h) {  // src/bc1.rs:1185-1196
  if (power % 2 == 1) return 0.0;
  double half = 0.5 * h;
  return 2.0 * powi(half, power + 1) / (double)(power + 1);
}

double integrate_square(const std::vector<Mono>& poly, double h) {  // src/bc1.rs:1162-1183
  double result = 0.0;
  for (const auto& a : poly)
    for (const auto& b : poly) {
      double ix = integrate_1d_monomial(a.i + b.i, h);
      double iy = integrate_1d_monomial(a.j + b.j, h);
      result += a.c * b.c * ix * iy;
    }
  return result;
}

double smooth_indicator(const std::vector<double>& coef, int r, double h) {  // src/bc1.rs:1070-1117
  if (r == 0) return 2.0 * h * h;
  double beta = 0.0;
  for (int ax = 0; ax <= r; ++ax)
    for (int ay = 0; ay <= r - ax; ++ay) {
      int order = ax + ay;
      if (order == 0) continue;
      auto der = derivative_coeffs(coef, r, ax, ay);
      double integral = integrate_square(der, h);
      double scale = powi(h, 2 * (order - 1));
      beta += scale * integral;
    }
  return beta;
}

// beta_quadratic_form, src/bc1.rs:62-81
std::vector<double> beta_quadratic_form(int r, double h) {
  int n = (r + 1) * (r + 2) / 2;
  std::vector<double> m((size_t)n * n, 0.0);
  for (int a = 0; a < n; ++a)
    for (int b = a; b < n; ++b) {
      std::vector<double> ea(n, 0.0), eb(n, 0.0), eab(n, 0.0);
      ea[a] = 1.0;
      eb[b] = 1.0;
      for (int q = 0; q < n; ++q) eab[q] = ea[q] + eb[q];
      double v = 0.5 * (smooth_indicator(eab, r, h) - smooth_indicator(ea, r, h) -
                        smooth_indicator(eb, r, h));
      m[(size_t)a * n + b] = v;
      m[(size_t)b * n + a] = v;
    }
  return m;
}

// Moore-Penrose pseudo-inverse of a rows x cols matrix (rows >= cols) by a
// one-sided (Hestenes) Jacobi SVD carried in long double, singular values
// <= tol treated as zero (nalgebra SVD::pseudo_inverse(1e-12) semantics).
// Output is cols x rows, row-major.
std::vector<double> pseudo_inverse(const std::vector<double>& a, int rows, int cols, double tol) {
  typedef long double LD;
  std::vector<LD> w((size_t)rows * cols), v((size_t)cols * cols, 0.0L);
  for (size_t q = 0; q < w.size(); ++q) w[q] = a[q];
  for (int c = 0; c < cols; ++c) v[(size_t)c * cols + c] = 1.0L;
  for (int sweep = 0; sweep < 100; ++sweep) {
    LD off = 0.0L;
    for (int p = 0; p < cols - 1; ++p)
      for (int q = p + 1; q < cols; ++q) {
        LD alpha = 0, beta = 0, gamma = 0;
        for (int r = 0; r < rows; ++r) {
          LD wp = w[(size_t)r * cols + p], wq = w[(size_t)r * cols + q];
          alpha += wp * wp;
          beta += wq * wq;
          gamma += wp * wq;
        }
        if (gamma == 0.0L) continue;
        LD rel = std::fabs(gamma) / std::sqrt(alpha * beta);
        if (rel > off) off = rel;
        if (rel < 1e-30L) continue;
        LD zeta = (beta - alpha) / (2.0L * gamma);
        LD t = (zeta >= 0 ? 1.0L : -1.0L) / (std::fabs(zeta) + std::sqrt(1.0L + zeta * zeta));
        LD c = 1.0L / std::sqrt(1.0L + t * t), s = c * t;
        for (int r = 0; r < rows; ++r) {
          LD wp = w[(size_t)r * cols + p], wq = w[(size_t)r * cols + q];
          w[(size_t)r * cols + p] = c * wp - s * wq;
          w[(size_t)r * cols + q] = s * wp + c * wq;
        }
        for (int r = 0; r < cols; ++r) {
          LD vp = v[(size_t)r * cols + p], vq = v[(size_t)r * cols + q];
          v[(size_t)r * cols + p] = c * vp - s * vq;
          v[(size_t)r * cols + q] = s * vp + c * vq;
        }
      }
    if (off < 1e-19L) break;
  }
  // sigma_c = |w_c|, u_c = w_c / sigma_c;  pinv = sum_c v_c u_c^T / sigma_c
  std::vector<LD> acc((size_t)cols * rows, 0.0L);
  for (int c = 0; c < cols; ++c) {
    LD s2 = 0;
    for (int r = 0; r < rows; ++r) s2 += w[(size_t)r * cols + c] * w[(size_t)r * cols + c];
    LD sigma = std::sqrt(s2);
    if (!(sigma > (LD)tol)) continue;
    LD inv2 = 1.0L / (sigma * sigma);  // v_c (w_c/sigma)^T / sigma
    for (int i = 0; i < cols; ++i) {
      LD vi = v[(size_t)i * cols + c] * inv2;
      for (int r = 0; r < rows; ++r) acc[(size_t)i * rows + r] += vi * w[(size_t)r * cols + c];
    }
  }
  std::vector<double> out(acc.size());
  for (size_t q = 0; q < acc.size(); ++q) out[q] = (double)acc[q];
  return out;
}

int64_t first_interior_i(const Problem& pb, const Projection& pr) {  // src/bc1.rs:1328-1350
  double q = (pr.point.x - pb.grid.x0) / pb.grid.dx;
  int64_t i = pr.normal.x > 0.0 ? (int64_t)std::ceil(q) - 1 : (int64_t)std::floor(q) + 1;
  return std::clamp<int64_t>(i, 0, pb.grid.nx - 1);
}

int64_t first_interior_j(const Problem& pb, const Projection& pr) {  // src/bc1.rs:1306-1326
  double q = (pr.point.y - pb.grid.y0) / pb.grid.dy;
  int64_t j = pr.normal.y > 0.0 ? (int64_t)std::ceil(q) - 1 : (int64_t)std::floor(q) + 1;
  return std::clamp<int64_t>(j, 0, pb.grid.ny - 1);
}

// shifted_{vertical,horizontal}_substencil_indices, src/bc1.rs:1489-1549
bool shifted_substencil(const Problem& pb, bool vertical, int64_t line, int64_t base, int64_t width,
                        std::vector<std::pair<int64_t, int64_t>>* out) {
  int64_t max_shift = (vertical ? pb.grid.ny : pb.grid.nx) + width;
  for (int64_t mag = 0; mag <= max_shift; ++mag) {
    int64_t shifts[2] = {-mag, mag};
    int count = mag == 0 ? 1 : 2;
    for (int s = 0; s < count; ++s) {
      int64_t start = base + shifts[s];
      bool ok = true;
      std::vector<std::pair<int64_t, int64_t>> idx;
      for (int64_t a = 0; a < width; ++a) {
        int64_t i = vertical ? line : start + a;
        int64_t j = vertical ? start + a : line;
        if (!pb.grid_in_domain(i, j) || !pb.is_in_domain(i, j)) {
          ok = false;
          break;
        }
        idx.emplace_back(i, j);
      }
      if (ok) {
        out->insert(out->end(), idx.begin(), idx.end());
        return true;
      }
    }
  }
  return false;
}

// weno_stencil_extractor_indices, src/bc1.rs:1374-1485
std::vector<std::pair<int64_t, int64_t>> weno_stencil_extractor_indices(const Problem& pb,
                                                                        const Projection& pr, int r) {
  Vec2 n = pr.normal;
  Point p0 = pr.point;
  const double dx = pb.grid.dx, dy = pb.grid.dy, x0 = pb.grid.x0, y0 = pb.grid.y0;
  int64_t width = r + 1;
  std::vector<std::pair<int64_t, int64_t>> e_r;
  if (std::fabs(n.x) >= std::fabs(n.y)) {
    if (!(std::fabs(n.x) > 1.0e-14)) fail(RH3T_ERR_GEOMETRY, "E_r: normal has no x component");
    int64_t i0 = first_interior_i(pb, pr);
    int64_t step = pr.normal.x > 0.0 ? -1 : 1;
    for (int l = 0; l <= r; ++l) {
      int64_t i_line = i0 + (int64_t)l * step;
      double x_line = pb.x(i_line);
      double t = (p0.x - x_line) / n.x;
      double y_star = p0.y - n.y * t;
      int64_t m = (int64_t)std::floor((y_star - y0) / dy);
      int64_t base = m - (r / 2);
      if (!shifted_substencil(pb, true, i_line, base, width, &e_r))
        fail(RH3T_ERR_GEOMETRY, "cannot fit paper substencil on a vertical grid line");
    }
  } else {
    if (!(std::fabs(n.y) > 1.0e-14)) fail(RH3T_ERR_GEOMETRY, "E_r: normal has no y component");
    int64_t j0 = first_interior_j(pb, pr);
    int64_t step = pr.normal.y > 0.0 ? -1 : 1;
    for (int l = 0; l <= r; ++l) {
      int64_t j_line = j0 + (int64_t)l * step;
      double y_line = pb.y(j_line);
      double t = (p0.y - y_line) / n.y;
      double x_star = p0.x - n.x * t;
      int64_t m = (int64_t)std::floor((x_star - x0) / dx);
      int64_t base = m - (r / 2);
      if (!shifted_substencil(pb, false, j_line, base, width, &e_r))
        fail(RH3T_ERR_GEOMETRY, "cannot fit paper substencil on a horizontal grid line");
    }
  }
  if ((int)e_r.size() != (r + 1) * (r + 1)) fail(RH3T_ERR_GEOMETRY, "E_r cardinality");
  return e_r;
}

// precompute_ghost_bc, src/bc1.rs:108-188
GhostBC precompute_ghost_bc(const Problem& pb, const Projection& pr,
                            const std::vector<std::vector<double>>& beta_forms) {
  GhostBC bc;
  std::memset(&bc, 0, sizeof(bc));
  find_nearest_grid_point(pb, pr, &bc.nearest_i, &bc.nearest_j);
  double h = std::sqrt(pb.grid.dx * pb.grid.dy);
  double d0 = 2.0 * powi(h, 4), d1 = 2.0 * powi(h, 3), d2 = 2.0 * powi(h, 2), d3 = 2.0 * h;
  bc.normal = pr.normal;
  bc.d[0] = d0;
  bc.d[1] = d1;
  bc.d[2] = d2;
  bc.d[3] = d3;
  bc.d[4] = 1.0 - d0 - d1 - d2 - d3;
  for (int r = 0; r <= 4; ++r) {
    int m = (r + 1) * (r + 1);
    int nt = (r + 1) * (r + 2) / 2;
    auto idxs = weno_stencil_extractor_indices(pb, pr, r);
    std::vector<double> a((size_t)m * nt);
    for (int row = 0; row < m; ++row) {
      int64_t gi = idxs[row].first, gj = idxs[row].second;
      bc.stencil[kStencilOff[r] + row] = (uint32_t)(gi * pb.grid.ny + gj);
      // Projection::gloabl2local_coord, src/geometry.rs:99-110
      double pdx = pb.x(gi) - pr.point.x, pdy = pb.y(gj) - pr.point.y;
      double nor = pr.normal.x * pdx + pr.normal.y * pdy;
      double tan = -pdx * pr.normal.y + pdy * pr.normal.x;
      int col = 0;
      for (int t = 0; t <= r; ++t)
        for (int i2 = 0; i2 <= t; ++i2) {
          int j2 = t - i2;
          a[(size_t)row * nt + col] = powi(nor, i2) * powi(tan, j2);
          ++col;
        }
    }
    auto pinv = pseudo_inverse(a, m, nt, 1.0e-12);  // nt x m
    if (r == 0) {
      bc.s[0] = 2.0 * h * h;
    } else {
      const auto& mr = beta_forms[r];
      // S_r = pinv^T M_r pinv  (m x m), accumulated in long double
      std::vector<long double> tmp((size_t)nt * m, 0.0L);  // M_r pinv
      for (int p = 0; p < nt; ++p)
        for (int b = 0; b < m; ++b) {
          long double s = 0;
          for (int q = 0; q < nt; ++q) s += (long double)mr[(size_t)p * nt + q] * pinv[(size_t)q * m + b];
          tmp[(size_t)p * m + b] = s;
        }
      for (int aa = 0; aa < m; ++aa)
        for (int b = 0; b < m; ++b) {
          long double s = 0;
          for (int p = 0; p < nt; ++p) s += (long double)pinv[(size_t)p * m + aa] * tmp[(size_t)p * m + b];
          bc.s[kSOff[r] + aa * m + b] = (double)s;
        }
    }
    for (int k = 0; k <= r; ++k) {
      int col = xk_coefficient_column(k);
      double kf = factorial(k);
      for (int aa = 0; aa < m; ++aa) bc.w[kWOff[r] + k * m + aa] = kf * pinv[(size_t)col * m + aa];
    }
  }
  return bc;
}

}  // namespace

void discover_ghost_indices(const Problem& pb, std::vector<std::pair<int64_t, int64_t>>* out) {
  // src/ghost.rs:352-376 with default_stencil_offsets() (radius-4 cross)
  std::set<std::pair<int64_t, int64_t>> set;
  for (int64_t i = 0; i < pb.grid.nx; ++i)
    for (int64_t j = 0; j < pb.grid.ny; ++j) {
      if (!pb.is_in_domain(i, j)) continue;
      for (int d = -4; d <= 4; ++d) {
        if (!pb.is_in_domain(i + d, j)) set.insert({i + d, j});
        if (!pb.is_in_domain(i, j + d)) set.insert({i, j + d});
      }
    }
  out->assign(set.begin(), set.end());  // lexicographic (i, j): sort_unstable order
}

void Problem::build_ghosts() {  // src/ghost.rs:50-118, 378-478
  std::vector<std::pair<int64_t, int64_t>> idx;
  discover_ghost_indices(*this, &idx);
  double h = std::sqrt(grid.dx * grid.dy);
  std::vector<std::vector<double>> beta_forms(5);
  for (int r = 0; r < 5; ++r) beta_forms[r] = beta_quadratic_form(r, h);

  // Periodic generalisation: the reference wraps only j with a hard-coded
  // period of 0.5 (src/bc1.rs:547-607).  Here a Periodic side wraps both
  // indices by the outer polygon's bounding-box extent (a no-op for an index
  // already inside the grid), so x-, y- and doubly-periodic boxes work.
  double bx0 = outer.points[0].x, bx1 = bx0, by0 = outer.points[0].y, by1 = by0;
  for (const auto& p : outer.points) {
    bx0 = std::fmin(bx0, p.x);
    bx1 = std::fmax(bx1, p.x);
    by0 = std::fmin(by0, p.y);
    by1 = std::fmax(by1, p.y);
  }
  int64_t np_x = (int64_t)std::round((bx1 - bx0) / grid.dx);
  int64_t np_y = (int64_t)std::round((by1 - by0) / grid.dy);

  ghosts.clear();
  heavy = std::make_shared<std::vector<GhostBC>>();
  ghosts.reserve(idx.size());
  for (const auto& ij : idx) {
    GhostInfo g;
    std::memset(&g, 0, sizeof(g));
    g.i = (int32_t)ij.first;
    g.j = (int32_t)ij.second;
    Point p{x(ij.first), y(ij.second)};
    bool outer_fluid = outer.is_fluid(p);
    g.boundary = outer_fluid ? 1 : 0;
    const Polygon& poly = g.boundary == 0 ? outer : inner;
    const auto& bcs = bc_list(g.boundary);
    Projection raw = rh3t::project(poly, p);
    find_nearest_grid_point(*this, raw, &g.nearest_i, &g.nearest_j);
    size_t side = select_boundary_side(raw.point, poly, bcs);
    g.side_id = (int32_t)side;
    const CircularArc* arc = arc_for_side(g.boundary, side);
    if (arc) {
      g.project = arc->project(p);
    } else {
      Vec2 n = poly.outward_normal_of_side(side);
      double dx = p.x - raw.point.x, dy = p.y - raw.point.y;
      g.project = Projection{raw.point, n, dx * n.x + dy * n.y};
    }
    const rh3t_bc& bc = bcs[side];
    g.bc_kind = bc.kind;
    g.heavy_id = -1;
    g.src_i = -1;
    g.src_j = -1;
    switch (bc.kind) {
      case RH3T_BC_WALL:
      case RH3T_BC_OUTFLOW:
      case RH3T_BC_FARFIELD:
        g.heavy_id = (int32_t)heavy->size();
        heavy->push_back(precompute_ghost_bc(*this, g.project, beta_forms));
        break;
      case RH3T_BC_REFLECTIVE_WALL:
      case RH3T_BC_ZEROTH_ORDER:
        if (!mirror_source(*this, g.i, g.j, g.project, &g.src_i, &g.src_j))
          fail(RH3T_ERR_GEOMETRY, "cannot find reflected interior point for ghost (" +
                                      std::to_string(g.i) + "," + std::to_string(g.j) + ")");
        break;
      case RH3T_BC_PERIODIC: {
        if (np_x <= 0 || np_y <= 0) fail(RH3T_ERR_GEOMETRY, "degenerate periodic extent");
        int64_t wi = g.i, wj = g.j;
        if (wi < 0 || wi >= grid.nx) wi = rem_euclid_i(wi, np_x);
        if (wj < 0 || wj >= grid.ny) wj = rem_euclid_i(wj, np_y);
        if (!grid_in_domain(wi, wj) || !is_in_domain(wi, wj))
          fail(RH3T_ERR_GEOMETRY, "periodic wrapped point is not fluid: ghost (" +
                                      std::to_string(g.i) + "," + std::to_string(g.j) + ")");
        g.src_i = (int32_t)wi;
        g.src_j = (int32_t)wj;
        break;
      }
      default:
        break;
    }
    ghosts.push_back(g);
  }
  ghosts_built = true;
}

}  // namespace rh3t

// =================================================================== C ABI

using rh3t::BuildError;
using rh3t::Point;
using rh3t::Problem;

namespace {

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const BuildError& e) {
    return e.status;
  } catch (...) {
    return RH3T_ERR_ARG;
  }
}

// RH3T_OK, RH3T_ERR_ARG (unknown kind) or RH3T_ERR_UNSUPPORTED (a
// TimeDependent BC without a built-in profile of this build)
int check_bc(const rh3t_bc& b) {
  if (b.kind < 0 || b.kind > 7) return RH3T_ERR_ARG;
  if (b.kind == RH3T_BC_TIME_DEPENDENT && b.profile != RH3T_TD_MOVING_FRONT) return RH3T_ERR_UNSUPPORTED;
  return RH3T_OK;
}

}  // namespace

extern "C" {

void rh3t_physics_default(rh3t_physics* p) {  // src/constant.rs:1-25
  if (!p) return;
  p->gamma_e = 1.4;
  p->gamma_i = 1.4;
  p->gamma_r = 1.4;
  p->kappa_e = 0.0;
  p->kappa_i = 0.0;
  p->kappa_r = 0.0;
  p->cve = 1.0;
  p->cvi = 1.0;
  p->a_rad = 1.0;
  p->omega_ei = 0.0;
  p->omega_er = 0.0;
  p->lambda = 0.5;
  p->weno_q = 10.0;
  p->eps = 1e-12;
}

int rh3t_problem_create(const rh3t_grid* grid, const double* outer_xy, int32_t n_outer,
                        const rh3t_bc* bc_outer, const double* inner_xy, int32_t n_inner,
                        const rh3t_bc* bc_inner, rh3t_problem** out) {
  if (!grid || !out || !outer_xy || !bc_outer || n_outer < 3) return RH3T_ERR_ARG;
  if (grid->nx <= 0 || grid->ny <= 0 || !(grid->dx > 0.0) || !(grid->dy > 0.0)) return RH3T_ERR_ARG;
  if (n_inner > 0 && (!inner_xy || !bc_inner)) return RH3T_ERR_ARG;
  for (int k = 0; k < n_outer; ++k)
    if (int st = check_bc(bc_outer[k])) return st;
  for (int k = 0; k < n_inner; ++k)
    if (int st = check_bc(bc_inner[k])) return st;
  *out = nullptr;
  return guarded([&]() {
    auto* pb = new rh3t_problem();
    Problem& p = pb->p;
    p.grid = *grid;
    p.outer.fluid = rh3t::FluidSide::Inside;
    for (int k = 0; k < n_outer; ++k) p.outer.points.push_back(Point{outer_xy[2 * k], outer_xy[2 * k + 1]});
    p.bc_outer.assign(bc_outer, bc_outer + n_outer);
    p.inner.fluid = rh3t::FluidSide::Outside;
    for (int k = 0; k < n_inner; ++k) p.inner.points.push_back(Point{inner_xy[2 * k], inner_xy[2 * k + 1]});
    p.bc_inner.assign(bc_inner, bc_inner + (n_inner > 0 ? n_inner : 0));
    p.build_fluid_mask();
    *out = pb;
    return (int)RH3T_OK;
  });
}

void rh3t_problem_destroy(rh3t_problem* pb) { delete pb; }

int rh3t_problem_add_arc(rh3t_problem* pb, int32_t boundary, int32_t side_start, int32_t side_end,
                         int32_t fluid_outside, const double s[2], const double m[2], const double e[2]) {
  if (!pb || !s || !m || !e || (boundary != 0 && boundary != 1)) return RH3T_ERR_ARG;
  size_t nsides = boundary == 0 ? pb->p.outer.points.size() : pb->p.inner.points.size();
  if (side_start < 0 || side_end < side_start || (size_t)side_end >= nsides) return RH3T_ERR_ARG;
  return guarded([&]() {
    rh3t::ArcOverride ov;
    ov.side_start = (size_t)side_start;
    ov.side_end = (size_t)side_end;
    ov.arc = rh3t::CircularArc::from_three_points(
        Point{s[0], s[1]}, Point{m[0], m[1]}, Point{e[0], e[1]},
        fluid_outside ? rh3t::FluidSide::Outside : rh3t::FluidSide::Inside);
    (boundary == 0 ? pb->p.outer_arcs : pb->p.inner_arcs).push_back(ov);
    pb->p.ghosts_built = false;
    return (int)RH3T_OK;
  });
}

int rh3t_problem_build_ghosts(rh3t_problem* pb) {
  if (!pb) return RH3T_ERR_ARG;
  return guarded([&]() {
    pb->p.build_ghosts();
    return (int)RH3T_OK;
  });
}

int rh3t_problem_init_cylinder(int32_t nx, double mach, rh3t_problem** out) {
  // src/init.rs:116-510.  h = 1/40 and nx = 121 in the reference; the grid
  // family keeps the domain [-3,0]x[-6,6] with h = 3/(nx-1).
  if (!out || nx < 9 || !(mach > 0.0)) return RH3T_ERR_ARG;
  *out = nullptr;
  return guarded([&]() {
    const double pi = 3.14159265358979323846;
    // 3.0/120.0 and the reference's 1.0/40.0 round to the same double.
    double h = 3.0 / (double)(nx - 1);
    double x0 = -3.0, y0 = -6.0, lx = 3.0, ly = 12.0;
    rh3t_grid g;
    g.nx = (int32_t)std::round(lx / h) + 1;
    g.ny = (int32_t)std::round(ly / h) + 1;
    g.dx = h;
    g.dy = h;
    g.x0 = x0;
    g.y0 = y0;
    if (g.nx != nx) throw BuildError{RH3T_ERR_ARG, "nx incompatible with h = 3/(nx-1)"};

    double uinf[6];
    {
      double gamma = 1.4, rho = 1.0, p = 1.0;
      double a = std::sqrt(gamma * p / rho);
      double ux = mach * a, uy = 0.0;
      double kinetic = 0.5 * rho * (ux * ux + uy * uy);  // euler_to_three_energy, init.rs:7-24
      double e_total = p / (gamma - 1.0) + kinetic;
      uinf[0] = rho;
      uinf[1] = rho * ux;
      uinf[2] = rho * uy;
      uinf[3] = e_total / 3.0;
      uinf[4] = e_total / 3.0;
      uinf[5] = e_total / 3.0;
    }
    rh3t_bc far;
    std::memset(&far, 0, sizeof(far));
    far.kind = RH3T_BC_FARFIELD;
    std::memcpy(far.state, uinf, sizeof(uinf));
    rh3t_bc refl;
    std::memset(&refl, 0, sizeof(refl));
    refl.kind = RH3T_BC_REFLECTIVE_WALL;

    std::vector<double> pts;
    std::vector<rh3t_bc> bcs;
    pts.insert(pts.end(), {-3.0, -6.0});
    pts.insert(pts.end(), {0.0, -6.0});
    bcs.push_back(far);
    pts.insert(pts.end(), {0.0, -1.0 + 0.0125});
    bcs.push_back(far);
    const int n_arc = 360;
    int arc_side_start = (int)bcs.size();
    for (int k = 1; k <= n_arc; ++k) {
      double s = (double)k / (double)n_arc;
      double theta = -0.5 * pi - pi * s;
      pts.push_back(std::cos(theta));
      pts.push_back(std::sin(theta) + 0.0125);
      bcs.push_back(refl);
    }
    int arc_side_end = (int)bcs.size() - 1;
    pts.insert(pts.end(), {0.0, 6.0});
    bcs.push_back(far);
    pts.insert(pts.end(), {-3.0, 6.0});
    bcs.push_back(far);
    bcs.push_back(far);
    const double inner[8] = {-1002.0, -1002.0, -1001.0, -1002.0, -1001.0, -1001.0, -1002.0, -1001.0};
    rh3t_bc wall;
    std::memset(&wall, 0, sizeof(wall));
    wall.kind = RH3T_BC_WALL;
    rh3t_bc bc_inner[4] = {wall, wall, wall, wall};
    rh3t_problem* pb = nullptr;
    int st = rh3t_problem_create(&g, pts.data(), (int32_t)(pts.size() / 2), bcs.data(), inner, 4, bc_inner, &pb);
    if (st != RH3T_OK) return st;
    const double s0[2] = {0.0, -1.0 + 0.0125}, m0[2] = {-1.0, 0.0 + 0.0125}, e0[2] = {0.0, 1.0 + 0.0125};
    st = rh3t_problem_add_arc(pb, 0, arc_side_start, arc_side_end, 1, s0, m0, e0);
    if (st == RH3T_OK) st = rh3t_problem_build_ghosts(pb);
    if (st != RH3T_OK) {
      delete pb;
      return st;
    }
    *out = pb;
    return (int)RH3T_OK;
  });
}

namespace {

// euler_to_three_energy, src/init.rs:7-24
void euler_to_three_energy(double rho, double ux, double uy, double p, double gamma, double s[6]) {
  double kinetic = 0.5 * rho * (ux * ux + uy * uy);
  double e_total = p / (gamma - 1.0) + kinetic;
  s[0] = rho;
  s[1] = rho * ux;
  s[2] = rho * uy;
  s[3] = e_total / 3.0;
  s[4] = e_total / 3.0;
  s[5] = e_total / 3.0;
}

// The Mach-10 states of init_double_mach (src/init.rs:45-49).
void double_mach_states(double pre[6], double post[6]) {
  euler_to_three_energy(1.4, 0.0, 0.0, 1.0, 1.4, pre);
  euler_to_three_energy(8.0, 8.25, 0.0, 116.5, 1.4, post);
}

}  // namespace

int rh3t_problem_init_double_mach(double h, rh3t_problem** out) {
  return rh3t_problem_init_double_mach_top(h, nullptr, out);
}

int rh3t_problem_init_double_mach_top(double h, const rh3t_bc* top, rh3t_problem** out) {
  // src/init.rs:26-113 (h = 1/320 in the reference): the polygon
  // A(-1/2 - sqrt3/12, 0) B(0, 0) C(23 sqrt3/12, 23/12) D(C.x, 23/12 +
  // sqrt3/2) E(A.x, D.y), Constant(post) on AB and EA, the 30-degree Wall on
  // BC, Constant(pre) on CD, ZerothOrder (or the caller's top BC) on DE.
  if (!out || !(h > 0.0) || h > 0.25) return RH3T_ERR_ARG;
  *out = nullptr;
  return guarded([&]() {
    const double sqrt3 = std::sqrt(3.0);
    const double ax = -0.5 - sqrt3 / 12.0, ay = 0.0;
    const double bx = 0.0, by = 0.0;
    const double cx = 23.0 * sqrt3 / 12.0, cy = 23.0 / 12.0;
    const double dx = cx, dy = 23.0 / 12.0 + sqrt3 / 2.0;
    const double ex = ax, ey = dy;
    const double lx = cx - ax, ly = dy;
    rh3t_grid g;
    g.nx = (int32_t)std::ceil(lx / h) + 1;
    g.ny = (int32_t)std::ceil(ly / h) + 1;
    g.dx = h;
    g.dy = h;
    g.x0 = ax;
    g.y0 = 0.0;
    double pre[6], post[6];
    double_mach_states(pre, post);
    rh3t_bc c_post, c_pre, wall, zeroth;
    std::memset(&c_post, 0, sizeof(c_post));
    c_post.kind = RH3T_BC_CONSTANT;
    std::memcpy(c_post.state, post, sizeof(post));
    c_pre = c_post;
    std::memcpy(c_pre.state, pre, sizeof(pre));
    std::memset(&wall, 0, sizeof(wall));
    wall.kind = RH3T_BC_WALL;
    std::memset(&zeroth, 0, sizeof(zeroth));
    zeroth.kind = RH3T_BC_ZEROTH_ORDER;
    const double pts[10] = {ax, ay, bx, by, cx, cy, dx, dy, ex, ey};
    const rh3t_bc bcs[5] = {c_post, wall, c_pre, top ? *top : zeroth, c_post};
    const double inner[8] = {-1002.0, -1002.0, -1001.0, -1002.0, -1001.0, -1001.0, -1002.0, -1001.0};
    const rh3t_bc bc_inner[4] = {wall, wall, wall, wall};
    rh3t_problem* pb = nullptr;
    int st = rh3t_problem_create(&g, pts, 5, bcs, inner, 4, bc_inner, &pb);
    if (st != RH3T_OK) return st;
    st = rh3t_problem_build_ghosts(pb);
    if (st != RH3T_OK) {
      delete pb;
      return st;
    }
    *out = pb;
    return (int)RH3T_OK;
  });
}

int rh3t_problem_double_mach_state(const rh3t_problem* pb, double* u) {
  // src/init.rs:97-104: post-shock where x <= 0, pre-shock elsewhere, on
  // the fluid cells (zero elsewhere)
  if (!pb || !u) return RH3T_ERR_ARG;
  double pre[6], post[6];
  double_mach_states(pre, post);
  const rh3t_grid& g = pb->p.grid;
  for (int i = 0; i < g.nx; ++i)
    for (int j = 0; j < g.ny; ++j) {
      const size_t l = (size_t)i * g.ny + j;
      const double x = g.x0 + (double)i * g.dx;
      const double* s = x <= 0.0 ? post : pre;
      for (int c = 0; c < 6; ++c) u[l * 6 + c] = pb->p.fluid[l] ? s[c] : 0.0;
    }
  return RH3T_OK;
}

int rh3t_problem_clone(const rh3t_problem* src, rh3t_problem** out) {
  if (!src || !out) return RH3T_ERR_ARG;
  *out = nullptr;
  return guarded([&]() {
    auto* pb = new rh3t_problem();
    pb->p = src->p;  // ghost tables shared through the heavy shared_ptr
    *out = pb;
    return (int)RH3T_OK;
  });
}

int rh3t_problem_set_bc_state(rh3t_problem* pb, int32_t kind, const double state[6]) {
  if (!pb || !state || (kind != RH3T_BC_FARFIELD && kind != RH3T_BC_CONSTANT)) return RH3T_ERR_ARG;
  int n = 0;
  for (auto* l : {&pb->p.bc_outer, &pb->p.bc_inner})
    for (auto& b : *l)
      if (b.kind == kind) {
        std::memcpy(b.state, state, sizeof(double) * 6);
        ++n;
      }
  return n > 0 ? RH3T_OK : RH3T_ERR_ARG;
}

int rh3t_problem_cylinder_state(const rh3t_problem* pb, double mach, double* u) {
  if (!pb || !u) return RH3T_ERR_ARG;
  double gamma = 1.4, rho = 1.0, p = 1.0;
  double a = std::sqrt(gamma * p / rho);
  double ux = mach * a, uy = 0.0;
  double kinetic = 0.5 * rho * (ux * ux + uy * uy);
  double e_total = p / (gamma - 1.0) + kinetic;
  double s[6] = {rho, rho * ux, rho * uy, e_total / 3.0, e_total / 3.0, e_total / 3.0};
  size_t n = (size_t)pb->p.grid.nx * pb->p.grid.ny;
  for (size_t l = 0; l < n; ++l)
    for (int c = 0; c < 6; ++c) u[l * 6 + c] = pb->p.fluid[l] ? s[c] : 0.0;
  return RH3T_OK;
}

int rh3t_problem_grid(const rh3t_problem* pb, rh3t_grid* out) {
  if (!pb || !out) return RH3T_ERR_ARG;
  *out = pb->p.grid;
  return RH3T_OK;
}

int rh3t_problem_counts(const rh3t_problem* pb, int64_t* n_fluid, int64_t* n_ghost, int64_t* n_heavy) {
  if (!pb) return RH3T_ERR_ARG;
  int64_t nf = 0;
  for (uint8_t f : pb->p.fluid) nf += f;
  if (n_fluid) *n_fluid = nf;
  if (n_ghost) *n_ghost = (int64_t)pb->p.ghosts.size();
  if (n_heavy) *n_heavy = (int64_t)pb->p.heavy->size();
  return pb->p.ghosts_built ? RH3T_OK : RH3T_ERR_ARG;
}

int rh3t_problem_fluid_mask(const rh3t_problem* pb, uint8_t* mask) {
  if (!pb || !mask) return RH3T_ERR_ARG;
  std::memcpy(mask, pb->p.fluid.data(), pb->p.fluid.size());
  return RH3T_OK;
}

int rh3t_problem_ghosts(const rh3t_problem* pb, rh3t_ghost* out) {
  if (!pb || !out || !pb->p.ghosts_built) return RH3T_ERR_ARG;
  for (size_t k = 0; k < pb->p.ghosts.size(); ++k) {
    const auto& g = pb->p.ghosts[k];
    rh3t_ghost& o = out[k];
    o.i = g.i;
    o.j = g.j;
    o.boundary = g.boundary;
    o.side_id = g.side_id;
    o.bc_kind = g.bc_kind;
    o.heavy_id = g.heavy_id;
    o.src_i = g.src_i;
    o.src_j = g.src_j;
    o.nearest_i = g.nearest_i;
    o.nearest_j = g.nearest_j;
    o.p0x = g.project.point.x;
    o.p0y = g.project.point.y;
    o.nx = g.project.normal.x;
    o.ny = g.project.normal.y;
    o.distance = g.project.distance;
  }
  return RH3T_OK;
}

int rh3t_problem_ghost_bc(const rh3t_problem* pb, int32_t heavy_id, rh3t_ghost_bc* out) {
  if (!pb || !out || heavy_id < 0 || (size_t)heavy_id >= pb->p.heavy->size()) return RH3T_ERR_ARG;
  const auto& h = (*pb->p.heavy)[heavy_id];
  out->nearest_i = h.nearest_i;
  out->nearest_j = h.nearest_j;
  out->normal[0] = h.normal.x;
  out->normal[1] = h.normal.y;
  std::memcpy(out->d, h.d, sizeof(h.d));
  std::memcpy(out->stencil, h.stencil, sizeof(h.stencil));
  std::memcpy(out->s, h.s, sizeof(h.s));
  std::memcpy(out->w, h.w, sizeof(h.w));
  return RH3T_OK;
}

int rh3t_problem_bc(const rh3t_problem* pb, int32_t boundary, int32_t side, rh3t_bc* out) {
  if (!pb || !out || (boundary != 0 && boundary != 1)) return RH3T_ERR_ARG;
  const auto& l = pb->p.bc_list(boundary);
  if (side < 0 || (size_t)side >= l.size()) return RH3T_ERR_ARG;
  *out = l[side];
  return RH3T_OK;
}

int rh3t_problem_access_table(const rh3t_problem* pb, uint32_t* access) {
  // GhostGrid::build access table, src/ghost.rs:68-107, with the sorted
  // 17-offset list of default_stencil_offsets().
  if (!pb || !access || !pb->p.ghosts_built) return RH3T_ERR_ARG;
  const Problem& p = pb->p;
  int64_t nx = p.grid.nx, ny = p.grid.ny;
  std::vector<std::pair<int, int>> offs;
  for (int d = -4; d <= 4; ++d) {
    offs.push_back({d, 0});
    offs.push_back({0, d});
  }
  std::sort(offs.begin(), offs.end());
  offs.erase(std::unique(offs.begin(), offs.end()), offs.end());
  // ghost id lookup through the sorted ghost list
  auto gid = [&](int64_t i, int64_t j) -> int64_t {
    auto it = std::lower_bound(p.ghosts.begin(), p.ghosts.end(), std::make_pair(i, j),
                               [](const rh3t::GhostInfo& g, const std::pair<int64_t, int64_t>& k) {
                                 return std::make_pair((int64_t)g.i, (int64_t)g.j) < k;
                               });
    if (it == p.ghosts.end() || it->i != i || it->j != j) return -1;
    return it - p.ghosts.begin();
  };
  std::memset(access, 0, sizeof(uint32_t) * (size_t)nx * ny * offs.size());
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j) {
      size_t l = (size_t)i * ny + j;
      if (!p.fluid[l]) continue;
      for (size_t k = 0; k < offs.size(); ++k) {
        int64_t ti = i + offs[k].first, tj = j + offs[k].second;
        int64_t t;
        if (p.grid_in_domain(ti, tj) && p.fluid[(size_t)ti * ny + tj]) {
          t = ti * ny + tj;
        } else {
          int64_t id = gid(ti, tj);
          if (id < 0) return RH3T_ERR_GEOMETRY;  // "ghost ... is not cached"
          t = id + nx * ny;
        }
        access[l * offs.size() + k] = (uint32_t)t;
      }
    }
  return RH3T_OK;
}

}  // extern "C"
