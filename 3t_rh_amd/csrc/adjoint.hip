// Reverse mode (VJP) of l() and of the SSP-RK3 step of colinhu123/3T_RH on
// gfx950, with a checkpoint-and-recompute schedule over time steps.
//
// For a stage state S with cotangent LAM of the stage output and a scale c
// (c = dt, dt/4, 2dt/3 for the three stages), one VJP evaluation forms
//   OUT = a LAM + b1 X1 + b2 X2 + c (dl/dS)^T LAM
// in gather form (every output written by exactly one lane, every sum in a
// fixed order: results are bitwise reproducible):
//   k_adj_iface      per interface (x and y; a lane pair each): the
//                    pullback of the characteristic WENO5 flux
//                    (src/weno.rs:685-802) through the eigensystem, the
//                    physical fluxes and the Derived values of its 6
//                    stencil points (36 state cotangents per interface,
//                    stored)
//   k_adj_gather     per fluid cell / ghost: sums its 12 interface
//                    contributions, the pullbacks of the noncon terms of the
//                    9-point lines through it (src/noncon.rs:288-332), of the
//                    diffusion fluxes (src/diffusion.rs:27-50) and of the
//                    exchange source (src/source.rs:7-21), then
//                    Derived::from_state's pullback (src/state.rs:241-275);
//                    per-block partial sums of the parameter cotangents
//   k_adj_ghost_T    transposes the cheap ghost maps (Periodic / Zeroth
//                    copies, ReflectiveWall mirror, src/bc1.rs:471-607,
//                    1015-1036) into their source cells
//   k_adj_param      sums the parameter partials of a call, in block order
// The parameters are (omega_ei, omega_er, kappa_e, kappa_i, kappa_r)
// (src/constant.rs:17-25); their cotangents are formed whether or not the
// forward skips the zero-valued terms.  The time step is not differentiated
// (dt is frozen).
//   k_adj_heavy      per FarField / Outflow / Wall ghost: the pullback of
//                    its value (src/bc1.rs:193-255, 691-928) to the states
//                    of its E_r points and its nearest point
//   k_adj_heavy_T    adds those into the cells (fixed order)
// TimeDependent ghosts are not covered.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "adjoint.hpp"
#include "engine.hpp"

using namespace rh3t_dev;
using namespace rh3t_eng;

namespace {

constexpr int kNParam = 5;

struct AdjDesc {
  Geo g;
  const uint8_t* pt;
  const double* S;     // stage state with its ghosts
  const double* D;     // its Derived planes (D_N)
  const double* LAM;   // cotangent of the stage output (fluid cells, 0 elsewhere)
  double* XB;          // x interfaces: [q][interface][f], q the stencil point
  double* YB;          // y interfaces: [q][interface][f]
  double* SB;          // 6 planes: state cotangents of ghost points
  double* OUT;         // 6 planes: result on fluid cells
  const double* X1;
  const double* X2;
  double a, b1, b2, rscale;
  double* PG;          // kNParam partial sums per gather block
};

// sqrt(rho)-weighted Roe average of stencil points 3 (slot c0) and 2
struct RoeAvg {
  double r1, r2, irs, q3[5], q2[5], qr[5];
};

__device__ __forceinline__ void roe_average(const double* S, const double* D, long P, long c0, long p2,
                                            RoeAvg& A) {
  A.r1 = sqrt(S[c0]);
  A.r2 = sqrt(S[p2]);
  A.irs = 1.0 / (A.r1 + A.r2);
  const int planes[5] = {D_U, D_V, D_EE, D_EI, D_ER};
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    A.q3[k] = D[planes[k] * P + c0];
    A.q2[k] = D[planes[k] * P + p2];
    A.qr[k] = (A.q3[k] * A.r1 + A.q2[k] * A.r2) * A.irs;
  }
}

// Interface flux pullback in two phases, two lanes per interface (32
// interfaces per 64-lane workgroup).  Characteristic phase: the WENO5 /
// local Lax-Friedrichs / eigenvector pullback of every characteristic
// (src/weno.rs:685-802); lane `half` of a pair holds stencil points
// 3 half .. 3 half + 2 in registers and evaluates one of the two WENO5
// pullbacks of each characteristic (characteristic_step_pair).  The
// split-flux cotangents of its points (6 per characteristic) go to LDS, the
// eigensystem cotangents to the Roe state, formed by both lanes.  Point
// phase: each lane pulls its three points' split-flux cotangents back
// through L (F and U projections), the physical fluxes (flux_from_derived,
// src/state.rs:56-85), the Roe average and Derived::from_state to their
// states (36 values per interface, stored for the gather).  Half the
// stencil per lane fits two wavefronts per SIMD (256 VGPRs, a few spilled);
// one lane per interface needs ~450 registers and runs one wavefront per
// SIMD, 4% slower end to end on the reverse mode.
constexpr int kIfBx = 64;
constexpr int kIfPair = 32;
constexpr int kCbPair = 36;

// Interfaces of a direction are numbered i-major, t = i * nj + j, and a
// workgroup takes 32 consecutive ones (a workgroup may span two lines).
template <int DIR>
__device__ __forceinline__ void adj_iface(const AdjDesc& A, const EigConst* __restrict__ Kp, const Phys& ph,
                                          int tile, double (*s_cb)[kIfBx]) {
  // not held to the forward's rounding sequence (unlike k_adj_heavy, which
  // recomputes the forward's values): products and sums contract to FMA
#pragma clang fp contract(fast)
  const Geo g = A.g;
  const int ni = DIR == 0 ? g.nx + 1 : g.nx;
  const int nj = DIR == 0 ? g.ny : g.ny + 1;
  const int half = threadIdx.x & 1;
  const long t = (long)tile * kIfPair + (threadIdx.x >> 1);
  if (t >= (long)ni * nj) return;  // both lanes of a pair
  const int i = (int)(t / nj), j = (int)(t - (long)i * nj);
  const long fpl = DIR == 0 ? (long)(g.nx + 1) * g.ny : (long)g.nx * (g.ny + 1);
  const long fidx = DIR == 0 ? (long)i * g.ny + j : (long)i * (g.ny + 1) + j;
  double* __restrict__ out = DIR == 0 ? A.XB : A.YB;
  const long c0 = pad_index(g, i, j);
  const long step = DIR == 0 ? g.PY : 1;
  const int q0 = 3 * half;
  if (A.pt[c0] != 1 && A.pt[c0 - step] != 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int f = 0; f < 6; ++f) out[((q0 + k) * fpl + fidx) * 6 + f] = 0.0;
    return;
  }
  const double* S = A.S;
  const double* D = A.D;
  const long P = g.plane;
  double* cbs = &s_cb[0][threadIdx.x];  // this lane's column, stride kIfBx
  RoeV R;
  {
    RoeAvg ra;
    roe_average(S, D, P, c0, c0 - step, ra);
    roe_values<DIR>(*fresh(Kp), ra.qr[0], ra.qr[1], ra.qr + 2, R);
  }
  double roeb[5];
  {
    // flux F[i] enters r(i) with +1/ds and r(i-1) with -1/ds (src/main.rs:266-270)
    const double inv = 1.0 / (DIR == 0 ? g.dx : g.dy);
    double ob[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) ob[f] = A.rscale * inv * (A.LAM[f * P + c0] - A.LAM[f * P + c0 - step]);
    double U[3][6], F[3][6];
#pragma unroll
    for (int k = 0; k < 3; ++k) load_point<DIR>(S, D, P, (unsigned)((c0 + (q0 + k - 3) * step) * 8), U[k], F[k]);
    RoeBar B;
    roe_bar_zero(B);
    // not unrolled: one characteristic's working set at a time (16 VGPRs
    // spilled instead of 38; 214 -> 193 us per call at 512^2)
#pragma nounroll
    for (int c = 0; c < 6; ++c) {
      double cb6[6];
      characteristic_step_pair<DIR>(c, half, *fresh(Kp), ph.eps, R, U, F, ob, cb6, B);
#pragma unroll
      for (int k = 0; k < 6; ++k) cbs[(c * 6 + k) * kIfBx] = cb6[k];
    }
    roe_bar_final<DIR>(*fresh(Kp), R, B, roeb);
  }
  // the split-flux cotangents are read back from LDS, not kept in registers
  reload_barrier();
  // ---- point phase, this lane's points q0 .. q0 + 2
  double L[6][6], a[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double col[6], lam;
    row_col<DIR>(*fresh(Kp), R, c, L[c], col, &lam);
    a[c] = fabs(lam);
  }
  // the Roe-average pullback enters point 2 (half 0) or point 3 (half 1)
  double bb[5], rb = 0.0, rr;
  {
    RoeAvg ra;
    roe_average(S, D, P, c0, c0 - step, ra);
    rr = half ? ra.r1 : ra.r2;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      bb[k] = roeb[k] * rr * ra.irs;
      rb += roeb[k] * ((half ? ra.q3[k] : ra.q2[k]) - ra.qr[k]) * ra.irs;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int q = q0 + k;
    const long p = c0 + (q - 3) * step;
    // cotangents of this point's physical flux (fb) and state (sb) through L
    double fb[6] = {0, 0, 0, 0, 0, 0}, sb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double fpb = cbs[(c * 6 + k) * kIfBx], fmb = cbs[(c * 6 + 3 + k) * kIfBx];
      const double cfb = 0.5 * (fpb + fmb), cub = a[c] * 0.5 * (fpb - fmb);
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        fb[f] += cfb * L[c][f];
        sb[f] += cub * L[c][f];
      }
    }
    double st[6], db[D_N];
#pragma unroll
    for (int f = 0; f < 6; ++f) st[f] = S[f * P + p];
#pragma unroll
    for (int m = 0; m < D_N; ++m) db[m] = 0.0;
    const double u = D[D_U * P + p], v = D[D_V * P + p];
    const double pk[3] = {D[D_PE * P + p], D[D_PI * P + p], D[D_PR * P + p]};
    const double rho = st[0], w = DIR == 0 ? u : v;
    double ub = 0.0, vb = 0.0, wb = fb[0] * rho;
    sb[0] += fb[0] * w;
    if (DIR == 0) {
      sb[0] += fb[1] * u * u + fb[2] * u * v;
      ub += fb[1] * 2.0 * rho * u + fb[2] * rho * v;
      vb += fb[2] * rho * u;
    } else {
      sb[0] += fb[1] * u * v + fb[2] * v * v;
      ub += fb[1] * rho * v;
      vb += fb[1] * rho * u + fb[2] * 2.0 * rho * v;
    }
    const double Pb = DIR == 0 ? fb[1] : fb[2];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      sb[3 + m] += fb[3 + m] * w;
      wb += fb[3 + m] * (st[3 + m] + pk[m]);
      db[D_PE + m] = Pb + fb[3 + m] * w;
    }
    if (DIR == 0)
      ub += wb;
    else
      vb += wb;
    db[D_U] = ub;
    db[D_V] = vb;
    if (half ? k == 0 : k == 2) {
      db[D_U] += bb[0];
      db[D_V] += bb[1];
      db[D_EE] = bb[2];
      db[D_EI] = bb[3];
      db[D_ER] = bb[4];
      sb[0] += rb * 0.5 / rr;
    }
    derived_vjp(ph, st, db, sb);
#pragma unroll
    for (int f = 0; f < 6; ++f) out[(q * fpl + fidx) * 6 + f] = sb[f];
  }
}

// x interfaces in workgroups [0, nbx), y interfaces after them: one launch,
// so that the two directions share a single partial last round of
// workgroups
__global__ __launch_bounds__(kIfBx) __attribute__((amdgpu_waves_per_eu(2))) void k_adj_iface(
    AdjDesc A, const EigConst* __restrict__ Kp, Phys ph, int nbx) {
  __shared__ double s_cb[kCbPair][kIfBx];
  if ((int)blockIdx.x < nbx)
    adj_iface<0>(A, Kp, ph, blockIdx.x, s_cb);
  else
    adj_iface<1>(A, Kp, ph, blockIdx.x - nbx, s_cb);
}

// Weights of N_t in jump(N[0..8)) and jump(N[1..9)) (n_jump_pre,
// src/noncon.rs:259-285, with the literal q_pulse weights of
// src/noncon.rs:49-57), of sixth_order_derivative's numerator
// (src/noncon.rs:26-44), and (2, -27, 270, -490, 270, -27, 2): the
// difference of the two diffusion fluxes of src/diffusion.rs:27-50 around a
// cell.  Used with compile-time indices only.
constexpr double kQc[7] = {-0.00714285714285723, 0.0595238095238118, -0.240476190476197, 0.759523809523819,
                           0.509523809523803,    -0.0904761904761891, 0.00952380952380977};
constexpr double jm_w(int t) { return (t >= 1 && t <= 7 ? kQc[7 - t] : 0.0) - (t <= 6 ? kQc[t] : 0.0); }
constexpr double jp_w(int t) { return t >= 1 ? jm_w(t - 1) : 0.0; }
constexpr double kS6[9] = {0.0, -1.0, 9.0, -45.0, 0.0, 45.0, -9.0, 1.0, 0.0};
constexpr double kLap[7] = {2.0, -27.0, 270.0, -490.0, 270.0, -27.0, 2.0};

// The gather stages, per (kGy x kGx) tile of points, the pressures, the two
// velocities, the scaled energy cotangents rb = rscale * LAM[3..6) and the
// point types of the tile plus a kGh-point halo in LDS; the noncon and
// diffusion line pullbacks read their 11-point lines from there.
constexpr int kGx = 32, kGy = 8, kGh = 5;
constexpr int kGr = kGy + 2 * kGh, kGc = kGx + 2 * kGh;
enum { T_PE = 0, T_PI, T_PR, T_U, T_V, T_R3, T_R4, T_R5, T_N };
struct GatherTile {
  double d[T_N][kGr][kGc];
  uint8_t pt[kGr][kGc];
};

// The tile planes along a line through a point, at offset o: from the LDS
// tile (TileLine) or straight from the padded planes (PlaneLine, same values).
template <int DIR>
struct TileLine {
  const GatherTile& T;
  int ty, tx;
  __device__ double d(int plane, int o) const {
    return T.d[plane][ty + kGh + (DIR == 0 ? o : 0)][tx + kGh + (DIR == 0 ? 0 : o)];
  }
  __device__ int pt(int o) const { return T.pt[ty + kGh + (DIR == 0 ? o : 0)][tx + kGh + (DIR == 0 ? 0 : o)]; }
};

template <int DIR>
struct PlaneLine {
  const AdjDesc& A;
  int ip, jp;
  __device__ bool at(int o, long& q) const {
    const int a = ip + (DIR == 0 ? o : 0), b = jp + (DIR == 0 ? 0 : o);
    q = (long)a * A.g.PY + b;
    return a >= 0 && a < A.g.nx + 8 && b >= 0 && b < A.g.PY;
  }
  __device__ double d(int plane, int o) const {
    long q;
    if (!at(o, q)) return 0.0;
    const long P = A.g.plane;
    switch (plane) {
      case T_PE: return A.D[D_PE * P + q];
      case T_PI: return A.D[D_PI * P + q];
      case T_PR: return A.D[D_PR * P + q];
      case T_U: return A.D[D_U * P + q];
      case T_V: return A.D[D_V * P + q];
      case T_R3: return A.rscale * A.LAM[3 * P + q];
      case T_R4: return A.rscale * A.LAM[4 * P + q];
      default: return A.rscale * A.LAM[5 * P + q];
    }
  }
  __device__ int pt(int o) const {
    long q;
    return at(o, q) ? A.pt[q] : 0;
  }
};

// One line (DIR 0: x, DIR 1: y) of cell-term pullbacks into the point with
// grid index (i, j): the noncon terms of the 9 cells whose stencils hold it
// (pressure and velocity cotangents) and the diffusion fluxes of the 7 cells
// around it (temperature cotangents, T_r^4 for the radiation field).
template <int DIR, class Line>
__device__ __forceinline__ void line_gather(const Line& T, const Geo& g, const Phys& ph, int i, int j,
                                            double db[D_N], double tb[3]) {
#pragma clang fp contract(fast)
  const double ds = DIR == 0 ? g.dx : g.dy;
  const double ids = 1.0 / ds;
  const int VP = DIR == 0 ? D_U : D_V;
  const int along = DIR == 0 ? i : j, across = DIR == 0 ? j : i;
  const int n_along = DIR == 0 ? g.nx : g.ny, n_across = DIR == 0 ? g.ny : g.nx;
  if (across < 0 || across >= n_across) return;
  // p_e, p_i, p_r and the velocity at offsets -5..5 (0 outside the padded
  // planes; such points are never in the stencil of a fluid cell)
  double Pv[3][11], V[11];
#pragma unroll
  for (int o = -5; o <= 5; ++o) {
#pragma unroll
    for (int m = 0; m < 3; ++m) Pv[m][o + 5] = T.d(T_PE + m, o);
    V[o + 5] = T.d(DIR == 0 ? T_U : T_V, o);
  }
  // rho_m(c) = sum_e rb_e(c) M[e][m], M = [[2,-1,-1],[-1,2,-1],[-1,-1,2]]
  double Pb[3] = {0.0, 0.0, 0.0};
  double rc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // rho_m of cells -1, 0, +1
  bool fc[3] = {false, false, false};
  const double kap[3] = {ph.kappa_e, ph.kappa_i, ph.kappa_r};
  const double dsc = 1.0 / (180.0 * ds * ds);
#pragma unroll
  for (int o = -4; o <= 4; ++o) {
    const int a = along + o;
    const bool fl = a >= 0 && a < n_along && T.pt(o) == 1;
    if (!fl) continue;
    const double r0 = T.d(T_R3, o), r1 = T.d(T_R4, o), r2 = T.d(T_R5, o);
    const double sr = r0 + r1 + r2;
    const double rh[3] = {3.0 * r0 - sr, 3.0 * r1 - sr, 3.0 * r2 - sr};
    const int t = 4 - o;  // position of the point in the stencil of cell o
    const double vc = V[o + 5], vm = V[o + 4], vp = V[o + 6];
    const double cl = fmax(fmax(vc, vm), 0.0) * (1.0 / 3.0);
    const double cr = fmin(fmin(vc, vp), 0.0) * (1.0 / 3.0);
    const double coef = (vc * (1.0 / 3.0) * kS6[t] * (1.0 / 60.0) + cl * jm_w(t) + cr * jp_w(t)) * ids;
#pragma unroll
    for (int m = 0; m < 3; ++m) Pb[m] += rh[m] * coef;
    if (o >= -1 && o <= 1) {
      fc[o + 1] = true;
#pragma unroll
      for (int m = 0; m < 3; ++m) rc[o + 1][m] = rh[m];
    }
    if (ph.diffusion && o >= -3 && o <= 3) {
      const double w = kLap[3 - o] * dsc;
      tb[0] += kap[0] * w * r0;
      tb[1] += kap[1] * w * r1;
      tb[2] += kap[2] * w * r2;
    }
  }
  db[D_PE] += Pb[0];
  db[D_PI] += Pb[1];
  db[D_PR] += Pb[2];
  // velocity: the v/3 factor of the point's own cell and the max/min
  // selections of cl (cells 0, +1) and cr (cells -1, 0).  sh: offset of a
  // cell's stencil origin in the line.
  auto jump_sum = [&](const double r[3], int sh, bool plus) {
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      double a = 0.0;
#pragma unroll
      for (int t = 0; t < 9; ++t) a += (plus ? jp_w(t) : jm_w(t)) * Pv[m][t + sh];
      acc += r[m] * a;
    }
    return acc * ids;
  };
  const double vs = V[5];
  double velb = 0.0;
  if (fc[1]) {
    double der = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      double d = 0.0;
#pragma unroll
      for (int t = 0; t < 9; ++t) d += kS6[t] * Pv[m][t + 1];
      der += rc[1][m] * d;
    }
    velb += der * (1.0 / 180.0) * ids;
    if (fmax(vs, V[4]) > 0.0 && vs >= V[4]) velb += jump_sum(rc[1], 1, false) * (1.0 / 3.0);
    if (fmin(vs, V[6]) < 0.0 && vs <= V[6]) velb += jump_sum(rc[1], 1, true) * (1.0 / 3.0);
  }
  if (fc[2] && fmax(V[6], vs) > 0.0 && !(V[6] >= vs)) velb += jump_sum(rc[2], 2, false) * (1.0 / 3.0);
  if (fc[0] && fmin(V[4], vs) < 0.0 && !(V[4] <= vs)) velb += jump_sum(rc[0], 0, true) * (1.0 / 3.0);
  db[VP] += velb;
}

// sum_t lap_t T(c + (t-3) st) / (180 ds^2) for the three temperatures
template <int DIR>
__device__ __forceinline__ void laplacian(const AdjDesc& A, long c, double lap[3]) {
#pragma clang fp contract(fast)
  const Geo& g = A.g;
  const long P = g.plane;
  const long st = DIR == 0 ? g.PY : 1;
  const double ds = DIR == 0 ? g.dx : g.dy;
  const double sc = 1.0 / (180.0 * ds * ds);
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < 7; ++t) {
      double T = A.D[(D_TE + e) * P + c + (t - 3) * st];
      if (e == 2) T = pow4(T);
      acc += kLap[t] * T;
    }
    lap[e] += acc * sc;
  }
}

// The gather of one point (ip, jp) of kind 1 (fluid: OUT) or 2 (ghost: its
// state cotangent SB); lx / ly: its x and y lines.  pg: the point's
// parameter cotangents.
template <class LX, class LY>
__device__ __forceinline__ void gather_point(const AdjDesc& A, const Phys& ph, int ip, int jp, int kind,
                                             const LX& lx, const LY& ly, double pg[kNParam]) {
  // a pullback with no forward values to reproduce: FMA contraction
#pragma clang fp contract(fast)
  const Geo& g = A.g;
  const long P = g.plane;
  const long p = (long)ip * g.PY + jp;
  const int i = ip - 4, j = jp - 4;
  double sb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // the 6 x and 6 y interfaces whose stencils hold p
  const long fxpl = (long)(g.nx + 1) * g.ny, fypl = (long)g.nx * (g.ny + 1);
  if (j >= 0 && j < g.ny) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int ix = i + 3 - q;
      if (ix < 0 || ix > g.nx) continue;
      const long fi = (long)ix * g.ny + j;
#pragma unroll
      for (int f = 0; f < 6; ++f) sb[f] += A.XB[(q * fxpl + fi) * 6 + f];
    }
  }
  if (i >= 0 && i < g.nx) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int jy = j + 3 - q;
      if (jy < 0 || jy > g.ny) continue;
      const long fi = (long)i * (g.ny + 1) + jy;
#pragma unroll
      for (int f = 0; f < 6; ++f) sb[f] += A.YB[(q * fypl + fi) * 6 + f];
    }
  }
  double db[D_N];
#pragma unroll
  for (int k = 0; k < D_N; ++k) db[k] = 0.0;
  double tb[3] = {0.0, 0.0, 0.0};
  line_gather<0>(lx, g, ph, i, j, db, tb);
  line_gather<1>(ly, g, ph, i, j, db, tb);
  if (ph.diffusion) {
    db[D_TE] += tb[0];
    db[D_TI] += tb[1];
    db[D_TR] += tb[2] * 4.0 * cube(A.D[D_TR * P + p]);
  }
  if (kind == 1) {
    const double rb3 = lx.d(T_R3, 0), rb4 = lx.d(T_R4, 0), rb5 = lx.d(T_R5, 0);
    const double te = A.D[D_TE * P + p], ti = A.D[D_TI * P + p], tr = A.D[D_TR * P + p];
    if (ph.source) {  // src/source.rs:7-21
      const double te3 = 4.0 * cube(te), tr3 = 4.0 * cube(tr);
      db[D_TE] += rb3 * (-ph.omega_ei - ph.omega_er * te3) + rb4 * ph.omega_ei + rb5 * ph.omega_er * te3;
      db[D_TI] += (rb3 - rb4) * ph.omega_ei;
      db[D_TR] += (rb3 - rb5) * ph.omega_er * tr3;
    }
    pg[0] = (te - ti) * (rb4 - rb3);
    pg[1] = (pow4(te) - pow4(tr)) * (rb5 - rb3);
    double lap[3] = {0.0, 0.0, 0.0};
    laplacian<0>(A, p, lap);
    laplacian<1>(A, p, lap);
    pg[2] = rb3 * lap[0];
    pg[3] = rb4 * lap[1];
    pg[4] = rb5 * lap[2];
  }
  double s[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) s[f] = A.S[f * P + p];
  derived_vjp(ph, s, db, sb);
  if (kind == 1) {
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      double v = A.a * A.LAM[f * P + p] + sb[f];
      if (A.X1) v += A.b1 * A.X1[f * P + p];
      if (A.X2) v += A.b2 * A.X2[f * P + p];
      A.OUT[f * P + p] = v;
    }
  } else {
#pragma unroll
    for (int f = 0; f < 6; ++f) A.SB[f * P + p] = sb[f];
  }
}

// Tiles of kGy x kGx points over the fluid box i in [0, nx), j in [0, ny)
// (its fluid cells and any ghosts inside it); the ghosts around the box go
// to k_adj_gather_ghosts.
__global__ __launch_bounds__(256) void k_adj_gather(AdjDesc A, Phys ph) {
  __shared__ GatherTile T;
  const Geo g = A.g;
  const long P = g.plane;
  const int tiles_c = (g.ny + kGx - 1) / kGx;
  const int ip0 = 4 + (int)(blockIdx.x / tiles_c) * kGy, jp0 = 4 + (int)(blockIdx.x % tiles_c) * kGx;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * kGx + tx;
  for (int idx = tid; idx < kGr * kGc; idx += kGx * kGy) {
    const int r = idx / kGc, c = idx % kGc;
    const int ip = ip0 - kGh + r, jp = jp0 - kGh + c;
    const bool in = ip >= 0 && ip < g.nx + 8 && jp >= 0 && jp < g.PY;
    const long q = (long)ip * g.PY + jp;
    T.pt[r][c] = in ? A.pt[q] : 0;
    T.d[T_PE][r][c] = in ? A.D[D_PE * P + q] : 0.0;
    T.d[T_PI][r][c] = in ? A.D[D_PI * P + q] : 0.0;
    T.d[T_PR][r][c] = in ? A.D[D_PR * P + q] : 0.0;
    T.d[T_U][r][c] = in ? A.D[D_U * P + q] : 0.0;
    T.d[T_V][r][c] = in ? A.D[D_V * P + q] : 0.0;
    T.d[T_R3][r][c] = in ? A.rscale * A.LAM[3 * P + q] : 0.0;
    T.d[T_R4][r][c] = in ? A.rscale * A.LAM[4 * P + q] : 0.0;
    T.d[T_R5][r][c] = in ? A.rscale * A.LAM[5 * P + q] : 0.0;
  }
  __syncthreads();
  const int ip = ip0 + ty, jp = jp0 + tx;
  const int kind = (ip < g.nx + 4 && jp < g.ny + 4) ? T.pt[ty + kGh][tx + kGh] : 0;
  double pg[kNParam] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (kind != 0)
    gather_point(A, ph, ip, jp, kind, TileLine<0>{T, ty, tx}, TileLine<1>{T, ty, tx}, pg);
  // block partial sums of the parameter cotangents (the tile's LDS is reused)
  __syncthreads();
  double(*red)[kGx * kGy] = reinterpret_cast<double(*)[kGx * kGy]>(&T.d[0][0][0]);
#pragma unroll
  for (int k = 0; k < kNParam; ++k) red[k][tid] = pg[k];
  __syncthreads();
  for (int w = kGx * kGy / 2; w > 0; w >>= 1) {
    if (tid < w)
#pragma unroll
      for (int k = 0; k < kNParam; ++k) red[k][tid] += red[k][tid + w];
    __syncthreads();
  }
  // each block owns its slots: partial sums accumulate over the VJPs of a
  // call in launch order
  if (tid < kNParam) A.PG[blockIdx.x * kNParam + tid] += red[tid][0];
}

// The ghosts outside the fluid box, one lane each (pads listed by the host).
__global__ __launch_bounds__(64) void k_adj_gather_ghosts(AdjDesc A, Phys ph, int n, const long* __restrict__ pads) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long p = pads[k];
  const int ip = (int)(p / A.g.PY), jp = (int)(p % A.g.PY);
  double pg[kNParam] = {0.0, 0.0, 0.0, 0.0, 0.0};
  gather_point(A, ph, ip, jp, 2, PlaneLine<0>{A, ip, jp}, PlaneLine<1>{A, ip, jp}, pg);
}

// One lane per source cell: adds the cotangents of the cheap ghosts copied
// from it (ghost-id order).
__global__ __launch_bounds__(256) void k_adj_ghost_T(int nrows, const long* __restrict__ row_cell,
                                                     const int* __restrict__ row_ptr, const long* __restrict__ ent_pad,
                                                     const int* __restrict__ ent_kind,
                                                     const double* __restrict__ ent_n, const double* __restrict__ SB,
                                                     double* __restrict__ OUT, long P) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const long c = row_cell[r];
  double acc[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) acc[f] = OUT[f * P + c];
  for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
    const long gp = ent_pad[e];
    double sb[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) sb[f] = SB[f * P + gp];
    if (ent_kind[e] == RH3T_BC_REFLECTIVE_WALL) {  // (I - 2 n n^T) is symmetric
      const double nx = ent_n[2 * e], ny = ent_n[2 * e + 1];
      const double mb = sb[1] * nx + sb[2] * ny;
      sb[1] -= 2.0 * mb * nx;
      sb[2] -= 2.0 * mb * ny;
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) acc[f] += sb[f];
  }
#pragma unroll
  for (int f = 0; f < 6; ++f) OUT[f * P + c] = acc[f];
}

// --------------------------------- heavy ghosts (FarField / Outflow / Wall)

// E_r geometry of the heavy-ghost tables (GhostBC, src/bc1.rs:193-255): the
// sub-stencil r holds (r+1)^2 points from slot sum_{i<r} (i+1)^2; S_r row a
// is at sum_{i<r} (i+1)^4 + a (r+1)^2, w_r row k at sum_{i<r} (i+1)^3 +
// k (r+1)^2.
__device__ __forceinline__ int er_of(int a) { return (a >= 1) + (a >= 5) + (a >= 14) + (a >= 30); }
__device__ __forceinline__ int er_seg(int r) { return r * (r + 1) * (2 * r + 1) / 6; }
__device__ __forceinline__ int er_soff(int r) { return r * (r + 1) * (2 * r + 1) * (3 * r * r + 3 * r - 1) / 30; }
__device__ __forceinline__ int er_woff(int r) { return r * r * (r + 1) * (r + 1) / 4; }

// The recomputed forward quantities of one ghost and their cotangents.
struct HeavyLds {
  double u[56][6];       // E_r point states in the (n, t) frame; 55: u_inf (FarField)
  double L[6][6], R[6][6], lam[6];  // rows of L, R[i][m], eigenvalues
  double d[6][56];       // L . u
  double sd[6][55];      // (S_r d)[a]
  double v[5][5][6];     // v_{k,r}[c] (0 for k > r)
  double beta[5][6], alpha[5][6], om[5][6], asum[6];
  double V[5][6], rhs[5][6];
  double wV[5][6], wR[6][6], wL[6][6], wom[5][6];
  double wv[5][5][6], wbeta[5][6];
  double wd[6][56];
  double wLw[6][6];      // Wall: cotangents of the rows of L through the ILW systems
};

// rhs_1[0] = l1 / lambda1 of the Outflow relaxation (src/bc1.rs:783-787)
// from u^(0) in the (n, t) frame, and its gradient g; false (g = 0) when the
// u - c characteristic is outgoing and nothing is replaced.
__device__ bool outflow_relax(const Phys& ph, const double uk[6], double p_inf, double sigma, double l_dom,
                              double* phi, double g[6]) {
  const double rho = uk[0], un = uk[1] / rho, ut = uk[2] / rho;
  const double w2 = un * un + ut * ut;
  const double e[3] = {uk[3] / rho - w2 / 6.0, uk[4] / rho - w2 / 6.0, uk[5] / rho - w2 / 6.0};
  const double gk[3] = {ph.ge, ph.gi, ph.gr};
  const double gam[3] = {ph.gamma_e, ph.gamma_i, ph.gamma_r};
  const double Pm = gk[0] * e[0] + gk[1] * e[1] + gk[2] * e[2];
  const double p0 = rho * Pm;
  const double a2 = gam[0] * gk[0] * e[0] + gam[1] * gk[1] * e[1] + gam[2] * gk[2] * e[2];
  const double a = sqrt(a2);
  const double lam1 = un - a;
#pragma unroll
  for (int f = 0; f < 6; ++f) g[f] = 0.0;
  if (!(lam1 < 0.0)) return false;
  const double K1 = sigma / l_dom, dp = p0 - p_inf;
  const double f1 = 1.0 - un * un / a2;
  const double ml = fabs(un / a);
  *phi = sigma * (1.0 - sq(ml)) * a / l_dom * dp / lam1;  // the forward's value
  // d phi / d (un, a2, p0)
  const double g_un = (K1 * a * dp * (-2.0 * un / a2) - *phi) / lam1;
  const double g_a2 = (K1 * dp * (a * un * un / (a2 * a2) + f1 / (2.0 * a)) + *phi / (2.0 * a)) / lam1;
  const double g_p0 = K1 * f1 * a / lam1;
  // to u^(0)
  double grho = g_p0 * Pm, gw2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double ge = g_p0 * rho * gk[k] + g_a2 * gam[k] * gk[k];
    g[3 + k] = ge / rho;
    grho -= ge * uk[3 + k] / (rho * rho);
    gw2 -= ge / 6.0;
  }
  const double gun = g_un + 2.0 * un * gw2, gut = 2.0 * ut * gw2;
  g[1] = gun / rho;
  g[2] = gut / rho;
  grho -= (gun * un + gut * ut) / rho;
  g[0] = grho;
  return true;
}

// Pullback of one FarField / Outflow / Wall ghost value
// (weno_extrapolation_pre and the BC branches, src/bc1.rs:193-255, 294-312,
// 691-928) to the states it reads: its 55 E_r points and its nearest point.  One 64-lane workgroup per
// ghost recomputes the forward quantities from the stage state with the
// forward kernel's own operations (so the derivative is taken at the very
// values the forward formed: the WENO weights d_r / (eps + beta_r)^q make
// it sensitive to the last bits of beta_r) and sweeps them in reverse.  The
// eigensystem is that of the Roe average of the nearest state with itself,
// whose derivative with respect to that state's (u, v, e_e, e_i, e_r) is the
// identity, so the eigensystem pullback chains to the state through
// derived_vjp.  The incoming-characteristic choice of
// FarField and the Outflow branch take the derivative of the branch the
// forward value takes; Wall's two ILW systems (x = A^-1 b) pull back as
// bbar = A^-T xbar, Abar = -bbar x^T, the ILW row's entries being functions
// of u^(0).  GB: 56 x 6 per ghost (the E_r points a = 0..54, then
// the nearest point), in the (x, y) frame.
__global__ __launch_bounds__(64) void k_adj_heavy(AdjDesc A, HeavyGhosts hg, const EigConst* __restrict__ Kp,
                                                  Phys ph, double* __restrict__ GB) {
  __shared__ HeavyLds T;
  const int h = blockIdx.x, lane = threadIdx.x;
  const long P = A.g.plane;
  const double* __restrict__ S = A.S;
  const long row = hg.tab[h];
  const double nx = hg.normal[2L * h], ny = hg.normal[2L * h + 1];
  const bool far = hg.kind[h] == RH3T_BC_FARFIELD, wall = hg.kind[h] == RH3T_BC_WALL;
  const double* Sh = hg.S + 979L * row;
  const double* Wh = hg.W + 225L * row;
  double s1[6];
  {
    const long q = hg.nearest[h];
    double val[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) val[f] = S[f * P + q];
    s1[0] = val[0];
    s1[1] = val[1] * nx + val[2] * ny;
    s1[2] = -val[1] * ny + val[2] * nx;
    s1[3] = val[3];
    s1[4] = val[4];
    s1[5] = val[5];
  }
  // the forward's eigensystem (roe_x, lx_row, rx_row_rt: the same values as
  // k_ghost_open), and its intermediates for the pullback
  RoeV Rv;
  RoeX Q;
  {
    double d1[D_N];
    derived_from_state(ph, s1, d1, false);
    const double e[3] = {d1[D_EE], d1[D_EI], d1[D_ER]};
    roe_values<0>(*fresh(Kp), d1[D_U], d1[D_V], e, Rv);
    roe_x(ph, s1[0], s1[0], d1[D_U], d1[D_V], d1[D_EE], d1[D_EI], d1[D_ER], d1[D_U], d1[D_V], d1[D_EE], d1[D_EI],
          d1[D_ER], Q);
  }
  if (lane < 6) {
    double rw[6];
    lx_row(ph, Q, lane, rw);
#pragma unroll
    for (int m = 0; m < 6; ++m) T.L[lane][m] = rw[m];
    rx_row_rt(ph, Q, lane, rw);
#pragma unroll
    for (int m = 0; m < 6; ++m) T.R[lane][m] = rw[m];
    T.lam[lane] = lamx(Q, lane);
  }
  if (lane < 56) {
    double s6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (lane < 55) {
      const long q = hg.stencil[55L * row + lane];
#pragma unroll
      for (int f = 0; f < 6; ++f) s6[f] = S[f * P + q];
    } else if (far) {
#pragma unroll
      for (int f = 0; f < 6; ++f) s6[f] = hg.bcstate[6L * h + f];
    }
    T.u[lane][0] = s6[0];
    T.u[lane][1] = s6[1] * nx + s6[2] * ny;
    T.u[lane][2] = -s6[1] * ny + s6[2] * nx;
    T.u[lane][3] = s6[3];
    T.u[lane][4] = s6[4];
    T.u[lane][5] = s6[5];
  }
  __syncthreads();
  if (lane < 56)
#pragma unroll
    for (int c = 0; c < 6; ++c) T.d[c][lane] = dot6(T.L[c], T.u[lane]);
  __syncthreads();
  if (lane < 55) {  // (S_r d)[a]
    const int a = lane, r = er_of(a), m = (r + 1) * (r + 1), seg = er_seg(r);
    const double* tr = Sh + er_soff(r) + (a - seg) * m;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double acc = 0.0;
      for (int b = 0; b < m; ++b) acc += tr[b] * T.d[c][seg + b];
      T.sd[c][a] = acc;
    }
  }
  if (lane < 25) {  // v_{k,r} = w_r row k . d
    const int r = lane / 5, k = lane % 5, m = (r + 1) * (r + 1), seg = er_seg(r);
    const double* tr = Wh + er_woff(r) + k * m;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double acc = 0.0;
      if (k <= r)
        for (int b = 0; b < m; ++b) acc += tr[b] * T.d[c][seg + b];
      T.v[k][r][c] = acc;
    }
  }
  __syncthreads();
  if (lane < 30) {
    const int r = lane / 6, c = lane % 6, m = (r + 1) * (r + 1), seg = er_seg(r);
    double beta = 0.0;
    for (int a = 0; a < m; ++a) beta += T.d[c][seg + a] * T.sd[c][seg + a];
    T.beta[r][c] = beta;
    T.alpha[r][c] = hg.d5[5L * row + r] / weno_pow(ph.eps + beta, ph.weno_q);
  }
  __syncthreads();
  if (lane < 6) {
    double as = 0.0;
#pragma unroll
    for (int r = 0; r < 5; ++r) as += T.alpha[r][lane];
    T.asum[lane] = as;
#pragma unroll
    for (int r = 0; r < 5; ++r) T.om[r][lane] = T.alpha[r][lane] / as;
  }
  __syncthreads();
  if (lane < 30) {
    const int k = lane / 6, c = lane % 6;
    double V = 0.0;
#pragma unroll
    for (int r = 0; r < 5; ++r) V += T.om[r][c] * T.v[k][r][c];
    T.V[k][c] = V;
  }
  __syncthreads();
  if (lane == 0) {
    // the ghost cotangent in the (n, t) frame and the Taylor weights
    const long pad = hg.pad[h];
    const double w1 = A.SB[1 * P + pad], w2 = A.SB[2 * P + pad];
    const double wug[6] = {A.SB[pad], w1 * nx + w2 * ny, -w1 * ny + w2 * nx, A.SB[3 * P + pad],
                           A.SB[4 * P + pad], A.SB[5 * P + pad]};
    const double dist = hg.dist[h];
    double coef[5];
    coef[0] = 1.0;
#pragma unroll
    for (int k = 1; k < 5; ++k) coef[k] = coef[k - 1] * (dist / (double)k);
    double wR[6][6], wrhs[5][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        wR[i][m] = 0.0;
        T.wLw[i][m] = 0.0;
      }
#pragma unroll
    for (int c = 0; c < 6; ++c) T.wd[c][55] = 0.0;
    if (wall) {
      // forward, src/bc1.rs:691-756: u^(0) from L with row 0 replaced by
      // the zero normal momentum, u^(1) from L with row 0 the ILW row of
      // u^(0) (build_ilw_row1, src/bc1.rs:448-465), both right-hand sides
      // V^(k) with entry 0 set to 0; u^(k) = R V^(k) for k >= 2
      double A0[6][6], A1[6][6], b[6], u0[6], u1[6];
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) A0[r][q] = A1[r][q] = T.L[r][q];
#pragma unroll
      for (int q = 0; q < 6; ++q) A0[0][q] = q == 1 ? 1.0 : 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) b[m] = m == 0 ? 0.0 : T.V[0][m];
      bool ok = solve6(A0, b, u0);
      const double rho = u0[0], mt = u0[2];
      A1[0][0] = -ph.gt * sq(mt) / (6.0 * sq(rho));
      A1[0][1] = 0.0;
      A1[0][2] = ph.gt * mt / (6.0 * rho);
      A1[0][3] = -ph.ge;
      A1[0][4] = -ph.gi;
      A1[0][5] = -ph.gr;
#pragma unroll
      for (int m = 0; m < 6; ++m) b[m] = m == 0 ? 0.0 : T.V[1][m];
      ok = ok && solve6(A1, b, u1);
      // reverse: k >= 2 through R
#pragma unroll
      for (int k = 2; k < 5; ++k)
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < 6; ++i) {
            acc += T.R[i][m] * coef[k] * wug[i];
            wR[i][m] += coef[k] * wug[i] * T.V[k][m];
          }
          wrhs[k][m] = acc;
        }
      // x = A^-1 b: bbar = A^-T xbar, Abar = -bbar x^T
      double At[6][6], xb[6], bb[6];
      double wu0[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) xb[i] = coef[1] * wug[i];
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) At[r][q] = A1[q][r];
      ok = ok && solve6(At, xb, bb);
#pragma unroll
      for (int m = 0; m < 6; ++m) wrhs[1][m] = m == 0 ? 0.0 : bb[m];
#pragma unroll
      for (int r = 1; r < 6; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) T.wLw[r][q] -= bb[r] * u1[q];
      {  // row 0 of A1 as a function of u^(0)
        const double a00 = -bb[0] * u1[0], a02 = -bb[0] * u1[2];
#pragma unroll
        for (int i = 0; i < 6; ++i) wu0[i] = wug[i];
        wu0[0] += a00 * ph.gt * sq(mt) / (3.0 * rho * rho * rho) - a02 * ph.gt * mt / (6.0 * rho * rho);
        wu0[2] += -a00 * ph.gt * mt / (3.0 * rho * rho) + a02 * ph.gt / (6.0 * rho);
      }
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) At[r][q] = A0[q][r];
      ok = ok && solve6(At, wu0, bb);
#pragma unroll
      for (int m = 0; m < 6; ++m) wrhs[0][m] = m == 0 ? 0.0 : bb[m];
#pragma unroll
      for (int r = 1; r < 6; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) T.wLw[r][q] -= bb[r] * u0[q];
      if (!ok)  // singular ILW system: the forward raised RH3T_ERR_SINGULAR
#pragma unroll
        for (int m = 0; m < 6; ++m) wrhs[0][m] = NAN;
#pragma unroll
      for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int m = 0; m < 6; ++m) T.wV[k][m] = wrhs[k][m];
    } else {
      // forward: rhs_k of the BC branch
#pragma unroll
      for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int m = 0; m < 6; ++m) T.rhs[k][m] = T.V[k][m];
      bool fix = false;
      double gphi[6];
      if (far) {
#pragma unroll
        for (int m = 0; m < 6; ++m)
          if (T.lam[m] < 0.0) {
            T.rhs[0][m] = T.d[m][55];
#pragma unroll
            for (int k = 1; k < 5; ++k) T.rhs[k][m] = 0.0;
          }
      } else {
        double u0[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) u0[i] = dot6(T.R[i], T.V[0]);
        double phi = 0.0;
        fix = outflow_relax(ph, u0, hg.outflow[3L * h], hg.outflow[3L * h + 1], hg.outflow[3L * h + 2], &phi,
                            gphi);
        if (fix) T.rhs[1][0] = phi;
      }
      // reverse through u^(k) = R rhs_k
      double wu0[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) wu0[i] = wug[i];
      for (int k = 4; k >= 0; --k) {
        double wuk[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) wuk[i] = k == 0 ? wu0[i] : coef[k] * wug[i];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < 6; ++i) {
            acc += T.R[i][m] * wuk[i];
            wR[i][m] += wuk[i] * T.rhs[k][m];
          }
          wrhs[k][m] = acc;
        }
        if (k == 1 && fix) {  // rhs_1[0] = phi(u^(0)), u^(0) = R V^(0)
#pragma unroll
          for (int i = 0; i < 6; ++i) wu0[i] += gphi[i] * wrhs[1][0];
          wrhs[1][0] = 0.0;
        }
      }
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const bool in = far && T.lam[m] < 0.0;
        if (in) T.wd[m][55] = wrhs[0][m];
#pragma unroll
        for (int k = 0; k < 5; ++k) T.wV[k][m] = in ? 0.0 : wrhs[k][m];
      }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int m = 0; m < 6; ++m) T.wR[i][m] = wR[i][m];
  }
  __syncthreads();
  if (lane < 30) {  // V_k = sum_r om_r v_{k,r}
    const int r = lane / 6, c = lane % 6;
    double wom = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      wom += T.wV[k][c] * T.v[k][r][c];
      T.wv[k][r][c] = T.om[r][c] * T.wV[k][c];
    }
    T.wom[r][c] = wom;
  }
  __syncthreads();
  if (lane < 30) {  // om_r = alpha_r / sum alpha, alpha_r = d_r / (eps + beta_r)^q
    const int r = lane / 6, c = lane % 6;
    double dot = 0.0;
#pragma unroll
    for (int q = 0; q < 5; ++q) dot += T.om[q][c] * T.wom[q][c];
    const double walpha = (T.wom[r][c] - dot) / T.asum[c];
    T.wbeta[r][c] = -walpha * ph.weno_q * T.alpha[r][c] / (ph.eps + T.beta[r][c]);
  }
  __syncthreads();
  if (lane < 55) {  // d at the E_r points: w rows and beta = d . S d
    const int p = lane, r = er_of(p), m = (r + 1) * (r + 1), seg = er_seg(r), b = p - seg;
    const double* sc = Sh + er_soff(r) + b;
    const double* wc = Wh + er_woff(r) + b;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double acc = 0.0;
      for (int k = 0; k <= r; ++k) acc += wc[k * m] * T.wv[k][r][c];
      double st = 0.0;
      for (int a = 0; a < m; ++a) st += sc[a * m] * T.d[c][seg + a];
      T.wd[c][p] = acc + T.wbeta[r][c] * (T.sd[c][p] + st);
    }
  }
  __syncthreads();
  if (lane < 55) {  // d = L . u: the E_r points' state cotangents, (x, y) frame
    double wu[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) acc += T.wd[c][lane] * T.L[c][f];
      wu[f] = acc;
    }
    double* o = GB + ((long)h * 56 + lane) * 6;
    o[0] = wu[0];
    o[1] = wu[1] * nx - wu[2] * ny;
    o[2] = wu[1] * ny + wu[2] * nx;
    o[3] = wu[3];
    o[4] = wu[4];
    o[5] = wu[5];
  }
  if (lane < 36) {
    const int c = lane / 6, f = lane % 6;
    double acc = T.wLw[c][f];
    for (int a = 0; a < 56; ++a) acc += T.wd[c][a] * T.u[a][f];
    T.wL[c][f] = acc;
  }
  __syncthreads();
  if (lane == 0) {  // the eigensystem to the nearest state
    RoeBar B;
    roe_bar_zero(B);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double Rb[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) Rb[i] = T.wR[i][c];
      row_col_bar<0>(*fresh(Kp), Rv, c, T.wL[c], Rb, 0.0, B);
    }
    double o5[5];
    roe_bar_final<0>(*fresh(Kp), Rv, B, o5);
    double db[D_N];
#pragma unroll
    for (int k = 0; k < D_N; ++k) db[k] = 0.0;
    db[D_U] = o5[0];
    db[D_V] = o5[1];
    db[D_EE] = o5[2];
    db[D_EI] = o5[3];
    db[D_ER] = o5[4];
    double sb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    derived_vjp(ph, s1, db, sb);
    double* o = GB + ((long)h * 56 + 55) * 6;
    o[0] = sb[0];
    o[1] = sb[1] * nx - sb[2] * ny;
    o[2] = sb[1] * ny + sb[2] * nx;
    o[3] = sb[3];
    o[4] = sb[4];
    o[5] = sb[5];
  }
}

// One lane per cell read by a heavy ghost: adds the ghosts' pullbacks to it
// (ghost, then point order).
__global__ __launch_bounds__(256) void k_adj_heavy_T(int nrows, const long* __restrict__ row_cell,
                                                     const int* __restrict__ row_ptr, const int* __restrict__ ent,
                                                     const double* __restrict__ GB, double* __restrict__ OUT,
                                                     long P) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const long c = row_cell[r];
  double acc[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) acc[f] = OUT[f * P + c];
  for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
    const double* g = GB + (long)ent[e] * 6;
#pragma unroll
    for (int f = 0; f < 6; ++f) acc[f] += g[f];
  }
#pragma unroll
  for (int f = 0; f < 6; ++f) OUT[f * P + c] = acc[f];
}

// acc[k] += sum over blocks of PG (fixed order: per-lane strided sums, then a
// tree); one workgroup.
__global__ __launch_bounds__(256) void k_adj_param(const double* __restrict__ PG, int nblocks, double* acc) {
  __shared__ double red[kNParam][256];
  double s[kNParam] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += 256)
#pragma unroll
    for (int k = 0; k < kNParam; ++k) s[k] += PG[b * kNParam + k];
#pragma unroll
  for (int k = 0; k < kNParam; ++k) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < kNParam; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kNParam) acc[threadIdx.x] += red[threadIdx.x][0];
}

__global__ void k_fill(double* p, long n, double v) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// host AoS (fluid cells) <-> padded SoA cotangent buffer
__global__ __launch_bounds__(256) void k_aos_to_buf(const double* __restrict__ aos, Geo g,
                                                    const uint8_t* __restrict__ pt, double* __restrict__ dst) {
  const long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lin >= (long)g.nx * g.ny) return;
  const long p = pad_index(g, lin / g.ny, lin % g.ny);
  const bool fl = pt[p] == 1;
#pragma unroll
  for (int f = 0; f < 6; ++f) dst[f * g.plane + p] = fl ? aos[6 * lin + f] : 0.0;
}

__global__ __launch_bounds__(256) void k_buf_to_aos(const double* __restrict__ src, Geo g,
                                                    const uint8_t* __restrict__ pt, double* __restrict__ aos) {
  const long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lin >= (long)g.nx * g.ny) return;
  const long p = pad_index(g, lin / g.ny, lin % g.ny);
  const bool fl = pt[p] == 1;
#pragma unroll
  for (int f = 0; f < 6; ++f) aos[6 * lin + f] = fl ? src[f * g.plane + p] : 0.0;
}

// dt of step n from the frozen schedule (and Field.time += dt when asked)
__global__ void k_set_dt(Scalars* sc, const double* __restrict__ dts, int n, int advance_time) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc->dt = dts[n];
    if (advance_time) sc->time = sc->time + dts[n];
  }
}

}  // namespace

// ======================================================================
// adjoint workspace of an Engine (single solve)
// ======================================================================

namespace rh3t_eng {

struct AdjWork {
  double* lam[4] = {nullptr, nullptr, nullptr, nullptr};  // cotangent buffers, 6 planes each
  double* XB = nullptr;
  double* YB = nullptr;
  double* SB = nullptr;
  double* PG = nullptr;
  int pg_blocks = 0;
  int n_out = 0;           // ghosts outside the fluid box
  long* out_pads = nullptr;
  double* pacc = nullptr;  // kNParam
  double* aos = nullptr;   // staging
  int nrows = 0;
  long* row_cell = nullptr;
  int* row_ptr = nullptr;
  long* ent_pad = nullptr;
  int* ent_kind = nullptr;
  double* ent_n = nullptr;
  // heavy ghosts: pullbacks (56 x 6 per ghost) and their transpose lists
  int n_heavy = 0;
  double* GB = nullptr;
  int hrows = 0;
  long* hrow_cell = nullptr;
  int* hrow_ptr = nullptr;
  int* hent = nullptr;
  double* dts = nullptr;   // frozen dt schedule (device)
  int dts_cap = 0;
  // the ghosts outside the fluid box are gathered on a second stream, beside
  // the tile gather (default priority; ordered by the two events)
  hipStream_t side = nullptr;
  hipEvent_t ev_iface = nullptr, ev_ghosts = nullptr;
  ~AdjWork() {
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipStreamDestroy(side);
    }
    if (ev_iface) (void)hipEventDestroy(ev_iface);
    if (ev_ghosts) (void)hipEventDestroy(ev_ghosts);
  }
};

void free_adjoint(Engine& E) {
  delete E.adj;  // device memory is owned by E.allocs
  E.adj = nullptr;
}

}  // namespace rh3t_eng

namespace {

int init_adjoint(Engine& E) {
  if (E.adj) return RH3T_OK;
  if (E.nsolve != 1) {
    std::snprintf(E.message, sizeof(E.message), "the adjoint covers single solves");
    return RH3T_ERR_UNSUPPORTED;
  }
  HIP_TRY(hipSetDevice(E.device));
  const SolveDesc& d = E.h_desc[0];
  const Geo& g = d.g;
  const SolveHost& sh = E.solves[0];
  auto is_fluid_slot = [&](long q) {
    const long i = q / g.PY - 4, j = q % g.PY - 4;
    return i >= 0 && i < g.nx && j >= 0 && j < g.ny && sh.fluid[i * g.ny + j];
  };
  // heavy-ghost transpose lists grouped by cell: (cell slot, ghost * 56 +
  // point), the points being the 55 E_r points and the nearest point
  std::vector<std::pair<long, int>> hents;
  for (int h = 0; h < E.heavy.n; ++h)
    for (int a = 0; a < 56; ++a) {
      const long q = a < 55 ? E.h_heavy_stencil[55L * E.h_heavy_tab[h] + a] : E.h_heavy_near[h];
      if (!is_fluid_slot(q)) {
        std::snprintf(E.message, sizeof(E.message), "heavy ghost reads a point that is not a fluid cell");
        return RH3T_ERR_UNSUPPORTED;
      }
      hents.push_back({q, h * 56 + a});
    }
  std::stable_sort(hents.begin(), hents.end(),
                   [](const std::pair<long, int>& a, const std::pair<long, int>& b) { return a.first < b.first; });
  std::vector<long> hrow_cell;
  std::vector<int> hrow_ptr, hent;
  for (size_t e = 0; e < hents.size(); ++e) {
    if (e == 0 || hents[e].first != hents[e - 1].first) {
      hrow_cell.push_back(hents[e].first);
      hrow_ptr.push_back((int)e);
    }
    hent.push_back(hents[e].second);
  }
  hrow_ptr.push_back((int)hents.size());
  // ghost transpose lists grouped by source cell (sources must be fluid)
  std::vector<std::pair<long, int>> ents;  // (source slot, ghost list index)
  for (int k = 0; k < E.cheap.n; ++k) {
    const int kind = E.h_cheap_kind[k];
    if (kind == RH3T_BC_TIME_DEPENDENT) {
      std::snprintf(E.message, sizeof(E.message), "the adjoint does not cover TimeDependent ghosts");
      return RH3T_ERR_UNSUPPORTED;
    }
    if (kind == RH3T_BC_CONSTANT) continue;
    const long src = E.h_cheap_src[k];
    const long i = src / g.PY - 4, j = src % g.PY - 4;
    if (i < 0 || i >= g.nx || j < 0 || j >= g.ny || !sh.fluid[i * g.ny + j]) {
      std::snprintf(E.message, sizeof(E.message), "ghost source is not a fluid cell");
      return RH3T_ERR_UNSUPPORTED;
    }
    ents.push_back({src, k});
  }
  std::stable_sort(ents.begin(), ents.end(),
                   [](const std::pair<long, int>& a, const std::pair<long, int>& b) { return a.first < b.first; });
  std::vector<long> row_cell, ent_pad;
  std::vector<int> row_ptr, ent_kind;
  std::vector<double> ent_n;
  for (size_t e = 0; e < ents.size(); ++e) {
    if (e == 0 || ents[e].first != ents[e - 1].first) {
      row_cell.push_back(ents[e].first);
      row_ptr.push_back((int)e);
    }
    const int k = ents[e].second;
    ent_pad.push_back(E.h_cheap_pad[k]);
    ent_kind.push_back(E.h_cheap_kind[k]);
    ent_n.push_back(E.h_cheap_normal[2 * k]);
    ent_n.push_back(E.h_cheap_normal[2 * k + 1]);
  }
  row_ptr.push_back((int)ents.size());
  // published in E.adj only once complete (device memory is owned by E.allocs)
  auto* W = new AdjWork();
  struct Guard {
    AdjWork* w;
    ~Guard() { delete w; }
  } guard{W};
  int st;
  const size_t plane6 = 6 * (size_t)g.plane;
  for (int b = 0; b < 4; ++b) {
    if ((st = dev_alloc(E, &W->lam[b], plane6))) return st;
    HIP_TRY(hipMemset(W->lam[b], 0, plane6 * sizeof(double)));
  }
  if ((st = dev_alloc(E, &W->SB, plane6))) return st;
  HIP_TRY(hipMemset(W->SB, 0, plane6 * sizeof(double)));
  if ((st = dev_alloc(E, &W->XB, 36 * (size_t)(g.nx + 1) * g.ny))) return st;
  if ((st = dev_alloc(E, &W->YB, 36 * (size_t)g.nx * (g.ny + 1)))) return st;
  W->pg_blocks = ((g.nx + kGy - 1) / kGy) * ((g.ny + kGx - 1) / kGx);
  std::vector<long> out_pads;
  for (int k = 0; k < E.cheap.n + E.heavy.n; ++k) {
    const long p = k < E.cheap.n ? E.h_cheap_pad[k] : E.h_heavy_pad[k - E.cheap.n];
    const long ip = p / g.PY, jp = p % g.PY;
    if (ip < 4 || ip >= g.nx + 4 || jp < 4 || jp >= g.ny + 4) out_pads.push_back(p);
  }
  W->n_out = (int)out_pads.size();
  if (W->n_out > 0 && (st = upload(E, &W->out_pads, out_pads))) return st;
  if ((st = dev_alloc(E, &W->PG, (size_t)kNParam * W->pg_blocks))) return st;
  if ((st = dev_alloc(E, &W->pacc, (size_t)kNParam))) return st;
  if ((st = dev_alloc(E, &W->aos, 6 * (size_t)g.nx * g.ny))) return st;
  W->nrows = (int)row_cell.size();
  if ((st = upload(E, &W->row_cell, row_cell))) return st;
  if ((st = upload(E, &W->row_ptr, row_ptr))) return st;
  if ((st = upload(E, &W->ent_pad, ent_pad))) return st;
  if ((st = upload(E, &W->ent_kind, ent_kind))) return st;
  if ((st = upload(E, &W->ent_n, ent_n))) return st;
  W->n_heavy = E.heavy.n;
  if (W->n_heavy > 0) {
    if ((st = dev_alloc(E, &W->GB, (size_t)W->n_heavy * 56 * 6))) return st;
    W->hrows = (int)hrow_cell.size();
    if ((st = upload(E, &W->hrow_cell, hrow_cell))) return st;
    if ((st = upload(E, &W->hrow_ptr, hrow_ptr))) return st;
    if ((st = upload(E, &W->hent, hent))) return st;
  }
  if (W->n_out > 0) {
    HIP_TRY(hipStreamCreateWithFlags(&W->side, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&W->ev_iface, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&W->ev_ghosts, hipEventDisableTiming));
  }
  guard.w = nullptr;
  E.adj = W;
  // the gather needs the temperatures of every stage state
  if (E.nplanes != D_N) {
    E.nplanes = D_N;
    E.dvalid = 0;  // the kept Derived planes lack the temperatures
    if (E.graph) {
      (void)hipGraphExecDestroy(E.graph);
      E.graph = nullptr;
    }
  }
  return RH3T_OK;
}

// OUT = a LAM + b1 X1 + b2 X2 + rscale (dl/dS)^T LAM for the state in
// buffer buf; parameter cotangents accumulate into W.pacc.  taped: the
// state's ghost points and Derived planes are still those the forward stage
// that read it formed (otherwise they are formed first).
void launch_vjp(Engine& E, int buf, bool taped, const double* LAM, double rscale, double a, const double* X1,
                double b1, const double* X2, double b2, double* OUT) {
  AdjWork& W = *E.adj;
  const SolveDesc& d = E.h_desc[0];
  const Geo& g = d.g;
  if (!taped) {
    launch_ghosts(E, buf, 2u, GT_FIELD);
    launch_derived(E, buf);
  }
  AdjDesc A;
  A.g = g;
  A.pt = d.pt;
  A.S = d.U[buf];
  A.D = d.D;
  A.LAM = LAM;
  A.XB = W.XB;
  A.YB = W.YB;
  A.SB = W.SB;
  A.OUT = OUT;
  A.X1 = X1;
  A.X2 = X2;
  A.a = a;
  A.b1 = b1;
  A.b2 = b2;
  A.rscale = rscale;
  A.PG = W.PG;
  {
    Launcher L(E, K_ADJ_IFACE);
    // a workgroup of kIfBx lanes covers kIfPair interfaces
    const int nbx = (int)(((long)(g.nx + 1) * g.ny + kIfPair - 1) / kIfPair);
    const int nby = (int)(((long)g.nx * (g.ny + 1) + kIfPair - 1) / kIfPair);
    k_adj_iface<<<nbx + nby, kIfBx, 0, E.stream>>>(A, E.d_eig, E.phys, nbx);
  }
  if (W.n_out > 0) {
    // ghost gather on the side stream, concurrently with the tile gather
    // (it needs few registers and no LDS, so it fits beside it)
    (void)hipEventRecord(W.ev_iface, E.stream);
    (void)hipStreamWaitEvent(W.side, W.ev_iface, 0);
    const hipStream_t main = E.stream;
    E.stream = W.side;
    {
      Launcher L(E, K_ADJ_GATHER);
      k_adj_gather_ghosts<<<(W.n_out + 63) / 64, 64, 0, E.stream>>>(A, E.phys, W.n_out, W.out_pads);
    }
    E.stream = main;
    (void)hipEventRecord(W.ev_ghosts, W.side);
  }
  {
    Launcher L(E, K_ADJ_GATHER);
    k_adj_gather<<<W.pg_blocks, dim3(kGx, kGy), 0, E.stream>>>(A, E.phys);
  }
  if (W.n_out > 0) (void)hipStreamWaitEvent(E.stream, W.ev_ghosts, 0);
  if (W.nrows > 0) {
    Launcher L(E, K_ADJ_GHOST);
    k_adj_ghost_T<<<(W.nrows + 255) / 256, 256, 0, E.stream>>>(W.nrows, W.row_cell, W.row_ptr, W.ent_pad,
                                                               W.ent_kind, W.ent_n, W.SB, OUT, g.plane);
  }
  if (W.n_heavy > 0) {
    Launcher L(E, K_ADJ_GHOST);
    k_adj_heavy<<<W.n_heavy, 64, 0, E.stream>>>(A, E.heavy, E.d_eig, E.phys, W.GB);
    k_adj_heavy_T<<<(W.hrows + 255) / 256, 256, 0, E.stream>>>(W.hrows, W.hrow_cell, W.hrow_ptr, W.hent, W.GB,
                                                               OUT, g.plane);
  }
}

int begin_adjoint_call(Engine& E) {
  int st = init_adjoint(E);
  if (st) return st;
  HIP_TRY(hipSetDevice(E.device));
  if ((st = reset_errors(E))) return st;
  E.dvalid = 0;  // adjoint stages write the state buffers without the kept planes
  AdjWork& W = *E.adj;
  const long npg = (long)kNParam * W.pg_blocks;
  k_fill<<<(unsigned)((npg + 255) / 256), 256, 0, E.stream>>>(W.PG, npg, 0.0);
  k_fill<<<1, 64, 0, E.stream>>>(W.pacc, kNParam, 0.0);
  HIP_TRY(hipGetLastError());
  return RH3T_OK;
}

int upload_cot(Engine& E, const double* aos, double* dst) {
  const SolveDesc& d = E.h_desc[0];
  const long ncell = (long)d.g.nx * d.g.ny;
  HIP_TRY(hipMemcpyAsync(E.adj->aos, aos, 6 * ncell * sizeof(double), hipMemcpyHostToDevice, E.stream));
  k_aos_to_buf<<<(unsigned)((ncell + 255) / 256), 256, 0, E.stream>>>(E.adj->aos, d.g, d.pt, dst);
  HIP_TRY(hipGetLastError());
  return RH3T_OK;
}

int download_cot(Engine& E, const double* src, double* aos, double* pgrad) {
  const SolveDesc& d = E.h_desc[0];
  const long ncell = (long)d.g.nx * d.g.ny;
  k_buf_to_aos<<<(unsigned)((ncell + 255) / 256), 256, 0, E.stream>>>(src, d.g, d.pt, E.adj->aos);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(aos, E.adj->aos, 6 * ncell * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  if (pgrad) {
    {
      Launcher L(E, K_ADJ_PARAM);
      k_adj_param<<<1, 256, 0, E.stream>>>(E.adj->PG, E.adj->pg_blocks, E.adj->pacc);
    }
    HIP_TRY(hipMemcpyAsync(pgrad, E.adj->pacc, kNParam * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  }
  HIP_TRY(hipStreamSynchronize(E.stream));
  return RH3T_OK;
}

int ensure_dts(Engine& E, int n) {
  AdjWork& W = *E.adj;
  if (W.dts_cap >= n) return RH3T_OK;
  int st = dev_alloc(E, &W.dts, (size_t)n);
  if (st) return st;
  W.dts_cap = n;
  return RH3T_OK;
}

// Reverse sweep of one RK3 step whose stage states sit in U[0] (U^n),
// U[1] (S1) and U[2] (S2):
//   S2b  = 2/3 lam + (2dt/3) J(S2)^T lam
//   S1b  = 1/4 S2b + (dt/4) J(S1)^T S2b
//   out  = 1/3 lam + 3/4 S2b + S1b + dt J(U^n)^T S1b
// on(b) is called before the VJP of the stage state in buffer b.
template <class Select>
void launch_rk3_reverse(Engine& E, bool taped, double dt, const double* lam, double* s2b, double* s1b,
                        double* out, Select on) {
  on(2);
  launch_vjp(E, 2, taped, lam, 2.0 * dt / 3.0, 2.0 / 3.0, nullptr, 0.0, nullptr, 0.0, s2b);
  on(1);
  launch_vjp(E, 1, taped, s2b, dt / 4.0, 0.25, nullptr, 0.0, nullptr, 0.0, s1b);
  on(0);
  launch_vjp(E, 0, taped, s1b, dt, 1.0, lam, 1.0 / 3.0, s2b, 0.75, out);
}

}  // namespace

// Buffers of a checkpointed rollout with checkpoint interval K.  Segment k
// (steps kK .. kK+K-1) is taped in set k % 2 of two tape sets.  Step n
// (segment k = n / K, slot m = n % K) reads its state U^n from base(n) and
// writes its stage states to s1, s2 of its set and slot and U^{n+1} to
// base(n + 1), where base(n) is the checkpoint ckpt[k] when m == 0, the
// set's mid[m] otherwise, and base(N) is the context's own state.  Every
// step and stage state b has its own device descriptor with these buffers as
// U[0..3] and the slot's Derived planes of state b as D, so neither the
// forward nor the recompute copies a state, and the reverse sweep finds each
// stage state's ghost values and Derived planes as its forward stage left
// them.  The forward leaves its last two segments on the tape; the backward
// recomputes each earlier segment on a second stream while the reverse sweep
// of the segment after it runs on the context's stream.
struct rh3t_adjoint {
  rh3t_ctx* ctx = nullptr;
  int every = 5;
  int nsteps = 0;
  std::vector<double> dts;
  rh3t_physics phys{};           // constants of the last forward (the backward
                                 // differentiates the forward it follows)
  std::vector<double*> ckpt;     // U^{kK}, k = 0, 1, ...
  std::vector<double*> s1, s2;   // stage states, [set * K + m]
  std::vector<double*> mid;      // U^{kK + m}, m = 1 .. K-1, [set * K + m] (m = 0 unused)
  std::vector<double*> dpl;      // Derived planes of stage state b, [(set * K + m) * 3 + b]
  std::vector<SolveDesc> h_steps;  // [3 n + b]
  SolveDesc* d_steps = nullptr;
  int steps_cap = 0;
  int tape_of[2] = {-1, -1};     // segment whose stage states each set holds
  hipStream_t side = nullptr;    // recompute stream (default priority)
  hipEvent_t ev_start = nullptr, ev_free[2] = {nullptr, nullptr}, ev_ready[2] = {nullptr, nullptr};
  hipEvent_t ev_join = nullptr;
  // device time of the last forward / backward, from after the host inputs
  // are on the device to before the results leave it
  hipEvent_t t_fwd[2] = {nullptr, nullptr}, t_bwd[2] = {nullptr, nullptr};
  bool timed_fwd = false, timed_bwd = false;
  std::vector<void*> allocs;
};

namespace {

int adj_alloc(rh3t_adjoint* A, double** p, size_t n) {
  Engine& E = A->ctx->e;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, n * sizeof(double));
  if (e != hipSuccess) {
    std::snprintf(E.message, sizeof(E.message), "hipMalloc(%zu): %s", n * sizeof(double), hipGetErrorString(e));
    return RH3T_ERR_HIP;
  }
  A->allocs.push_back(q);
  // cells outside the fluid and its ghosts are never written by a step
  e = hipMemset(q, 0, n * sizeof(double));
  if (e != hipSuccess) {
    std::snprintf(E.message, sizeof(E.message), "hipMemset: %s", hipGetErrorString(e));
    return RH3T_ERR_HIP;
  }
  *p = (double*)q;
  return RH3T_OK;
}

// While in scope, launches run on the buffers of step n; at(b) selects the
// descriptor whose Derived planes are those of stage state b.
struct OnStep {
  Engine& E;
  const rh3t_adjoint* A;
  int n;
  SolveDesc* d0;
  SolveDesc h0;
  OnStep(Engine& e, const rh3t_adjoint* a, int step) : E(e), A(a), n(step), d0(e.d_desc), h0(e.h_desc[0]) {
    at(0);
  }
  void at(int b) {
    E.d_desc = A->d_steps + 3 * n + b;
    E.h_desc[0] = A->h_steps[3 * n + b];
  }
  ~OnStep() {
    E.d_desc = d0;
    E.h_desc[0] = h0;
  }
};

// While in scope, launches go to stream st.
struct OnStream {
  Engine& E;
  hipStream_t s0;
  OnStream(Engine& e, hipStream_t st) : E(e), s0(e.stream) { E.stream = st; }
  ~OnStream() { E.stream = s0; }
};

const StageSpec kStage1{1, 0, 0, 1, 3}, kStage2{2, 0, 1, 2, 5}, kStage3{3, 0, 2, 3, 7};

// Recomputes segment k from its checkpoint onto tape set k % 2 (U^{n1} is
// the next checkpoint already; stage 3 of the last step only fills what the
// reverse sweep reads of S2).
void recompute_segment(Engine& E, const rh3t_adjoint* A, int k) {
  const int n0 = k * A->every, n1 = std::min(A->nsteps, n0 + A->every);
  for (int n = n0; n < n1; ++n) {
    k_set_dt<<<1, 64, 0, E.stream>>>(E.d_sc, E.adj->dts, n, 0);
    OnStep on(E, A, n);
    launch_stage(E, 0, kStage1);
    on.at(1);
    launch_stage(E, 1, kStage2);
    on.at(2);
    if (n + 1 < n1) {
      launch_stage(E, 2, kStage3);
    } else {
      launch_ghosts(E, 2, 6u, GT_STAGE3);
      launch_derived(E, 2);
    }
  }
}

}  // namespace

extern "C" {

int rh3t_ctx_set_physics(rh3t_ctx* ctx, const rh3t_physics* phys);

int rh3t_l_vjp(rh3t_ctx* ctx, const double* rbar, double* ubar, double* pgrad) {
  if (!ctx || !rbar || !ubar) return RH3T_ERR_ARG;
  Engine& E = ctx->e;
  int st = begin_adjoint_call(E);
  if (st) return st;
  AdjWork& W = *E.adj;
  if ((st = upload_cot(E, rbar, W.lam[0]))) return st;
  launch_vjp(E, 0, false, W.lam[0], 1.0, 0.0, nullptr, 0.0, nullptr, 0.0, W.lam[1]);
  HIP_TRY(hipGetLastError());
  if ((st = download_cot(E, W.lam[1], ubar, pgrad))) return st;
  return check_errors(E);
}

int rh3t_rk3_vjp(rh3t_ctx* ctx, double dt, const double* lam, double* ubar, double* pgrad) {
  if (!ctx || !lam || !ubar) return RH3T_ERR_ARG;
  if (!(dt > 0.0)) return RH3T_ERR_DT;
  Engine& E = ctx->e;
  int st = begin_adjoint_call(E);
  if (st) return st;
  AdjWork& W = *E.adj;
  if ((st = ensure_dts(E, 1))) return st;
  HIP_TRY(hipMemcpyAsync(W.dts, &dt, sizeof(double), hipMemcpyHostToDevice, E.stream));
  k_set_dt<<<1, 64, 0, E.stream>>>(E.d_sc, W.dts, 0, 0);
  // forward stages 1 and 2 (U[0] is left untouched)
  launch_stage(E, 0, StageSpec{1, 0, 0, 1, 3});
  launch_stage(E, 1, StageSpec{2, 0, 1, 2, 5});
  launch_capture(E, 2, 5);
  if ((st = upload_cot(E, lam, W.lam[0]))) return st;
  launch_rk3_reverse(E, false, dt, W.lam[0], W.lam[1], W.lam[2], W.lam[3], [](int) {});
  HIP_TRY(hipGetLastError());
  if ((st = download_cot(E, W.lam[3], ubar, pgrad))) return st;
  return check_errors(E);
}

int rh3t_adjoint_create(rh3t_ctx* ctx, int32_t checkpoint_every, rh3t_adjoint** out) {
  if (!ctx || !out || checkpoint_every < 1) return RH3T_ERR_ARG;
  *out = nullptr;
  Engine& E = ctx->e;
  int st = init_adjoint(E);
  if (st) return st;
  auto* A = new rh3t_adjoint();
  A->ctx = ctx;
  A->every = checkpoint_every;
  const size_t plane = (size_t)E.h_desc[0].g.plane;
  const int K = checkpoint_every;
  A->s1.assign(2 * K, nullptr);
  A->s2.assign(2 * K, nullptr);
  A->mid.assign(2 * K, nullptr);
  A->dpl.assign(6 * K, nullptr);
  for (int q = 0; q < 2 * K && !st; ++q) {
    st = adj_alloc(A, &A->s1[q], 6 * plane);
    if (!st) st = adj_alloc(A, &A->s2[q], 6 * plane);
    if (!st && q % K > 0) st = adj_alloc(A, &A->mid[q], 6 * plane);
    for (int b = 0; b < 3 && !st; ++b) st = adj_alloc(A, &A->dpl[3 * q + b], (size_t)D_N * plane);
  }
  hipError_t e = hipSuccess;
  if (!st) {
    e = hipStreamCreateWithFlags(&A->side, hipStreamNonBlocking);
    hipEvent_t* evs[6] = {&A->ev_start, &A->ev_free[0], &A->ev_free[1], &A->ev_ready[0], &A->ev_ready[1],
                          &A->ev_join};
    for (int q = 0; q < 6 && e == hipSuccess; ++q) e = hipEventCreateWithFlags(evs[q], hipEventDisableTiming);
    hipEvent_t* tev[4] = {&A->t_fwd[0], &A->t_fwd[1], &A->t_bwd[0], &A->t_bwd[1]};
    for (int q = 0; q < 4 && e == hipSuccess; ++q) e = hipEventCreate(tev[q]);
    if (e != hipSuccess) {
      std::snprintf(E.message, sizeof(E.message), "stream/event: %s", hipGetErrorString(e));
      st = RH3T_ERR_HIP;
    }
  }
  if (st) {
    rh3t_adjoint_destroy(A);
    return st;
  }
  *out = A;
  return RH3T_OK;
}

void rh3t_adjoint_destroy(rh3t_adjoint* A) {
  if (!A) return;
  (void)hipSetDevice(A->ctx->e.device);
  (void)hipStreamSynchronize(A->ctx->e.stream);
  if (A->side) {
    (void)hipStreamSynchronize(A->side);
    (void)hipStreamDestroy(A->side);
  }
  for (hipEvent_t ev : {A->ev_start, A->ev_free[0], A->ev_free[1], A->ev_ready[0], A->ev_ready[1], A->ev_join,
                        A->t_fwd[0], A->t_fwd[1], A->t_bwd[0], A->t_bwd[1]})
    if (ev) (void)hipEventDestroy(ev);
  for (void* p : A->allocs) (void)hipFree(p);
  delete A;
}

int rh3t_adjoint_forward(rh3t_adjoint* A, int32_t nsteps, const double* dt) {
  if (!A || nsteps < 1 || !dt) return RH3T_ERR_ARG;
  for (int n = 0; n < nsteps; ++n)
    if (!(dt[n] > 0.0)) return RH3T_ERR_DT;
  Engine& E = A->ctx->e;
  int st = begin_adjoint_call(E);
  if (st) return st;
  const SolveDesc d = E.h_desc[0];
  const size_t plane6 = 6 * (size_t)d.g.plane;
  const int K = A->every;
  const int nck = (nsteps + K - 1) / K;
  while ((int)A->ckpt.size() < nck) {
    double* p;
    if ((st = adj_alloc(A, &p, plane6))) return st;
    A->ckpt.push_back(p);
  }
  if (A->steps_cap < nsteps) {
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, sizeof(SolveDesc) * 3 * nsteps));
    A->allocs.push_back(q);
    A->d_steps = (SolveDesc*)q;
    A->steps_cap = nsteps;
  }
  auto slot = [&](int n) { return ((n / K) % 2) * K + n % K; };
  auto base = [&](int n) { return n == nsteps ? d.U[0] : (n % K == 0 ? A->ckpt[n / K] : A->mid[slot(n)]); };
  std::vector<SolveDesc> steps(3 * (size_t)nsteps, d);
  for (int n = 0; n < nsteps; ++n)
    for (int b = 0; b < 3; ++b) {
      SolveDesc& sd = steps[3 * n + b];
      sd.U[0] = base(n);
      sd.U[1] = A->s1[slot(n)];
      sd.U[2] = A->s2[slot(n)];
      sd.U[3] = base(n + 1);
      sd.D = A->dpl[3 * slot(n) + b];
    }
  if (steps.size() != A->h_steps.size() ||
      std::memcmp(steps.data(), A->h_steps.data(), sizeof(SolveDesc) * steps.size()) != 0) {
    HIP_TRY(hipStreamSynchronize(E.stream));
    HIP_TRY(hipStreamSynchronize(A->side));
    HIP_TRY(hipMemcpy(A->d_steps, steps.data(), sizeof(SolveDesc) * steps.size(), hipMemcpyHostToDevice));
    A->h_steps.swap(steps);
  }
  if ((st = ensure_dts(E, nsteps))) return st;
  A->dts.assign(dt, dt + nsteps);
  A->nsteps = nsteps;
  A->phys = E.src_phys;
  A->tape_of[0] = A->tape_of[1] = -1;
  // a backward still recomputing on the side stream must finish first
  HIP_TRY(hipEventRecord(A->ev_join, A->side));
  HIP_TRY(hipStreamWaitEvent(E.stream, A->ev_join, 0));
  HIP_TRY(hipMemcpyAsync(E.adj->dts, dt, nsteps * sizeof(double), hipMemcpyHostToDevice, E.stream));
  HIP_TRY(hipEventRecord(A->t_fwd[0], E.stream));
  HIP_TRY(hipMemcpyAsync(A->ckpt[0], d.U[0], plane6 * sizeof(double), hipMemcpyDeviceToDevice, E.stream));
  for (int n = 0; n < nsteps; ++n) {
    k_set_dt<<<1, 64, 0, E.stream>>>(E.d_sc, E.adj->dts, n, 1);
    OnStep on(E, A, n);
    // each stage's own buffer descriptors: capture its phases before the
    // next stage's descriptors replace them
    launch_stage(E, 0, kStage1);
    launch_capture(E, 2, 3);
    on.at(1);
    launch_stage(E, 1, kStage2);
    launch_capture(E, 4, 5);
    on.at(2);
    launch_stage(E, 2, kStage3);
    launch_capture(E, 6, 7);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(A->t_fwd[1], E.stream));
  A->timed_fwd = true;
  // the last two segments are on the tape
  A->tape_of[(nck - 1) % 2] = nck - 1;
  if (nck >= 2) A->tape_of[(nck - 2) % 2] = nck - 2;
  E.dt_partials_valid = true;
  return check_errors(E);
}

namespace {

bool same_physics(const rh3t_physics& a, const rh3t_physics& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

int adjoint_backward(rh3t_adjoint* A, const double* lam_final, double* lam_initial, double* pgrad);

}  // namespace

// The backward differentiates the last forward with the constants that
// forward ran with, even if the context's constants were replaced since
// (rh3t_ctx_set_physics); the context's own constants are restored after.
int rh3t_adjoint_backward(rh3t_adjoint* A, const double* lam_final, double* lam_initial, double* pgrad) {
  if (!A || !lam_final || !lam_initial) return RH3T_ERR_ARG;
  if (A->nsteps < 1) return RH3T_ERR_ARG;
  Engine& E = A->ctx->e;
  if (same_physics(A->phys, E.src_phys)) return adjoint_backward(A, lam_final, lam_initial, pgrad);
  const rh3t_physics own = E.src_phys;
  HIP_TRY(hipSetDevice(E.device));
  HIP_TRY(hipStreamSynchronize(E.stream));
  HIP_TRY(hipStreamSynchronize(A->side));
  int st = apply_physics(E, A->phys);
  if (st == RH3T_OK) st = adjoint_backward(A, lam_final, lam_initial, pgrad);
  // the backward's kernels read the device constants: let them finish
  (void)hipStreamSynchronize(A->side);
  (void)hipStreamSynchronize(E.stream);
  int st2 = apply_physics(E, own);
  return st != RH3T_OK ? st : st2;
}

namespace {

int adjoint_backward(rh3t_adjoint* A, const double* lam_final, double* lam_initial, double* pgrad) {
  Engine& E = A->ctx->e;
  int st = begin_adjoint_call(E);
  if (st) return st;
  AdjWork& W = *E.adj;
  const hipStream_t main = E.stream;
  HIP_TRY(hipMemcpyAsync(W.dts, A->dts.data(), A->nsteps * sizeof(double), hipMemcpyHostToDevice, main));
  // the side stream starts after everything queued so far; both tape sets
  // are free until a reverse sweep reads them
  HIP_TRY(hipEventRecord(A->ev_start, main));
  HIP_TRY(hipStreamWaitEvent(A->side, A->ev_start, 0));
  HIP_TRY(hipEventRecord(A->ev_free[0], main));
  HIP_TRY(hipEventRecord(A->ev_free[1], main));
  bool pending[2] = {false, false};
  // enqueue the recompute of segment k on the side stream unless its set
  // holds it already
  auto ensure = [&](int k) -> int {
    if (k < 0 || A->tape_of[k % 2] == k) return RH3T_OK;
    const int set = k % 2;
    A->tape_of[set] = -1;
    HIP_TRY(hipStreamWaitEvent(A->side, A->ev_free[set], 0));
    {
      OnStream os(E, A->side);
      recompute_segment(E, A, k);
    }
    HIP_TRY(hipEventRecord(A->ev_ready[set], A->side));
    A->tape_of[set] = k;
    pending[set] = true;
    return RH3T_OK;
  };
  const int nseg = (A->nsteps + A->every - 1) / A->every;
  if ((st = ensure(nseg - 1)) || (st = ensure(nseg - 2))) return st;
  double* lam = W.lam[0];
  double* nxt = W.lam[3];
  if ((st = upload_cot(E, lam_final, lam))) return st;
  HIP_TRY(hipEventRecord(A->t_bwd[0], main));
  for (int k = nseg - 1; k >= 0; --k) {
    const int n0 = k * A->every, n1 = std::min(A->nsteps, n0 + A->every);
    const int set = k % 2;
    if (pending[set]) {
      HIP_TRY(hipStreamWaitEvent(main, A->ev_ready[set], 0));
      pending[set] = false;
    }
    // reverse sweep over the segment: every stage state on the tape still
    // holds the ghost values and Derived planes of its forward stage
    for (int n = n1 - 1; n >= n0; --n) {
      OnStep on(E, A, n);
      launch_rk3_reverse(E, true, A->dts[n], lam, W.lam[1], W.lam[2], nxt, [&](int b) { on.at(b); });
      std::swap(lam, nxt);
    }
    HIP_TRY(hipEventRecord(A->ev_free[set], main));
    if ((st = ensure(k - 2))) return st;
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(A->ev_join, A->side));
  HIP_TRY(hipStreamWaitEvent(main, A->ev_join, 0));
  HIP_TRY(hipEventRecord(A->t_bwd[1], main));
  A->timed_bwd = true;
  if ((st = download_cot(E, lam, lam_initial, pgrad))) return st;
  // the recompute's stage 3 overwrote the dt minima
  E.dt_partials_valid = false;
  return check_errors(E);
}

}  // namespace

int rh3t_adjoint_device_ms(const rh3t_adjoint* A, double* forward_ms, double* backward_ms) {
  if (!A || !forward_ms || !backward_ms || !A->timed_fwd || !A->timed_bwd) return RH3T_ERR_ARG;
  Engine& E = A->ctx->e;
  HIP_TRY(hipSetDevice(E.device));
  float f = 0.0f, b = 0.0f;
  HIP_TRY(hipEventSynchronize(A->t_fwd[1]));
  HIP_TRY(hipEventSynchronize(A->t_bwd[1]));
  HIP_TRY(hipEventElapsedTime(&f, A->t_fwd[0], A->t_fwd[1]));
  HIP_TRY(hipEventElapsedTime(&b, A->t_bwd[0], A->t_bwd[1]));
  *forward_ms = f;
  *backward_ms = b;
  return RH3T_OK;
}

}  // extern "C"
