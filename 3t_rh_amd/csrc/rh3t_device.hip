// HIP (gfx950) implementation of l() and rk3_ssp() of colinhu123/3T_RH and
// the device half of the C ABI in include/rh3t.h.
//
// Per l() evaluation of a state S (src/main.rs:135-305):
//   k_ghost_cheap   Periodic / Reflective / Zeroth / Constant ghosts
//                   (src/bc1.rs:471-636, 930-1037): gather from a static
//                   source cell into the ghost's padded slot.
//   k_ghost_heavy   FarField / Outflow / Wall ghosts (src/bc1.rs:609-928,
//                   193-255): one wavefront per ghost; the E_r gather, the
//                   characteristic projection and the S_r quadratic forms are
//                   spread over the 64 lanes, the reductions keep the
//                   reference's summation order.
//   k_derived       Derived::from_state on every fluid cell and ghost
//                   (src/state.rs:241-275, src/main.rs:149-153).
//   k_flux<0|1>     characteristic WENO5 interface flux in x / y
//                   (src/main.rs:165-204 -> src/weno.rs:685-802), one
//                   interface per lane, streamed over the 6 characteristic
//                   fields so the 6-point U and F stencils stay in VGPRs.
//   k_assemble      flux differences + noncon x/y + source + diffusion
//                   (src/main.rs:251-304) fused with the RK stage combination
//                   and the admissibility check (src/main.rs:307-349).
//   k_dt_partial / k_dt_final   calc_global_dt as a two-level fmin
//                   reduction (src/main.rs:375-408, src/dt.rs:5-34) and the
//                   dt = 0.05*dt_cfl rule of src/main.rs:461, on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "flux.hpp"
#include "kernels.hpp"
#include "problem.hpp"

using namespace rh3t_dev;

namespace {

constexpr int kFluxBlockJ = 64;
constexpr int kFluxBlockI = 4;
constexpr int kStOff[6] = {0, 1, 5, 14, 30, 55};
constexpr int kSOffD[6] = {0, 1, 17, 98, 354, 979};
constexpr int kWOffD[6] = {0, 1, 9, 36, 100, 225};

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; give
// each XCD a contiguous run of tiles so that halo rows shared by
// neighbouring tiles stay in one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int full = (nb / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

__device__ __forceinline__ void raise_error(Scalars* sc, unsigned phase, unsigned kind, unsigned idx) {
  unsigned long long key = err_key(sc->step, phase, kind, idx);
  atomicMin(&sc->err, key);
}

// ------------------------------------------------------------ ghosts

struct CheapGhosts {
  int n;
  const long* pad;      // destination slot
  const long* src;      // source slot (reflective / zeroth / periodic)
  const int* kind;      // rh3t_bc_kind
  const double* normal; // 2 per ghost
  const double* cstate; // 6 per ghost (Constant)
  const int* gid;       // sorted ghost id
};

__global__ __launch_bounds__(256) void k_ghost_cheap(double* __restrict__ S, Geo g, CheapGhosts cg,
                                                     Scalars* sc, unsigned phase) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cg.n) return;
  int kind = cg.kind[k];
  long dst = cg.pad[k];
  double v[6];
  if (kind == RH3T_BC_CONSTANT) {
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = cg.cstate[6 * k + f];
  } else {
    long s = cg.src[k];
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = S[f * g.plane + s];
    if (kind == RH3T_BC_REFLECTIVE_WALL) {  // src/bc1.rs:1015-1036
      double nx = cg.normal[2 * k], ny = cg.normal[2 * k + 1];
      double mom_n = v[1] * nx + v[2] * ny;
      v[1] = v[1] - 2.0 * mom_n * nx;
      v[2] = v[2] - 2.0 * mom_n * ny;
    }
  }
  bool ok = true;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    S[f * g.plane + dst] = v[f];
    ok = ok && isfinite(v[f]);
  }
  if (!ok) raise_error(sc, phase, RH3T_ERR_GHOST, (unsigned)cg.gid[k]);
}

struct HeavyGhosts {
  int n;
  const long* pad;       // destination slot
  const int* kind;       // Wall / Outflow / FarField
  const int* gid;        // sorted ghost id
  const long* nearest;   // slot of GhostBC::nearest
  const double* normal;  // 2 per ghost
  const double* dist;    // signed distance D
  const double* bcstate; // 6 per ghost (FarField u_inf)
  const double* outflow; // 3 per ghost (p_inf, sigma, l_domain)
  const double* d5;      // 5 per ghost
  const long* stencil;   // 55 slots per ghost
  const double* S;       // 979 per ghost
  const double* W;       // 225 per ghost
};

// Small dense helpers for the per-ghost BC algebra (lane 0 only).
__device__ __forceinline__ void matvec6(const double M[6][6], const double x[6], double y[6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) y[i] = dot6(M[i], x);
}

// solve6 (LAPACK dgesv): LU with partial pivoting, first maximal pivot.
__device__ bool solve6(const double A0[6][6], const double b0[6], double x[6]) {
  double A[6][6], b[6];
  for (int i = 0; i < 6; ++i) {
    b[i] = b0[i];
    for (int j = 0; j < 6; ++j) A[i][j] = A0[i][j];
  }
  for (int c = 0; c < 6; ++c) {
    int p = c;
    double best = fabs(A[c][c]);
    for (int r = c + 1; r < 6; ++r)
      if (fabs(A[r][c]) > best) {
        best = fabs(A[r][c]);
        p = r;
      }
    if (best == 0.0) return false;
    if (p != c) {
      for (int q = 0; q < 6; ++q) {
        double t = A[c][q];
        A[c][q] = A[p][q];
        A[p][q] = t;
      }
      double t = b[c];
      b[c] = b[p];
      b[p] = t;
    }
    for (int r = c + 1; r < 6; ++r) {
      double f = A[r][c] / A[c][c];
      A[r][c] = f;
      for (int q = c + 1; q < 6; ++q) A[r][q] -= f * A[c][q];
      b[r] -= f * b[c];
    }
  }
  for (int r = 5; r >= 0; --r) {
    double s = b[r];
    for (int q = r + 1; q < 6; ++q) s -= A[r][q] * x[q];
    x[r] = s / A[r][r];
  }
  return true;
}

// One 64-lane workgroup per heavy ghost.  set_ghost_point_value's fast
// path (src/bc1.rs:663-928) evaluates weno_extrapolation_pre five times
// (k = 0..4) on identical data; the projected data and the smoothness
// indicators beta_r do not depend on k, so they are formed once here (the
// values are the same, the work is a fifth).
__global__ __launch_bounds__(64) void k_ghost_heavy(double* __restrict__ S, Geo g, HeavyGhosts hg,
                                                    Phys ph, Scalars* sc, unsigned phase) {
  const int h = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ double s_data[6][56];   // L . u6 at the 55 E_r points
  __shared__ double s_part[330];     // data[c][a] * (S_r data[c])[a]
  __shared__ double s_vk[5][6][5];   // [k][c][r]
  __shared__ double s_alpha[6][5];   // [c][r]
  __shared__ double s_v[5][6];       // V^(k) per characteristic

  const double nx = hg.normal[2 * h], ny = hg.normal[2 * h + 1];
  // build_local_plain, src/bc1.rs:294-312: rotate the nearest state to the
  // (n, t) frame and take the X-direction eigenvectors of the Roe average of
  // that state with itself.
  Eigen E;
  {
    long q = hg.nearest[h];
    double val[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) val[f] = S[f * g.plane + q];
    double s1[6] = {val[0], val[1] * nx + val[2] * ny, -val[1] * ny + val[2] * nx, val[3], val[4], val[5]};
    double d1[D_N];
    derived_from_state(ph, s1, d1, false);
    build_eigen<0>(ph, s1[0], s1[0], d1[D_U], d1[D_V], d1[D_EE], d1[D_EI], d1[D_ER], d1[D_U], d1[D_V],
                   d1[D_EE], d1[D_EI], d1[D_ER], E);
  }
  // data[c][a] = L[c] . (rho, m_n, m_t, ee, ei, er) at E_r point a
  if (lane < 55) {
    long q = hg.stencil[55L * h + lane];
    double s[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) s[f] = S[f * g.plane + q];
    double u6[6] = {s[0], s[1] * nx + s[2] * ny, -s[1] * ny + s[2] * nx, s[3], s[4], s[5]};
#pragma unroll
    for (int c = 0; c < 6; ++c) s_data[c][lane] = dot6(E.L[c], u6);
  }
  __syncthreads();
  const double* Sh = hg.S + 979L * h;
  const double* Wh = hg.W + 225L * h;
  // partial products of beta_r = sum_a data_a * (sum_b S_ab data_b)
  for (int q = lane; q < 330; q += 64) {
    // q enumerates (r, c, a) with r-major blocks of 6*m entries
    int r = 0, base = 0;
    while (q >= base + 6 * (r + 1) * (r + 1)) {
      base += 6 * (r + 1) * (r + 1);
      ++r;
    }
    int m = (r + 1) * (r + 1);
    int c = (q - base) / m, a = (q - base) % m;
    const double* row = Sh + kSOffD[r] + a * m;
    const double* dc = &s_data[c][kStOff[r]];
    double acc = 0.0;
    for (int b = 0; b < m; ++b) acc += row[b] * dc[b];
    s_part[q] = dc[a] * acc;
  }
  // v_k = w_rk . data  (k <= r), zero otherwise: 30 (r, c) x 5 k
  for (int q = lane; q < 150; q += 64) {
    int k = q / 30, rc = q % 30, r = rc / 6, c = rc % 6;
    int m = (r + 1) * (r + 1);
    double vk = 0.0;
    if (k <= r) {
      const double* w = Wh + kWOffD[r] + k * m;
      const double* dc = &s_data[c][kStOff[r]];
      for (int a = 0; a < m; ++a) vk += w[a] * dc[a];
    }
    s_vk[k][c][r] = vk;
  }
  __syncthreads();
  if (lane < 30) {  // beta_r and alpha_r = d_r / (eps + beta_r)^q
    int r = lane / 6, c = lane % 6;
    int m = (r + 1) * (r + 1);
    int base = 0;
    for (int rr = 0; rr < r; ++rr) base += 6 * (rr + 1) * (rr + 1);
    const double* p = &s_part[base + c * m];
    double beta = 0.0;
    for (int a = 0; a < m; ++a) beta += p[a];
    s_alpha[c][r] = hg.d5[5L * h + r] / pow(ph.eps + beta, ph.weno_q);
  }
  __syncthreads();
  if (lane < 30) {  // V^(k)_c = sum_r (alpha_r / sum alpha) v_k,r
    int k = lane / 6, c = lane % 6;
    double asum = 0.0;
#pragma unroll
    for (int r = 0; r < 5; ++r) asum += s_alpha[c][r];
    double res = 0.0;
#pragma unroll
    for (int r = 0; r < 5; ++r) res += (s_alpha[c][r] / asum) * s_vk[k][c][r];
    s_v[k][c] = res;
  }
  __syncthreads();
  if (lane != 0) return;

  double uk[5][6];
  const int kind = hg.kind[h];
  bool ok = true;
  if (kind == RH3T_BC_FARFIELD) {  // src/bc1.rs:829-924
    const double* ui = hg.bcstate + 6L * h;
    double uloc[6] = {ui[0], ui[1] * nx + ui[2] * ny, -ui[1] * ny + ui[2] * nx, ui[3], ui[4], ui[5]};
    double vinf[6];
    matvec6(E.L, uloc, vinf);
    double rhs[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) rhs[m] = E.lam[m] < 0.0 ? vinf[m] : s_v[0][m];
    matvec6(E.R, rhs, uk[0]);
    for (int k = 1; k <= 4; ++k) {
#pragma unroll
      for (int m = 0; m < 6; ++m) rhs[m] = E.lam[m] < 0.0 ? 0.0 : s_v[k][m];
      matvec6(E.R, rhs, uk[k]);
    }
  } else if (kind == RH3T_BC_OUTFLOW) {  // src/bc1.rs:760-828
    const double* of = hg.outflow + 3L * h;
    matvec6(E.R, s_v[0], uk[0]);
    double rho0 = uk[0][0];
    double un0 = uk[0][1] / rho0, ut0 = uk[0][2] / rho0;
    double w2 = un0 * un0 + ut0 * ut0;  // local_pressure_and_a2, src/bc1.rs:424-440
    double ee1 = uk[0][3] / rho0 - w2 / 6.0;
    double ei1 = uk[0][4] / rho0 - w2 / 6.0;
    double er1 = uk[0][5] / rho0 - w2 / 6.0;
    double p0 = rho0 * (ph.ge * ee1 + ph.gi * ei1 + ph.gr * er1);
    double a2 = ph.gamma_e * ph.ge * ee1 + ph.gamma_i * ph.gi * ei1 + ph.gamma_r * ph.gr * er1;
    double a = sqrt(a2);
    double lambda1 = un0 - a;
    double rhs1[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) rhs1[m] = s_v[1][m];
    if (lambda1 < 0.0) {
      double ml = fabs(un0 / a);
      double l1 = of[1] * (1.0 - sq(ml)) * a / of[2] * (p0 - of[0]);
      rhs1[0] = l1 / lambda1;
    }
    matvec6(E.R, rhs1, uk[1]);
    for (int k = 2; k <= 4; ++k) matvec6(E.R, s_v[k], uk[k]);
  } else {  // Wall-ILW, src/bc1.rs:691-756
    double A[6][6], rhs[6];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) A[i][j] = E.L[i][j];
    for (int c = 0; c < 6; ++c) A[0][c] = c == 1 ? 1.0 : 0.0;
    for (int m = 0; m < 6; ++m) rhs[m] = s_v[0][m];
    rhs[0] = 0.0;
    ok = solve6(A, rhs, uk[0]);
    double u1 = uk[0][0], u3 = uk[0][2];
    A[0][0] = -ph.gt * sq(u3) / (6.0 * sq(u1));  // build_ilw_row1, src/bc1.rs:448-465
    A[0][1] = 0.0;
    A[0][2] = ph.gt * u3 / (6.0 * u1);
    A[0][3] = -ph.ge;
    A[0][4] = -ph.gi;
    A[0][5] = -ph.gr;
    for (int m = 0; m < 6; ++m) rhs[m] = s_v[1][m];
    rhs[0] = 0.0;
    ok = ok && solve6(A, rhs, uk[1]);
    for (int k = 2; k <= 4; ++k) matvec6(E.R, s_v[k], uk[k]);
  }
  // Taylor expansion to the ghost point, rotation back to (x, y)
  double ug[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) ug[c] = uk[0][c];
  double coef = 1.0;
  const double dist = hg.dist[h];
  for (int k = 1; k <= 4; ++k) {
    coef *= dist / (double)k;
#pragma unroll
    for (int c = 0; c < 6; ++c) ug[c] = ug[c] + coef * uk[k][c];
  }
  double out[6] = {ug[0], ug[1] * nx - ug[2] * ny, ug[1] * ny + ug[2] * nx, ug[3], ug[4], ug[5]};
  long dst = hg.pad[h];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    S[f * g.plane + dst] = out[f];
    ok = ok && isfinite(out[f]);
  }
  if (!ok) raise_error(sc, phase, RH3T_ERR_GHOST, (unsigned)hg.gid[h]);
}

// ----------------------------------------------------------- derived

__global__ __launch_bounds__(256) void k_derived(const double* __restrict__ S, double* __restrict__ D,
                                                 const uint8_t* __restrict__ pt, long total, Geo g, Phys ph,
                                                 int nplanes) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total || pt[p] == 0) return;
  double s[6], d[D_N];
#pragma unroll
  for (int f = 0; f < 6; ++f) s[f] = S[f * g.plane + p];
  derived_from_state(ph, s, d, nplanes > 8);
#pragma unroll
  for (int f = 0; f < D_N; ++f)
    if (f < nplanes) D[f * g.plane + p] = d[f];
}

// ------------------------------------------------------- WENO flux

// Interface flux of one direction.  DIR 0: interface i-1/2 of row j,
// i in [0, nx], stencil points i-3..i+2.  DIR 1: interface j-1/2 of column
// i, j in [0, ny].  Output planes are i-major: x at i*ny + j, y at
// i*(ny+1) + j (the reference stores fy j-major; the layout is internal).
#ifndef RH3T_FLUX_WAVES
#define RH3T_FLUX_WAVES 2
#endif
template <int DIR, bool FAST>
__global__ __launch_bounds__(256, RH3T_FLUX_WAVES) void k_flux(const double* __restrict__ S, const double* __restrict__ D,
                                              double* __restrict__ F, const uint8_t* __restrict__ pt, Geo g,
                                              EigConst K, double eps, int tiles_j) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = (t % tiles_j) * kFluxBlockJ + threadIdx.x;
  const int i = (t / tiles_j) * kFluxBlockI + threadIdx.y;
  const int ni = DIR == 0 ? g.nx + 1 : g.nx;
  const int nj = DIR == 0 ? g.ny : g.ny + 1;
  if (i >= ni || j >= nj) return;
  const long fpl = DIR == 0 ? (long)(g.nx + 1) * g.ny : (long)g.nx * (g.ny + 1);
  const long fidx = DIR == 0 ? (long)i * g.ny + j : (long)i * (g.ny + 1) + j;
  const long c0 = pad_index(g, i, j);
  const long step = DIR == 0 ? g.PY : 1;
  if (pt[c0] != 1 && pt[c0 - step] != 1) {
#pragma unroll
    for (int f = 0; f < 6; ++f) F[f * fpl + fidx] = 0.0;
    return;
  }
  // 6-point stencils of U and of the physical flux (flux_from_derived,
  // src/state.rs:56-85)
  double U[6][6], Fl[6][6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const long p = c0 + (q - 3) * step;
#pragma unroll
    for (int f = 0; f < 6; ++f) U[q][f] = S[f * g.plane + p];
    const double u = D[D_U * g.plane + p], v = D[D_V * g.plane + p];
    const double pe = D[D_PE * g.plane + p], pi = D[D_PI * g.plane + p], pr = D[D_PR * g.plane + p];
    const double ptot = pe + pi + pr;
    const double w = DIR == 0 ? u : v;
    Fl[q][0] = U[q][0] * w;
    Fl[q][1] = DIR == 0 ? U[q][0] * u * u + ptot : U[q][0] * u * v;
    Fl[q][2] = DIR == 0 ? U[q][0] * u * v : U[q][0] * v * v + ptot;
    Fl[q][3] = (U[q][3] + pe) * w;
    Fl[q][4] = (U[q][4] + pi) * w;
    Fl[q][5] = (U[q][5] + pr) * w;
  }
  const long p2 = c0 - step;
  const double e3[3] = {D[D_EE * g.plane + c0], D[D_EI * g.plane + c0], D[D_ER * g.plane + c0]};
  const double e2[3] = {D[D_EE * g.plane + p2], D[D_EI * g.plane + p2], D[D_ER * g.plane + p2]};
  double out[6];
  interface_flux<DIR, FAST>(K, eps, U, Fl, D[D_U * g.plane + c0], D[D_V * g.plane + c0], e3,
                            D[D_U * g.plane + p2], D[D_V * g.plane + p2], e2, out);
#pragma unroll
  for (int f = 0; f < 6; ++f) F[f * fpl + fidx] = out[f];
}

// --------------------------------------------------------- assemble

struct StageArgs {
  int mode;               // 0: write rhs, 1..3: SSP-RK3 stage
  const double* base;     // u^n
  const double* add;      // u1 / u2 (stages 2, 3)
  double* dst;            // stage output
  double* rhs;            // mode 0 output, compact SoA nx*ny
  unsigned phase;
};

__global__ __launch_bounds__(256) void k_assemble(const double* __restrict__ D, const double* __restrict__ FX,
                                                  const double* __restrict__ FY,
                                                  const uint8_t* __restrict__ pt, Geo g, Phys ph,
                                                  StageArgs sa, Scalars* sc, int tiles_j) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int j = (t % tiles_j) * kFluxBlockJ + threadIdx.x;
  const int i = (t / tiles_j) * kFluxBlockI + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const long c0 = pad_index(g, i, j);
  const long lin = (long)i * g.ny + j;
  const long ncell = (long)g.nx * g.ny;
  if (pt[c0] != 1) {
    if (sa.mode == 0)
#pragma unroll
      for (int f = 0; f < 6; ++f) sa.rhs[f * ncell + lin] = 0.0;
    return;
  }
  const long fxpl = (long)(g.nx + 1) * g.ny, fypl = (long)g.nx * (g.ny + 1);
  const long ixl = (long)i * g.ny + j, ixr = ixl + g.ny;
  const long iyb = (long)i * (g.ny + 1) + j, iyt = iyb + 1;
  const double inv_dx = 1.0 / g.dx, inv_dy = 1.0 / g.dy;
  double r[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const double fxt = -(FX[f * fxpl + ixr] - FX[f * fxpl + ixl]) * inv_dx;
    const double fyt = -(FY[f * fypl + iyt] - FY[f * fypl + iyb]) * inv_dy;
    r[f] = fxt + fyt;
  }
  // non-conservative terms, src/noncon.rs:288-332
  double ncx[3], ncy[3];
  {
    double vel[9], pe[9], pi[9], pr[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const long p = c0 + (long)(k - 4) * g.PY;
      vel[k] = D[D_U * g.plane + p];
      pe[k] = D[D_PE * g.plane + p];
      pi[k] = D[D_PI * g.plane + p];
      pr[k] = D[D_PR * g.plane + p];
    }
    noncon(vel, pe, pi, pr, g.dx, ncx);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const long p = c0 + (k - 4);
      vel[k] = D[D_V * g.plane + p];
      pe[k] = D[D_PE * g.plane + p];
      pi[k] = D[D_PI * g.plane + p];
      pr[k] = D[D_PR * g.plane + p];
    }
    noncon(vel, pe, pi, pr, g.dy, ncy);
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) r[3 + e] = r[3 + e] + ncx[e] + ncy[e];
  if (ph.source) {  // source, src/source.rs:7-21
    const double te = D[D_TE * g.plane + c0], ti = D[D_TI * g.plane + c0], tr = D[D_TR * g.plane + c0];
    r[3] = r[3] + (-ph.omega_ei * (te - ti) - ph.omega_er * (pow4(te) - pow4(tr)));
    r[4] = r[4] + ph.omega_ei * (te - ti);
    r[5] = r[5] + ph.omega_er * (pow4(te) - pow4(tr));
  }
  if (ph.diffusion) {  // src/main.rs:209-246, 288-295 with src/diffusion.rs:27-50
    const double kap[3] = {ph.kappa_e, ph.kappa_i, ph.kappa_r};
    const double cx = -1.0 / (g.dx * g.dx), cy = -1.0 / (g.dy * g.dy);
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      double tx[7], ty[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        double a = D[(D_TE + e) * g.plane + c0 + (long)(k - 3) * g.PY];
        double b = D[(D_TE + e) * g.plane + c0 + (k - 3)];
        tx[k] = e == 2 ? pow4(a) : a;
        ty[k] = e == 2 ? pow4(b) : b;
      }
      const double dl = diffusion_flux(kap[e], tx), dr = diffusion_flux(kap[e], tx + 1);
      const double db = diffusion_flux(kap[e], ty), dt_ = diffusion_flux(kap[e], ty + 1);
      r[3 + e] = r[3 + e] + ((-(dr - dl) * cx) + (-(dt_ - db) * cy));
    }
  }
  if (sa.mode == 0) {
#pragma unroll
    for (int f = 0; f < 6; ++f) sa.rhs[f * ncell + lin] = r[f];
    return;
  }
  // stage_update_rhs / stage_update_comb, src/main.rs:307-349
  const double dt = sc->dt;
  double v[6];
  if (sa.mode == 1) {
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = sa.base[f * g.plane + c0] + r[f] * dt;
  } else if (sa.mode == 2) {
    const double coef = dt / 4.0;
#pragma unroll
    for (int f = 0; f < 6; ++f)
      v[f] = sa.base[f * g.plane + c0] * 0.75 + sa.add[f * g.plane + c0] * 0.25 + r[f] * coef;
  } else {
    const double coef = 2.0 * dt / 3.0;
#pragma unroll
    for (int f = 0; f < 6; ++f)
      v[f] = sa.base[f * g.plane + c0] * (1.0 / 3.0) + sa.add[f * g.plane + c0] * (2.0 / 3.0) + r[f] * coef;
  }
#pragma unroll
  for (int f = 0; f < 6; ++f) sa.dst[f * g.plane + c0] = v[f];
  if (!admissible(v)) raise_error(sc, sa.phase, RH3T_ERR_NONPHYSICAL, (unsigned)lin);
}

// Copies the state of the first failure of `phase` into the scalars block
// (one thread; runs right after the kernel that can fail).
__global__ void k_capture(const double* __restrict__ buf, Geo g, Scalars* sc, unsigned phase,
                          const long* __restrict__ ghost_pad) {
  unsigned long long key = sc->err;
  if (key == ~0ULL) return;
  if (((key >> 44) & 0xFFFFF) != (sc->step & 0xFFFFF) || ((key >> 40) & 0xF) != phase) return;
  unsigned idx = (unsigned)(key & 0xFFFFFFFFULL);
  long p;
  if (phase % 2 == 0) {
    p = ghost_pad[idx];
  } else {
    p = pad_index(g, idx / g.ny, idx % g.ny);
  }
  for (int f = 0; f < 6; ++f) sc->err_state[f] = buf[f * g.plane + p];
}

// ------------------------------------------------------------- dt

__global__ __launch_bounds__(256) void k_dt_partial(const double* __restrict__ S, const uint8_t* __restrict__ pt,
                                                    Geo g, Phys ph, double* __restrict__ partial) {
  __shared__ double red[256];
  long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double m = INFINITY;
  if (lin < (long)g.nx * g.ny) {
    long i = lin / g.ny, j = lin % g.ny;
    long p = pad_index(g, i, j);
    if (pt[p] == 1) {
      double s[6];
#pragma unroll
      for (int f = 0; f < 6; ++f) s[f] = S[f * g.plane + p];
      m = local_dt(ph, s, g.dx, g.dy);
    }
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// fmin is exact, so the two-level reduction equals the serial f64::min of
// src/main.rs:402-403 whatever the order.
__global__ __launch_bounds__(256) void k_dt_final(const double* __restrict__ partial, int n, Scalars* sc,
                                                  int advance) {
  __shared__ double red[256];
  double m = INFINITY;
  for (int k = threadIdx.x; k < n; k += 256) m = fmin(m, partial[k]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double dt_cfl = red[0];
    sc->dt_cfl = dt_cfl;
    if (advance) {
      sc->step = sc->step + 1;
      double dt = sc->dt_fixed > 0.0 ? sc->dt_fixed : sc->cfl_factor * dt_cfl;
      sc->dt = dt;
      if (!(dt > 0.0)) atomicMin(&sc->err, err_key(sc->step, 1, RH3T_ERR_DT, 0));
      sc->time = sc->time + dt;
    }
  }
}

__global__ void k_set_err_clean(Scalars* sc) { sc->err = ~0ULL; }

// ----------------------------------------------------------- layout I/O

// host AoS (i*ny+j, 6 doubles) <-> padded SoA on fluid cells
__global__ __launch_bounds__(256) void k_aos_to_pad(const double* __restrict__ aos, double* __restrict__ S,
                                                    const uint8_t* __restrict__ pt, Geo g) {
  long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lin >= (long)g.nx * g.ny) return;
  long p = pad_index(g, lin / g.ny, lin % g.ny);
  if (pt[p] != 1) return;
#pragma unroll
  for (int f = 0; f < 6; ++f) S[f * g.plane + p] = aos[6 * lin + f];
}

__global__ __launch_bounds__(256) void k_pad_to_aos(const double* __restrict__ S, double* __restrict__ aos,
                                                    const uint8_t* __restrict__ pt, Geo g) {
  long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lin >= (long)g.nx * g.ny) return;
  long p = pad_index(g, lin / g.ny, lin % g.ny);
  if (pt[p] != 1) return;
#pragma unroll
  for (int f = 0; f < 6; ++f) aos[6 * lin + f] = S[f * g.plane + p];
}

__global__ __launch_bounds__(256) void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos,
                                                    long n) {
  long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lin >= n) return;
#pragma unroll
  for (int f = 0; f < 6; ++f) aos[6 * lin + f] = soa[f * n + lin];
}

__global__ __launch_bounds__(256) void k_gather_ghosts(const double* __restrict__ S, const long* __restrict__ gpad,
                                                       int n, Geo g, double* __restrict__ out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
#pragma unroll
  for (int f = 0; f < 6; ++f) out[6L * k + f] = S[f * g.plane + gpad[k]];
}

}  // namespace

// ======================================================================
// Context
// ======================================================================

enum KernelId { K_GHOST_CHEAP, K_GHOST_HEAVY, K_DERIVED, K_FLUX_X, K_FLUX_Y, K_ASSEMBLE, K_DT_PARTIAL,
                K_DT_FINAL, K_CAPTURE, K_COUNT };
static const char* kKernelNames[K_COUNT] = {"ghost_cheap", "ghost_heavy", "derived", "flux_x", "flux_y",
                                            "assemble", "dt_partial", "dt_final", "capture"};

struct ProfRec {
  int kernel;
  hipEvent_t a, b;
};

struct rh3t_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Geo geo{};
  Phys phys{};
  EigConst eig{};
  bool fast = true;  // flux arithmetic variant (see flux.hpp)
  int nplanes = 8;  // derived planes stored
  long n_ghost = 0;
  std::vector<uint8_t> host_fluid;
  std::vector<double> host_nonfluid;  // AoS values of non-fluid cells (kept as given)

  uint8_t* d_pt = nullptr;
  double* d_U[3] = {nullptr, nullptr, nullptr};
  double* d_D = nullptr;
  double* d_FX = nullptr;
  double* d_FY = nullptr;
  double* d_RHS = nullptr;
  double* d_aos = nullptr;   // staging for AoS transfers (nx*ny*6)
  double* d_partial = nullptr;
  int n_partial = 0;
  Scalars* d_sc = nullptr;
  long* d_gpad = nullptr;    // all ghosts, sorted-id order
  double* d_gvals = nullptr;

  CheapGhosts cheap{};
  HeavyGhosts heavy{};
  std::vector<void*> allocs;

  hipGraphExec_t graph = nullptr;
  bool profiling = false;
  std::vector<ProfRec> prof;
  double prof_ms[K_COUNT] = {0};
  long long prof_n[K_COUNT] = {0};

  rh3t_error last{};
};

namespace {

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      std::snprintf(ctx->last.message, sizeof(ctx->last.message), "%s: %s", #expr,      \
                    hipGetErrorString(e_));                                             \
      ctx->last.kind = RH3T_ERR_HIP;                                                    \
      return RH3T_ERR_HIP;                                                              \
    }                                                                                   \
  } while (0)

template <class T>
int dev_alloc(rh3t_ctx* ctx, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) {
    std::snprintf(ctx->last.message, sizeof(ctx->last.message), "hipMalloc(%zu): %s", count * sizeof(T),
                  hipGetErrorString(e));
    return RH3T_ERR_HIP;
  }
  ctx->allocs.push_back(q);
  *p = (T*)q;
  return RH3T_OK;
}

template <class T>
int upload(rh3t_ctx* ctx, T** p, const std::vector<T>& v) {
  int st = dev_alloc(ctx, p, v.size());
  if (st) return st;
  if (!v.empty()) HIP_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return RH3T_OK;
}

// Launch helper with optional per-kernel event timing.
struct Launcher {
  rh3t_ctx* ctx;
  int kid;
  hipEvent_t a = nullptr, b = nullptr;
  Launcher(rh3t_ctx* c, int k) : ctx(c), kid(k) {
    if (ctx->profiling) {
      hipEventCreate(&a);
      hipEventCreate(&b);
      hipEventRecord(a, ctx->stream);
    }
  }
  ~Launcher() {
    if (ctx->profiling) {
      hipEventRecord(b, ctx->stream);
      ctx->prof.push_back(ProfRec{kid, a, b});
    }
  }
};

inline int flux_tiles_j(const rh3t_ctx* ctx, int nj) { return (nj + kFluxBlockJ - 1) / kFluxBlockJ; }

int launch_ghosts(rh3t_ctx* ctx, double* S, unsigned phase) {
  if (ctx->cheap.n > 0) {
    Launcher L(ctx, K_GHOST_CHEAP);
    k_ghost_cheap<<<(ctx->cheap.n + 255) / 256, 256, 0, ctx->stream>>>(S, ctx->geo, ctx->cheap, ctx->d_sc, phase);
  }
  if (ctx->heavy.n > 0) {
    Launcher L(ctx, K_GHOST_HEAVY);
    k_ghost_heavy<<<ctx->heavy.n, 64, 0, ctx->stream>>>(S, ctx->geo, ctx->heavy, ctx->phys, ctx->d_sc, phase);
  }
  if (ctx->n_ghost > 0) {
    Launcher L(ctx, K_CAPTURE);
    k_capture<<<1, 1, 0, ctx->stream>>>(S, ctx->geo, ctx->d_sc, phase, ctx->d_gpad);
  }
  return RH3T_OK;
}

// ghosts -> derived -> x/y fluxes -> assemble (one l() + stage)
int launch_stage(rh3t_ctx* ctx, double* S, const StageArgs& sa) {
  const Geo& g = ctx->geo;
  launch_ghosts(ctx, S, sa.mode == 0 ? 2u : (unsigned)(2 * sa.mode));
  {
    Launcher L(ctx, K_DERIVED);
    long total = g.plane;
    k_derived<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(S, ctx->d_D, ctx->d_pt, total, g,
                                                                        ctx->phys, ctx->nplanes);
  }
  dim3 blk(kFluxBlockJ, kFluxBlockI);
  {
    Launcher L(ctx, K_FLUX_X);
    int tj = flux_tiles_j(ctx, g.ny), ti = (g.nx + 1 + kFluxBlockI - 1) / kFluxBlockI;
    if (ctx->fast)
      k_flux<0, true><<<tj * ti, blk, 0, ctx->stream>>>(S, ctx->d_D, ctx->d_FX, ctx->d_pt, g, ctx->eig,
                                                        ctx->phys.eps, tj);
    else
      k_flux<0, false><<<tj * ti, blk, 0, ctx->stream>>>(S, ctx->d_D, ctx->d_FX, ctx->d_pt, g, ctx->eig,
                                                         ctx->phys.eps, tj);
  }
  {
    Launcher L(ctx, K_FLUX_Y);
    int tj = flux_tiles_j(ctx, g.ny + 1), ti = (g.nx + kFluxBlockI - 1) / kFluxBlockI;
    if (ctx->fast)
      k_flux<1, true><<<tj * ti, blk, 0, ctx->stream>>>(S, ctx->d_D, ctx->d_FY, ctx->d_pt, g, ctx->eig,
                                                        ctx->phys.eps, tj);
    else
      k_flux<1, false><<<tj * ti, blk, 0, ctx->stream>>>(S, ctx->d_D, ctx->d_FY, ctx->d_pt, g, ctx->eig,
                                                         ctx->phys.eps, tj);
  }
  {
    Launcher L(ctx, K_ASSEMBLE);
    int tj = flux_tiles_j(ctx, g.ny), ti = (g.nx + kFluxBlockI - 1) / kFluxBlockI;
    k_assemble<<<tj * ti, blk, 0, ctx->stream>>>(ctx->d_D, ctx->d_FX, ctx->d_FY, ctx->d_pt, g, ctx->phys, sa,
                                                 ctx->d_sc, tj);
  }
  if (sa.mode > 0) {
    Launcher L(ctx, K_CAPTURE);
    k_capture<<<1, 1, 0, ctx->stream>>>(sa.dst, g, ctx->d_sc, sa.phase, ctx->d_gpad);
  }
  return RH3T_OK;
}

int launch_dt(rh3t_ctx* ctx, int advance) {
  const Geo& g = ctx->geo;
  {
    Launcher L(ctx, K_DT_PARTIAL);
    k_dt_partial<<<ctx->n_partial, 256, 0, ctx->stream>>>(ctx->d_U[0], ctx->d_pt, g, ctx->phys, ctx->d_partial);
  }
  {
    Launcher L(ctx, K_DT_FINAL);
    k_dt_final<<<1, 256, 0, ctx->stream>>>(ctx->d_partial, ctx->n_partial, ctx->d_sc, advance);
  }
  return RH3T_OK;
}

// rk3_ssp, src/main.rs:351-373, with the state kept in U0:
//   stage 1: U1 = U0 + dt L(U0)
//   stage 2: U2 = 3/4 U0 + 1/4 U1 + dt/4 L(U1)
//   stage 3: U0 = 1/3 U0 + 2/3 U2 + 2dt/3 L(U2)   (in place: each cell
//            reads its own U0 value before writing it)
int launch_rk3(rh3t_ctx* ctx) {
  StageArgs s1{1, ctx->d_U[0], nullptr, ctx->d_U[1], nullptr, 3};
  launch_stage(ctx, ctx->d_U[0], s1);
  StageArgs s2{2, ctx->d_U[0], ctx->d_U[1], ctx->d_U[2], nullptr, 5};
  launch_stage(ctx, ctx->d_U[1], s2);
  StageArgs s3{3, ctx->d_U[0], ctx->d_U[2], ctx->d_U[0], nullptr, 7};
  launch_stage(ctx, ctx->d_U[2], s3);
  return RH3T_OK;
}

int check_error(rh3t_ctx* ctx) {
  Scalars sc;
  HIP_TRY(hipMemcpyAsync(&sc, ctx->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (sc.err == ~0ULL) return RH3T_OK;
  unsigned phase = (unsigned)((sc.err >> 40) & 0xF);
  unsigned kind = (unsigned)((sc.err >> 32) & 0xFF);
  unsigned idx = (unsigned)(sc.err & 0xFFFFFFFFULL);
  rh3t_error& e = ctx->last;
  std::memset(&e, 0, sizeof(e));
  e.kind = (int)kind;
  e.stage = phase >= 2 ? (int)((phase - 2) / 2 + 1) : 0;
  e.ghost_id = -1;
  std::memcpy(e.state, sc.err_state, sizeof(e.state));
  if (kind == RH3T_ERR_DT) {
    std::snprintf(e.message, sizeof(e.message), "non-positive dt at step %u", (unsigned)(sc.err >> 44));
  } else if (phase % 2 == 0) {
    e.ghost_id = (int)idx;
    std::snprintf(e.message, sizeof(e.message), "NON-FINITE GHOST STATE ghost id %u", idx);
  } else {
    e.i = (int)(idx / ctx->geo.ny);
    e.j = (int)(idx % ctx->geo.ny);
    std::snprintf(e.message, sizeof(e.message), "non-physical state at (%d, %d) in RK%d state", e.i, e.j,
                  e.stage);
  }
  return (int)kind;
}

int reset_error(rh3t_ctx* ctx) {
  k_set_err_clean<<<1, 1, 0, ctx->stream>>>(ctx->d_sc);
  HIP_TRY(hipGetLastError());
  return RH3T_OK;
}

Phys make_phys(const rh3t_physics& p) {
  Phys ph;
  ph.gamma_e = p.gamma_e;
  ph.gamma_i = p.gamma_i;
  ph.gamma_r = p.gamma_r;
  ph.gi = p.gamma_i - 1.0;
  ph.ge = p.gamma_e - 1.0;
  ph.gr = p.gamma_r - 1.0;
  ph.gt = ph.gi + ph.ge + ph.gr;
  ph.kappa_e = p.kappa_e;
  ph.kappa_i = p.kappa_i;
  ph.kappa_r = p.kappa_r;
  ph.cve = p.cve;
  ph.cvi = p.cvi;
  ph.a_rad = p.a_rad;
  ph.omega_ei = p.omega_ei;
  ph.omega_er = p.omega_er;
  ph.lambda = p.lambda;
  ph.weno_q = p.weno_q;
  ph.eps = p.eps;
  ph.diffusion = (p.kappa_e != 0.0 || p.kappa_i != 0.0 || p.kappa_r != 0.0) ? 1 : 0;
  ph.source = (p.omega_ei != 0.0 || p.omega_er != 0.0) ? 1 : 0;
  return ph;
}

// Uniform prefixes of the eigensystem expressions, evaluated with the
// reference's association (src/weno.rs:56-84): identical roundings to the
// per-interface evaluation.
EigConst make_eig(const Phys& ph) {
  EigConst K;
  const double g[3] = {ph.ge, ph.gi, ph.gr};
  const double gam[3] = {ph.gamma_e, ph.gamma_i, ph.gamma_r};
  const double gt = ph.gt;
  for (int k = 0; k < 3; ++k) {
    K.g[k] = g[k];
    K.gam[k] = gam[k];
    K.gamg[k] = gam[k] * g[k];
    K.g_over_gt[k] = g[k] / gt;
    K.g_gt[k] = g[k] * gt;
    K.two_g_gt[k] = 2.0 * g[k] * gt;
    K.six_g[k] = 6.0 * g[k];
    K.gam_gt[k] = gam[k] * gt;
    K.inv6g[k] = 1.0 / (6.0 * g[k]);
    for (int j = 0; j < 3; ++j) {
      K.num[k][j] = j == k ? -(g[k] * g[k]) : -g[k] * g[j];
      K.coef[k][j] = j == k ? 6.0 * g[k] : 6.0 * g[k] * g[j];
    }
  }
  K.h6[0] = 6.0 * ph.gi * ph.gr;
  K.h6[1] = 6.0 * ph.ge * ph.gr;
  K.h6[2] = 6.0 * ph.ge * ph.gi;
  K.c36 = 36.0 * ph.ge * ph.gi * ph.gr;
  K.m2gt = -2.0 * gt;
  K.gt = gt;
  K.inv_gt = 1.0 / gt;
  return K;
}

void free_ctx(rh3t_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->graph) hipGraphExecDestroy(ctx->graph);
  for (auto& r : ctx->prof) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (void* p : ctx->allocs) hipFree(p);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

int build_ctx(rh3t_ctx* ctx, const rh3t::Problem& P) {
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  Geo& g = ctx->geo;
  g.nx = P.grid.nx;
  g.ny = P.grid.ny;
  g.PY = ((g.ny + 8 + 15) / 16) * 16;
  g.plane = (long)(g.nx + 8) * g.PY;
  g.dx = P.grid.dx;
  g.dy = P.grid.dy;
  ctx->nplanes = (ctx->phys.diffusion || ctx->phys.source) ? D_N : 8;
  const long ncell = (long)g.nx * g.ny;
  ctx->host_fluid = P.fluid;

  // point types: 1 fluid, 2 ghost
  std::vector<uint8_t> pt(g.plane, 0);
  for (long i = 0; i < g.nx; ++i)
    for (long j = 0; j < g.ny; ++j)
      if (P.fluid[i * g.ny + j]) pt[pad_index(g, i, j)] = 1;
  ctx->n_ghost = (long)P.ghosts.size();
  std::vector<long> gpad(P.ghosts.size());
  for (size_t k = 0; k < P.ghosts.size(); ++k) {
    const auto& gh = P.ghosts[k];
    if (gh.i < -4 || gh.i >= g.nx + 4 || gh.j < -4 || gh.j >= g.ny + 4) {
      std::snprintf(ctx->last.message, sizeof(ctx->last.message), "ghost outside the halo");
      return RH3T_ERR_GEOMETRY;
    }
    gpad[k] = pad_index(g, gh.i, gh.j);
    pt[gpad[k]] = 2;
    if (gh.bc_kind == RH3T_BC_TIME_DEPENDENT) {
      std::snprintf(ctx->last.message, sizeof(ctx->last.message),
                    "TimeDependent boundary conditions are not part of this path");
      return RH3T_ERR_UNSUPPORTED;
    }
  }
  int st;
  if ((st = upload(ctx, &ctx->d_pt, pt))) return st;
  if ((st = upload(ctx, &ctx->d_gpad, gpad))) return st;
  for (int b = 0; b < 3; ++b) {
    if ((st = dev_alloc(ctx, &ctx->d_U[b], 6 * (size_t)g.plane))) return st;
    HIP_TRY(hipMemset(ctx->d_U[b], 0, 6 * (size_t)g.plane * sizeof(double)));
  }
  if ((st = dev_alloc(ctx, &ctx->d_D, (size_t)ctx->nplanes * g.plane))) return st;
  HIP_TRY(hipMemset(ctx->d_D, 0, (size_t)ctx->nplanes * g.plane * sizeof(double)));
  if ((st = dev_alloc(ctx, &ctx->d_FX, 6 * (size_t)(g.nx + 1) * g.ny))) return st;
  if ((st = dev_alloc(ctx, &ctx->d_FY, 6 * (size_t)g.nx * (g.ny + 1)))) return st;
  if ((st = dev_alloc(ctx, &ctx->d_RHS, 6 * (size_t)ncell))) return st;
  if ((st = dev_alloc(ctx, &ctx->d_aos, 6 * (size_t)ncell))) return st;
  ctx->n_partial = (int)((ncell + 255) / 256);
  if ((st = dev_alloc(ctx, &ctx->d_partial, (size_t)ctx->n_partial))) return st;
  if ((st = dev_alloc(ctx, &ctx->d_gvals, 6 * (size_t)std::max<long>(ctx->n_ghost, 1)))) return st;
  if ((st = dev_alloc(ctx, &ctx->d_sc, 1))) return st;
  {
    Scalars sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.err = ~0ULL;
    sc.cfl_factor = 0.05;
    HIP_TRY(hipMemcpy(ctx->d_sc, &sc, sizeof(sc), hipMemcpyHostToDevice));
  }

  // ghost tables
  std::vector<long> c_pad, c_src, h_pad, h_near, h_st;
  std::vector<int> c_kind, c_gid, h_kind, h_gid;
  std::vector<double> c_n, c_state, h_n, h_dist, h_state, h_of, h_d5, h_S, h_W;
  for (size_t k = 0; k < P.ghosts.size(); ++k) {
    const auto& gh = P.ghosts[k];
    const auto& bc = P.bc_list(gh.boundary)[gh.side_id];
    if (gh.heavy_id < 0) {
      c_pad.push_back(gpad[k]);
      c_src.push_back(gh.src_i >= 0 ? pad_index(g, gh.src_i, gh.src_j) : 0);
      c_kind.push_back(gh.bc_kind);
      c_gid.push_back((int)k);
      c_n.push_back(gh.project.normal.x);
      c_n.push_back(gh.project.normal.y);
      for (int f = 0; f < 6; ++f) c_state.push_back(bc.state[f]);
    } else {
      const auto& hb = P.heavy[gh.heavy_id];
      h_pad.push_back(gpad[k]);
      h_kind.push_back(gh.bc_kind);
      h_gid.push_back((int)k);
      h_near.push_back(pad_index(g, hb.nearest_i, hb.nearest_j));
      h_n.push_back(gh.project.normal.x);
      h_n.push_back(gh.project.normal.y);
      h_dist.push_back(gh.project.distance);
      for (int f = 0; f < 6; ++f) h_state.push_back(bc.state[f]);
      h_of.push_back(bc.p_inf);
      h_of.push_back(bc.sigma);
      h_of.push_back(bc.l_domain);
      for (int q = 0; q < 5; ++q) h_d5.push_back(hb.d[q]);
      for (int q = 0; q < 55; ++q) {
        uint32_t l = hb.stencil[q];
        h_st.push_back(pad_index(g, l / g.ny, l % g.ny));
      }
      h_S.insert(h_S.end(), hb.s, hb.s + 979);
      h_W.insert(h_W.end(), hb.w, hb.w + 225);
    }
  }
  ctx->cheap.n = (int)c_pad.size();
  ctx->heavy.n = (int)h_pad.size();
  long* lp;
  int* ip;
  double* dp;
  if ((st = upload(ctx, &lp, c_pad))) return st;
  ctx->cheap.pad = lp;
  if ((st = upload(ctx, &lp, c_src))) return st;
  ctx->cheap.src = lp;
  if ((st = upload(ctx, &ip, c_kind))) return st;
  ctx->cheap.kind = ip;
  if ((st = upload(ctx, &ip, c_gid))) return st;
  ctx->cheap.gid = ip;
  if ((st = upload(ctx, &dp, c_n))) return st;
  ctx->cheap.normal = dp;
  if ((st = upload(ctx, &dp, c_state))) return st;
  ctx->cheap.cstate = dp;
  if ((st = upload(ctx, &lp, h_pad))) return st;
  ctx->heavy.pad = lp;
  if ((st = upload(ctx, &ip, h_kind))) return st;
  ctx->heavy.kind = ip;
  if ((st = upload(ctx, &ip, h_gid))) return st;
  ctx->heavy.gid = ip;
  if ((st = upload(ctx, &lp, h_near))) return st;
  ctx->heavy.nearest = lp;
  if ((st = upload(ctx, &dp, h_n))) return st;
  ctx->heavy.normal = dp;
  if ((st = upload(ctx, &dp, h_dist))) return st;
  ctx->heavy.dist = dp;
  if ((st = upload(ctx, &dp, h_state))) return st;
  ctx->heavy.bcstate = dp;
  if ((st = upload(ctx, &dp, h_of))) return st;
  ctx->heavy.outflow = dp;
  if ((st = upload(ctx, &dp, h_d5))) return st;
  ctx->heavy.d5 = dp;
  if ((st = upload(ctx, &lp, h_st))) return st;
  ctx->heavy.stencil = lp;
  if ((st = upload(ctx, &dp, h_S))) return st;
  ctx->heavy.S = dp;
  if ((st = upload(ctx, &dp, h_W))) return st;
  ctx->heavy.W = dp;

  ctx->host_nonfluid.assign(6 * (size_t)ncell, 0.0);
  return RH3T_OK;
}

}  // namespace

// ======================================================================
// C ABI
// ======================================================================

extern "C" {

int rh3t_device_count(int32_t* n) {
  if (!n) return RH3T_ERR_ARG;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return RH3T_OK;
}

int rh3t_ctx_create(const rh3t_problem* pb, const rh3t_physics* phys, int32_t device, rh3t_ctx** out) {
  if (!pb || !out || !pb->p.ghosts_built) return RH3T_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RH3T_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return RH3T_ERR_ARG;
  rh3t_physics p;
  if (phys)
    p = *phys;
  else
    rh3t_physics_default(&p);
  auto* ctx = new rh3t_ctx();
  ctx->device = device;
  ctx->phys = make_phys(p);
  ctx->eig = make_eig(ctx->phys);
  int st = build_ctx(ctx, pb->p);
  if (st != RH3T_OK) {
    free_ctx(ctx);
    return st;
  }
  *out = ctx;
  return RH3T_OK;
}

void rh3t_ctx_destroy(rh3t_ctx* ctx) { free_ctx(ctx); }

int rh3t_set_state(rh3t_ctx* ctx, const double* u, double time) {
  if (!ctx || !u) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  const Geo& g = ctx->geo;
  const long ncell = (long)g.nx * g.ny;
  for (long l = 0; l < ncell; ++l)
    if (!ctx->host_fluid[l])
      for (int f = 0; f < 6; ++f) ctx->host_nonfluid[6 * l + f] = u[6 * l + f];
  HIP_TRY(hipMemcpyAsync(ctx->d_aos, u, 6 * ncell * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  k_aos_to_pad<<<(unsigned)((ncell + 255) / 256), 256, 0, ctx->stream>>>(ctx->d_aos, ctx->d_U[0], ctx->d_pt, g);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&ctx->d_sc->time, &time, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return RH3T_OK;
}

int rh3t_get_state(rh3t_ctx* ctx, double* u, double* time) {
  if (!ctx || !u) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  const Geo& g = ctx->geo;
  const long ncell = (long)g.nx * g.ny;
  k_pad_to_aos<<<(unsigned)((ncell + 255) / 256), 256, 0, ctx->stream>>>(ctx->d_U[0], ctx->d_aos, ctx->d_pt, g);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(u, ctx->d_aos, 6 * ncell * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  double t = 0.0;
  HIP_TRY(hipMemcpyAsync(&t, &ctx->d_sc->time, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (long l = 0; l < ncell; ++l)
    if (!ctx->host_fluid[l])
      for (int f = 0; f < 6; ++f) u[6 * l + f] = ctx->host_nonfluid[6 * l + f];
  if (time) *time = t;
  return RH3T_OK;
}

int rh3t_l(rh3t_ctx* ctx, double* rhs_aos, double* ghost_vals) {
  if (!ctx || !rhs_aos) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  int st = reset_error(ctx);
  if (st) return st;
  StageArgs sa{0, nullptr, nullptr, nullptr, ctx->d_RHS, 2};
  launch_stage(ctx, ctx->d_U[0], sa);
  HIP_TRY(hipGetLastError());
  const long ncell = (long)ctx->geo.nx * ctx->geo.ny;
  k_soa_to_aos<<<(unsigned)((ncell + 255) / 256), 256, 0, ctx->stream>>>(ctx->d_RHS, ctx->d_aos, ncell);
  HIP_TRY(hipMemcpyAsync(rhs_aos, ctx->d_aos, 6 * ncell * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (ghost_vals && ctx->n_ghost > 0) {
    k_gather_ghosts<<<(unsigned)((ctx->n_ghost + 255) / 256), 256, 0, ctx->stream>>>(
        ctx->d_U[0], ctx->d_gpad, (int)ctx->n_ghost, ctx->geo, ctx->d_gvals);
    HIP_TRY(hipMemcpyAsync(ghost_vals, ctx->d_gvals, 6 * ctx->n_ghost * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
  }
  return check_error(ctx);
}

int rh3t_rk3_step(rh3t_ctx* ctx, double dt) {
  if (!ctx) return RH3T_ERR_ARG;
  if (!(dt > 0.0)) return RH3T_ERR_DT;
  HIP_TRY(hipSetDevice(ctx->device));
  int st = reset_error(ctx);
  if (st) return st;
  HIP_TRY(hipMemcpyAsync(&ctx->d_sc->dt, &dt, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  launch_rk3(ctx);
  HIP_TRY(hipGetLastError());
  // Field.time advances by dt (u3.time = u.time + dt, src/main.rs:370)
  double t = 0.0;
  HIP_TRY(hipMemcpyAsync(&t, &ctx->d_sc->time, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  t = t + dt;
  HIP_TRY(hipMemcpyAsync(&ctx->d_sc->time, &t, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  return check_error(ctx);
}

int rh3t_calc_global_dt(rh3t_ctx* ctx, double* dt_cfl) {
  if (!ctx || !dt_cfl) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  launch_dt(ctx, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(dt_cfl, &ctx->d_sc->dt_cfl, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return RH3T_OK;
}

int rh3t_advance(rh3t_ctx* ctx, int32_t nsteps, double cfl_factor, double dt_fixed, double* time) {
  if (!ctx || nsteps < 0) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  int st = reset_error(ctx);
  if (st) return st;
  HIP_TRY(hipMemcpyAsync(&ctx->d_sc->cfl_factor, &cfl_factor, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(&ctx->d_sc->dt_fixed, &dt_fixed, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  if (ctx->profiling) {
    for (int s = 0; s < nsteps; ++s) {
      launch_dt(ctx, 1);
      launch_rk3(ctx);
    }
    HIP_TRY(hipGetLastError());
  } else {
    if (!ctx->graph) {
      hipGraph_t gr;
      HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      launch_dt(ctx, 1);
      launch_rk3(ctx);
      hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
      if (e != hipSuccess) {
        std::snprintf(ctx->last.message, sizeof(ctx->last.message), "graph capture: %s", hipGetErrorString(e));
        return RH3T_ERR_HIP;
      }
      e = hipGraphInstantiate(&ctx->graph, gr, nullptr, nullptr, 0);
      hipGraphDestroy(gr);
      if (e != hipSuccess) {
        ctx->graph = nullptr;
        std::snprintf(ctx->last.message, sizeof(ctx->last.message), "graph instantiate: %s",
                      hipGetErrorString(e));
        return RH3T_ERR_HIP;
      }
    }
    for (int s = 0; s < nsteps; ++s) HIP_TRY(hipGraphLaunch(ctx->graph, ctx->stream));
  }
  if (time) HIP_TRY(hipMemcpyAsync(time, &ctx->d_sc->time, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  return check_error(ctx);
}

int rh3t_ctx_set_arith(rh3t_ctx* ctx, int32_t mode) {
  if (!ctx || (mode != RH3T_ARITH_EXACT && mode != RH3T_ARITH_FAST)) return RH3T_ERR_ARG;
  ctx->fast = mode == RH3T_ARITH_FAST;
  return RH3T_OK;
}

int rh3t_last_error(const rh3t_ctx* ctx, rh3t_error* err) {
  if (!ctx || !err) return RH3T_ERR_ARG;
  *err = ctx->last;
  return RH3T_OK;
}

int rh3t_profile(rh3t_ctx* ctx, int32_t enable) {
  if (!ctx) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->prof) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  ctx->prof.clear();
  for (int k = 0; k < K_COUNT; ++k) {
    ctx->prof_ms[k] = 0;
    ctx->prof_n[k] = 0;
  }
  ctx->profiling = enable != 0;
  return RH3T_OK;
}

int rh3t_profile_read(rh3t_ctx* ctx, int32_t max_kernels, char* names, int32_t names_cap, double* avg_ms,
                      int64_t* launches, int32_t* n_kernels) {
  if (!ctx || !n_kernels) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->prof) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
    ctx->prof_ms[r.kernel] += ms;
    ctx->prof_n[r.kernel] += 1;
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  ctx->prof.clear();
  int n = std::min<int>(K_COUNT, max_kernels);
  std::string s;
  for (int k = 0; k < n; ++k) {
    s += kKernelNames[k];
    s.push_back('\0');
    if (avg_ms) avg_ms[k] = ctx->prof_n[k] ? ctx->prof_ms[k] / (double)ctx->prof_n[k] : 0.0;
    if (launches) launches[k] = ctx->prof_n[k];
  }
  if (names && names_cap > 0) {
    int m = std::min<int>((int)s.size(), names_cap);
    std::memcpy(names, s.data(), m);
  }
  *n_kernels = n;
  return RH3T_OK;
}

}  // extern "C"
