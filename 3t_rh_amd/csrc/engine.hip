// HIP (gfx950) implementation of l() and rk3_ssp() of colinhu123/3T_RH over
// a ragged batch of independent solves, and the device half of the C ABI in
// include/rh3t.h.  A plain solve (rh3t_ctx) is a batch of one.
//
// Per l() evaluation of the stage state S of every solve
// (src/main.rs:135-305):
//   k_ghost_cheap   Periodic / Reflective / Zeroth / Constant ghosts
//                   (src/bc1.rs:471-636, 930-1037): gather from a static
//                   source cell into the ghost's padded slot (it rides in
//                   k_ghost_prep's launch when a problem has heavy ghosts).
//   k_ghost_prep    the heavy ghosts' frames (the Roe eigensystem of each
//                   distinct nearest point and normal, src/bc1.rs:294-312)
//                   and their E_r points packed 6 doubles per point.
//   k_ghost_open /  FarField / Outflow and Wall ghosts (src/bc1.rs:609-928,
//   k_ghost_wall    193-255): two ghosts per wavefront; the E_r projection
//                   and the S_r / w_r rows spread over a 32-lane group,
//                   sums in the reference's order; the BC branch (Wall: the
//                   two ILW systems) and the Taylor sum.
//   k_derived       Derived::from_state on every fluid cell and ghost
//                   (src/state.rs:241-275, src/main.rs:149-153).
//   k_flux<0>       characteristic WENO5 interface flux in x
//                   (src/main.rs:165-204 -> src/weno.rs:685-802), one
//                   interface per lane (flux.hpp).
//   k_fy_assemble   the y interface fluxes (as k_flux<1>) with flux
//                   differences + noncon x/y + source + diffusion
//                   (src/main.rs:251-304), the RK stage combination and the
//                   admissibility check (src/main.rs:307-349) in one pass
//                   (k_flux<1> + k_assemble for launch shapes it does not
//                   take).
//   k_capture       the state of a call's first failure (in the time loop
//                   k_dt_final takes it for the previous step).
//   k_dt_final      calc_global_dt (src/main.rs:375-408, src/dt.rs:5-34) as
//                   wavefront fmin reductions: the last RK stage's assembly
//                   leaves one minimum of get_local_dt of the new state per
//                   workgroup (k_dt_partial does the same from U0 when the
//                   state came from the host), one workgroup per solve
//                   reduces them and sets dt = cfl_factor * dt_cfl
//                   (src/main.rs:461), on the device.
// Every grid-shaped kernel enumerates the tiles of all solves in one launch:
// a per-kernel prefix table maps a (remapped) workgroup index to its solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "engine.hpp"
#include "problem.hpp"

using namespace rh3t_dev;
using namespace rh3t_eng;

namespace {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; give
// each XCD a contiguous run of tiles so that halo rows shared by
// neighbouring tiles stay in one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int full = (nb / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

// start[0] = 0 <= ... <= start[n]: the solve owning tile t
__device__ __forceinline__ int find_solve(const int* __restrict__ start, int n, int t) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (start[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void raise_error(Scalars* sc, unsigned phase, unsigned kind, unsigned idx) {
  atomicMin(&sc->err, err_key(sc->step, phase, kind, idx));
}

// ------------------------------------------------------------ ghosts

// Derived::from_state of the state s written to slot p of solve sd: its
// first wd planes into sd.D (the planes of the buffer the slot belongs to)
__device__ __forceinline__ void put_derived(const SolveDesc& sd, const Phys& ph, const double s[6], long p, int wd) {
  double d[D_N];
  derived_from_state(ph, s, d, wd > 8);
  double* __restrict__ Dp = sd.D;  // per lane in the ghost kernels: no scalar base
#pragma unroll
  for (int f = 0; f < D_N; ++f)
    if (f < wd) Dp[f * sd.g.plane + p] = d[f];
}

// wd > 0: also Derived of the ghost's new state, its first wd planes
__device__ __forceinline__ void ghost_cheap(const SolveDesc* __restrict__ descs, const CheapGhosts& cg, int buf,
                                            unsigned phase, int gt, int k, const Phys& ph, int wd) {
  const SolveDesc& sd = descs[cg.sid[k]];
  double* S = sd.U[buf];
  const long plane = sd.g.plane;
  int kind = cg.kind[k];
  long dst = cg.pad[k];
  double v[6];
  if (kind == RH3T_BC_CONSTANT) {
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = cg.cstate[6L * k + f];
  } else if (kind == RH3T_BC_TIME_DEPENDENT) {
    // the moving front profile at the time of the state being evaluated:
    // Field.time of an l() call; t, t + dt, t + 0.5 dt for the three l()
    // calls of rk3_ssp (src/main.rs:356-366); the closure compares P0.x
    // with front0 + speed * t (src/init.rs:61-65: 10.0 * t)
    const Scalars* sc = sd.sc;
    const double t = gt == GT_FIELD ? sc->time
                                    : (gt == GT_STAGE1 ? sc->t_n
                                                       : (gt == GT_STAGE2 ? sc->t_n + sc->dt : sc->t_n + 0.5 * sc->dt));
    const double xf = cg.front[2L * k + 1] + cg.front[2L * k] * t;
    const double* st = cg.p0x[k] <= xf ? cg.cstate + 6L * k : cg.cstate2 + 6L * k;
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = st[f];
  } else {
    long s = cg.src[k];
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = S[f * plane + s];
    if (kind == RH3T_BC_REFLECTIVE_WALL) {  // src/bc1.rs:1015-1036
      double nx = cg.normal[2L * k], ny = cg.normal[2L * k + 1];
      double mom_n = v[1] * nx + v[2] * ny;
      v[1] = v[1] - 2.0 * mom_n * nx;
      v[2] = v[2] - 2.0 * mom_n * ny;
    }
  }
  bool ok = true;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    S[f * plane + dst] = v[f];
    ok = ok && isfinite(v[f]);
  }
  if (!ok) raise_error(sd.sc, phase, RH3T_ERR_GHOST, (unsigned)cg.gid[k]);
  if (wd > 0) put_derived(sd, ph, v, dst, wd);
}

__global__ __launch_bounds__(256) void k_ghost_cheap(const SolveDesc* __restrict__ descs, CheapGhosts cg, int buf,
                                                     unsigned phase, int gt, Phys ph, int wd) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < cg.n) ghost_cheap(descs, cg, buf, phase, gt, k, ph, wd);
}

// Derived::from_state of the heavy ghost slots of buffer buf, their first
// wd planes, in (solve, slot) order so that a wavefront's lanes touch
// neighbouring slots of one solve: used instead of forming them in the
// ghost kernels when the heavy ghosts are many (see ghosts_on).
__global__ __launch_bounds__(256) void k_derived_ghosts(const SolveDesc* __restrict__ descs, HeavyGhosts hg,
                                                        Phys ph, int buf, int wd) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= hg.n) return;
  const int hk = hg.dord[k];
  const SolveDesc& sd = descs[hg.sid[hk]];
  const long p = hg.pad[hk];
  const double* __restrict__ S = sd.U[buf];
  double st[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) st[f] = S[f * sd.g.plane + p];
  put_derived(sd, ph, st, p, wd);
}

// y = R x of the X eigensystem with the entries formed on the fly (one dot6
// per row, rows in order).
// Heavy ghosts per workgroup: at most this many wavefronts (two ghosts each)
constexpr int kMaxGhostsPerBlock = 4;
// from this many heavy ghosts on, their Derived values get a pass of their
// own (ghosts_on)
constexpr int kHeavyDerivedPass = 130000;

// One frame: build_local_plain (src/bc1.rs:294-312), the nearest
// state rotated to the (n, t) frame and the X-direction eigensystem of its
// Roe average with itself -- the rows of L and R and the eigenvalues, the
// values k_ghost_open's ghosts of this frame apply.
__device__ __forceinline__ void ghost_frame(const SolveDesc* __restrict__ descs, const GhostFrames& fr,
                                            const Phys& ph, int buf, int k) {
  const SolveDesc& sd = descs[fr.sid[k]];
  const double* S = sd.U[buf];
  const long plane = sd.g.plane;
  const double nx = fr.normal[2L * k], ny = fr.normal[2L * k + 1];
  const long q = fr.nearest[k];
  double val[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) val[f] = S[f * plane + q];
  double s1[6] = {val[0], val[1] * nx + val[2] * ny, -val[1] * ny + val[2] * nx, val[3], val[4], val[5]};
  double d1[D_N];
  derived_from_state(ph, s1, d1, false);
  RoeX Q;
  roe_x(ph, s1[0], s1[0], d1[D_U], d1[D_V], d1[D_EE], d1[D_EI], d1[D_ER], d1[D_U], d1[D_V], d1[D_EE], d1[D_EI],
        d1[D_ER], Q);
  double* o = fr.out + (long)kFrameLen * k;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double row[6];
    lx_row(ph, Q, c, row);
#pragma unroll
    for (int m = 0; m < 6; ++m) o[6 * c + m] = row[m];
  }
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int m = 0; m < 6; ++m) o[36 + 6 * a + m] = rx_entry(ph, Q, a, m);
#pragma unroll
  for (int c = 0; c < 6; ++c) o[72 + c] = lamx(Q, c);
}

// One packed point: its 6 state values, contiguous.
__device__ __forceinline__ void pack_point(const SolveDesc* __restrict__ descs, const PackedPoints& pk, int buf,
                                           int e) {
  const SolveDesc& sd = descs[pk.sid[e]];
  const double* S = sd.U[buf];
  const long q = pk.slot[e];
  double v[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) v[f] = S[f * sd.g.plane + q];
  double2* o = reinterpret_cast<double2*>(pk.out + 6L * e);
  o[0] = make_double2(v[0], v[1]);
  o[1] = make_double2(v[2], v[3]);
  o[2] = make_double2(v[4], v[5]);
}

// The cheap ghosts and the FarField / Outflow ghosts' inputs in one launch:
// workgroups [0, nbf) form the frames, [nbf, nbp) pack the E_r points, the
// others take the cheap ghosts (which no frame or E_r point reads).
__global__ __launch_bounds__(64) void k_ghost_prep(const SolveDesc* __restrict__ descs, GhostFrames fr,
                                                   PackedPoints pk, CheapGhosts cg, Phys ph, int buf,
                                                   unsigned phase, int gt, int nbf, int nbp, int wd) {
  const int b = blockIdx.x;
  if (b < nbf) {
    const int k = b * 64 + threadIdx.x;
    if (k < fr.n) ghost_frame(descs, fr, ph, buf, k);
  } else if (b < nbp) {
    const int e = (b - nbf) * 64 + threadIdx.x;
    if (e < pk.n) pack_point(descs, pk, buf, e);
  } else {
    const int k = (b - nbp) * 64 + threadIdx.x;
    if (k < cg.n) ghost_cheap(descs, cg, buf, phase, gt, k, ph, wd);
  }
}

// FarField / Outflow ghosts, kOpenPerWave of them per 64-lane wavefront:
// lane group g = lane / LG (LG = 64 / kOpenPerWave lanes) takes one ghost.
// The eigensystem of the ghost's frame comes from k_ghost_prep (formed
// once per nearest point and normal, about one in four ghosts); what else is
// uniform over a ghost -- the Outflow relaxation, the Taylor sum -- is
// evaluated once per lane group, so the wavefront's instructions serve
// kOpenPerWave ghosts at a time.  The 70 table rows of a
// ghost (55 rows of S_r, the 15 rows of w_r) run on the group's 32 lanes in
// two segments of 25 and 16 steps (41 instead of the 75 that three rounds
// of 25-entry rows would take), from the row block swizzled on the host so
// that each step's loads are coalesced.  Every sum keeps the order of
// weno_extrapolation_pre and the BC branches (src/bc1.rs:193-255,
// 760-928), so the values are the same as one ghost per wavefront.
constexpr int kOpenPerWave = 2;
constexpr int kOpenLG = 64 / kOpenPerWave;
// data rows of 57 doubles (55 E_r points, u_inf, one pad): a stride of 56
// put rows c and c + 2 on the same LDS banks
constexpr int kRow = 57;
// per-ghost scalars read by the later phases (see ghost_grp)
constexpr int kScal = 14;
constexpr int kOpenLds = 6 * kRow + 90 + kFrameLen + kScal;  // doubles per ghost
// Wall ghosts keep the L rows for their ILW systems: alpha gets its own 30
constexpr int grp_lds(bool wall) { return kOpenLds + (wall ? 30 : 0); }
static_assert(kOpenLG == 32, "the swizzled table schedule is for 32-lane groups");
// slot of v_k,r (k <= r; v_k,r = 0 for k > r is never stored)
__device__ __forceinline__ int vk_slot(int k, int r) { return r * (r + 1) / 2 + k; }

// wd > 0: also Derived of each new ghost value, its first wd planes
template <bool WALL>
__device__ __forceinline__ void ghost_grp(const SolveDesc* __restrict__ descs, const HeavyGhosts& hg,
                                          const Phys& ph, int buf, unsigned phase, int h0, int count, int wd) {
  constexpr int LG = kOpenLG;
  const int lane = threadIdx.x;
  const int grp = lane / LG, l = lane % LG;
  const int w = threadIdx.y;
  const int hl = (xcd_tile(blockIdx.x, gridDim.x) * blockDim.y + w) * kOpenPerWave + grp;
  // per-ghost LDS (kOpenLds doubles): data[6][kRow] (L . u at the E_r points,
  // then, once the table rows have been applied, the products data * (S_r
  // data) in the same slots; slot 55 keeps L . u_inf), vk[15][6] (v_k,r for
  // k <= r), the frame (rows of L, rows of R, eigenvalues).  Once read, the
  // L rows hold alpha[6][5], and data slots that are no longer read hold the
  // normalised weights om[6][5] and V^(k) v[5][6] (slot 55 of each row is
  // kept).  The per-ghost scalars s_sc (14 doubles): the distance D, the
  // destination slot, the Outflow constants, the weights d_r, the state
  // plane pointer, the plane size and the Derived plane pointer.  524
  // doubles per ghost.
  extern __shared__ double s_dyn[];
  double* const s_data = s_dyn + ((long)w * kOpenPerWave + grp) * grp_lds(WALL);
  double* const s_part = s_data;
  double* const s_vk = s_data + 6 * kRow;
  double* const s_F = s_vk + 90;
  double* const s_alpha = WALL ? s_F + kFrameLen : s_F;  // Wall keeps the L rows
  double* const s_sc = s_F + kFrameLen + (WALL ? 30 : 0);
  double* const s_v = s_data;
  double* const s_om = s_data + 60;
  const bool active = hl < count;
  const int h = h0 + (active ? hl : 0);
  // Every global input of the ghost is requested before the first barrier
  // in two dependent rounds, with no branch between the loads: first what
  // the ghost index gives (table row, solve, frame, kind, normal, the packed
  // indices of its E_r points, D, the destination slot, the Outflow
  // constants), then what those give (the packed E_r states -- 3 x 16 bytes
  // per point, packed by k_ghost_prep -- its frame's eigensystem, the table
  // entries of the first segment, d_r, the state plane).  Their latencies
  // overlap, and the scalars that later phases use wait in LDS instead of
  // being read from global memory once the arithmetic has started.  Loads of
  // lanes or slots that are not used read valid entries (index 54 for the
  // slots past the 55 E_r points, ghost h0 for an idle group).
  constexpr int NP = (56 + LG - 1) / LG;
  const long row = hg.tab[h];
  const int sid = hg.sid[h];
  const int frm = hg.frame[h];
  const double nx = hg.normal[2L * h], ny = hg.normal[2L * h + 1];
  const bool far = !WALL && hg.kind[h] == RH3T_BC_FARFIELD;
  int gp[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) gp[p] = hg.gpk[55L * h + min(l + p * LG, 54)];
  using u64 = unsigned long long;
  // lanes 0..4: D, the destination slot, p_inf, sigma, l_domain
  const u64* sc1p = l == 0   ? reinterpret_cast<const u64*>(hg.dist + h)
                    : l == 1 ? reinterpret_cast<const u64*>(hg.pad + h)
                             : reinterpret_cast<const u64*>(hg.outflow + 3L * h + min(max(l - 2, 0), 2));
  const u64 sc1 = *sc1p;
  // lanes 0..4: d_0 .. d_4 of the row; 5: the state plane pointer; 6: the
  // plane size; 7: the Derived plane pointer
  const SolveDesc* sdp = descs + sid;
  const u64* sc2p = l < 5    ? reinterpret_cast<const u64*>(hg.d5 + 5L * row + l)
                    : l == 5 ? reinterpret_cast<const u64*>(sdp->U + buf)
                    : l == 6 ? reinterpret_cast<const u64*>(&sdp->g.plane)
                             : reinterpret_cast<const u64*>(&sdp->D);
  const double* Fg = hg.F + (long)kFrameLen * frm;
  double fv[(kFrameLen + LG - 1) / LG];
#pragma unroll
  for (int q = 0; q < (kFrameLen + LG - 1) / LG; ++q) fv[q] = Fg[min(l + q * LG, kFrameLen - 1)];
  const u64 sc2 = *sc2p;
  const double2* T = reinterpret_cast<const double2*>(hg.T + (long)kSwzLen * row) + l;
  double tvA[kSwzA];
  // L . u at the E_r points (a < 55) and, for a FarField ghost, L . u_inf
  // (src/bc1.rs:829-924) in the free slot 55: point a = l + p LG, p = 0, 1
  {
    double u6[NP][6];
    bool has[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int a = l + p * LG;
      has[p] = active && (a < 55 || (a == 55 && far));
      const double2* pp = reinterpret_cast<const double2*>(a == 55 ? hg.bcstate + 6L * h : hg.P + 6L * gp[p]);
      const double2 x0 = pp[0], x1 = pp[1], x2 = pp[2];
      u6[p][0] = x0.x;
      u6[p][1] = x0.y * nx + x1.x * ny;
      u6[p][2] = -x0.y * ny + x1.x * nx;
      u6[p][3] = x1.y;
      u6[p][4] = x2.x;
      u6[p][5] = x2.y;
    }
#pragma unroll
    for (int q = 0; q < kSwzPairsA; ++q) {
      const double2 t2 = T[32 * q];
      tvA[2 * q] = t2.x;
      if (2 * q + 1 < kSwzA) tvA[2 * q + 1] = t2.y;
    }
#pragma unroll
    for (int q = 0; q < (kFrameLen + LG - 1) / LG; ++q)
      if (l + q * LG < kFrameLen) s_F[l + q * LG] = fv[q];
    if (l < 5) s_sc[l] = __longlong_as_double((long long)sc1);
    if (l < 8) s_sc[5 + l] = __longlong_as_double((long long)sc2);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double lr[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) lr[m] = s_F[6 * c + m];
#pragma unroll
      for (int p = 0; p < NP; ++p)
        if (has[p]) s_data[c * kRow + l + p * LG] = dot6(lr, u6[p]);
    }
  }
  __syncthreads();
  if (active) {
    // The 70 table rows on a fixed two-segment schedule (kSwz* in
    // engine.hpp): step s of the swizzled row block holds the entry each of
    // the group's 32 lanes applies at that step, two steps per 16-byte load,
    // so every load of the block is 512 consecutive bytes.  Each row sums over b in increasing order
    // from 0.0, as weno_extrapolation_pre.  An S row's product data[a] * (S
    // data)[a] goes straight into data slot a once no lane reads that E_r
    // segment any more (the group is one wavefront: LDS accesses complete in
    // program order), a w row's sum into vk.
    // segment B's table entries in flight while segment A runs
    double tv[kSwzB];
#pragma unroll
    for (int q = 0; q < kSwzPairsB; ++q) {
      const double2 t2 = T[32 * (kSwzPairsA + q)];
      tv[2 * q] = t2.x;
      tv[2 * q + 1] = t2.y;
    }
    {  // segment A: S_4 rows (lanes 0..24), w_4 rows (25..29), S_0 / w_0 (30, 31)
      const int m = l < 30 ? 25 : 1, d0 = l < 30 ? 30 : 0;
      double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int b = 0; b < kSwzA; ++b)
        if (b < m)
#pragma unroll
          for (int c = 0; c < 6; ++c) acc[c] += tvA[b] * s_data[c * kRow + d0 + b];
      if (l < 25 || l == 30) {
        const int a = l < 25 ? 30 + l : 0;
#pragma unroll
        for (int c = 0; c < 6; ++c) s_data[c * kRow + a] = s_data[c * kRow + a] * acc[c];
      } else {
        const int q = l < 30 ? vk_slot(l - 25, 4) : vk_slot(0, 0);
#pragma unroll
        for (int c = 0; c < 6; ++c) s_vk[q * 6 + c] = acc[c];
      }
    }
    {  // segment B: S_3 / w_3 rows (lanes 0..19) over 16 steps; S_2 / w_2
       // rows (20..31) over the first 9, then S_1 / w_1 rows (20..25) over 4
      const int d0 = l < 20 ? 14 : 5;
      double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int b = 0; b < 9; ++b)
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] += tv[b] * s_data[c * kRow + d0 + b];
      if (l >= 20) {
        if (l < 29) {
          const int a = 5 + (l - 20);
#pragma unroll
          for (int c = 0; c < 6; ++c) s_data[c * kRow + a] = s_data[c * kRow + a] * acc[c];
        } else {
#pragma unroll
          for (int c = 0; c < 6; ++c) s_vk[vk_slot(l - 29, 2) * 6 + c] = acc[c];
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] = 0.0;
      }
      const int m = l < 20 ? 16 : (l < 26 ? 13 : 9), d1 = l < 20 ? 14 : -8;
#pragma unroll
      for (int b = 9; b < kSwzB; ++b)
        if (b < m)
#pragma unroll
          for (int c = 0; c < 6; ++c) acc[c] += tv[b] * s_data[c * kRow + d1 + b];
      if (l < 16 || (l >= 20 && l < 24)) {
        const int a = l < 16 ? 14 + l : 1 + (l - 20);
#pragma unroll
        for (int c = 0; c < 6; ++c) s_data[c * kRow + a] = s_data[c * kRow + a] * acc[c];
      } else if (l < 26) {
        const int q = l < 20 ? vk_slot(l - 16, 3) : vk_slot(l - 24, 1);
#pragma unroll
        for (int c = 0; c < 6; ++c) s_vk[q * 6 + c] = acc[c];
      }
    }
  }
  __syncthreads();
  if (active) {  // beta_r and alpha_r = d_r / (eps + beta_r)^q, task (r, c)
    for (int t = l; t < 30; t += LG) {
      const int r = t / 6, c = t % 6;
      const int m = (r + 1) * (r + 1);
      const double* p = &s_part[c * kRow + r * (r + 1) * (2 * r + 1) / 6];
      double beta = 0.0;
      for (int a = 0; a < m; ++a) beta += p[a];
      s_alpha[c * 5 + r] = s_sc[5 + r] / weno_pow(ph.eps + beta, ph.weno_q);
    }
  }
  __syncthreads();
  if (active) {  // om_r = alpha_r / sum alpha, task (r, c)
    for (int t = l; t < 30; t += LG) {
      const int r = t / 6, c = t % 6;
      double asum = 0.0;
#pragma unroll
      for (int q = 0; q < 5; ++q) asum += s_alpha[c * 5 + q];
      s_om[c * 5 + r] = s_alpha[c * 5 + r] / asum;
    }
  }
  __syncthreads();
  if (active) {  // V^(k)_c = sum_r om_r v_k,r, task (k, c); the terms of
                 // r < k are +0.0 products added to the +0.0 start, which
                 // leave it unchanged, so the sum starts at r = k
    for (int t = l; t < 30; t += LG) {
      const int k = t / 6, c = t % 6;
      double res = 0.0;
#pragma unroll
      for (int r = 0; r < 5; ++r)
        if (r >= k) res += s_om[c * 5 + r] * s_vk[vk_slot(k, r) * 6 + c];
      s_v[k * 6 + c] = res;
    }
  }
  __syncthreads();
  const double dist = s_sc[0];
  double* const s_uk = s_vk;        // [k][i]
  double* const s_ug = s_vk + 30;   // Taylor sums
  double* const s_fix = s_vk + 42;  // Outflow: flag, rhs_1[0]; Wall: solved
  if constexpr (WALL) {
    // Wall-ILW, src/bc1.rs:691-756: u^(0) from L with row 0 replaced by the
    // zero normal momentum, u^(1) from L with row 0 the ILW row of u^(0)
    // (build_ilw_row1, src/bc1.rs:448-465), both right-hand sides V^(k)
    // with entry 0 set to 0 (lane 0 of the group, two solve6); u^(k) = R
    // V^(k) for k >= 2 on the other lanes
    if (active && l == 0) {
      double A[6][6], rhs[6], u0[6], u1[6];
#pragma unroll
      for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int m = 0; m < 6; ++m) A[c][m] = s_F[6 * c + m];
#pragma unroll
      for (int c = 0; c < 6; ++c) A[0][c] = c == 1 ? 1.0 : 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) rhs[m] = s_v[m];
      rhs[0] = 0.0;
      bool solved = solve6(A, rhs, u0);
      const double r0 = u0[0], m3 = u0[2];
      A[0][0] = -ph.gt * sq(m3) / (6.0 * sq(r0));
      A[0][1] = 0.0;
      A[0][2] = ph.gt * m3 / (6.0 * r0);
      A[0][3] = -ph.ge;
      A[0][4] = -ph.gi;
      A[0][5] = -ph.gr;
#pragma unroll
      for (int m = 0; m < 6; ++m) rhs[m] = s_v[6 + m];
      rhs[0] = 0.0;
      solved = solved && solve6(A, rhs, u1);
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        s_uk[m] = u0[m];
        s_uk[6 + m] = u1[m];
      }
      s_fix[0] = solved ? 1.0 : 0.0;
    }
    if (active) {
      for (int t = 12 + l; t < 30; t += LG) {
        const int k = t / 6, i = t % 6;
        double r[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) r[m] = s_F[36 + 6 * i + m];
        s_uk[k * 6 + i] = dot6(r, s_v + k * 6);
      }
    }
    __syncthreads();
  } else {
    // FarField, src/bc1.rs:829-924: incoming characteristics of u^(0) from
    // u_inf, of u^(k >= 1) zero.  Outflow, src/bc1.rs:760-828: the u - c
    // characteristic of u^(1) from the pressure relaxation when incoming.
    if (active && l < 6 && !far) {    // Outflow: u^(0) = R V^(0)
      double r[6], x[6];
  #pragma unroll
      for (int c = 0; c < 6; ++c) {
        r[c] = s_F[36 + 6 * l + c];
        x[c] = s_v[c];
      }
      s_uk[l] = dot6(r, x);
    }
    __syncthreads();
    if (active && l == 0 && !far) {
      const double* uk = s_uk;
      double rho0 = uk[0];
      double un0 = uk[1] / rho0, ut0 = uk[2] / rho0;
      double w2 = un0 * un0 + ut0 * ut0;  // local_pressure_and_a2, src/bc1.rs:424-440
      double ee1 = uk[3] / rho0 - w2 / 6.0;
      double ei1 = uk[4] / rho0 - w2 / 6.0;
      double er1 = uk[5] / rho0 - w2 / 6.0;
      double p0 = rho0 * (ph.ge * ee1 + ph.gi * ei1 + ph.gr * er1);
      double a2 = ph.gamma_e * ph.ge * ee1 + ph.gamma_i * ph.gi * ei1 + ph.gamma_r * ph.gr * er1;
      double a = sqrt(a2);
      double lambda1 = un0 - a;
      s_fix[0] = 0.0;
      if (lambda1 < 0.0) {
        const double p_inf = s_sc[2], sigma = s_sc[3], l_dom = s_sc[4];
        double ml = fabs(un0 / a);
        double l1 = sigma * (1.0 - sq(ml)) * a / l_dom * (p0 - p_inf);
        s_fix[0] = 1.0;
        s_fix[1] = l1 / lambda1;
      }
    }
    __syncthreads();
    if (active) {  // u^(k) = R rhs_k, task (k, i) (u^(0) of Outflow done)
      for (int t = l; t < 30; t += LG) {
        if (!far && t < 6) continue;
        const int k = t / 6, i = t % 6;
        double rhs[6], r[6];
  #pragma unroll
        for (int m = 0; m < 6; ++m) rhs[m] = s_v[k * 6 + m];
        if (far) {
  #pragma unroll
          for (int m = 0; m < 6; ++m)
            if (s_F[72 + m] < 0.0) rhs[m] = k == 0 ? s_data[m * kRow + 55] : 0.0;
        } else if (k == 1 && s_fix[0] != 0.0) {
          rhs[0] = s_fix[1];
        }
  #pragma unroll
        for (int m = 0; m < 6; ++m) r[m] = s_F[36 + 6 * i + m];
        s_uk[k * 6 + i] = dot6(r, rhs);
      }
    }
    __syncthreads();
  }
  if (active && l < 6) {  // Taylor expansion to the ghost point
    double ug = s_uk[l];
    double coef = 1.0;
    for (int k = 1; k <= 4; ++k) {
      coef *= dist / (double)k;
      ug = ug + coef * s_uk[k * 6 + l];
    }
    s_ug[l] = ug;
  }
  __syncthreads();
  if (active && l < 6) {  // rotation back to (x, y)
    const int f = l;
    const double out = f == 1 ? s_ug[1] * nx - s_ug[2] * ny : (f == 2 ? s_ug[1] * ny + s_ug[2] * nx : s_ug[f]);
    double* const S = reinterpret_cast<double*>(__double_as_longlong(s_sc[10]));
    const long plane = (long)__double_as_longlong(s_sc[11]), pad = (long)__double_as_longlong(s_sc[1]);
    S[f * plane + pad] = out;
    if (WALL && s_fix[0] == 0.0) {
      if (f == 0) raise_error(descs[sid].sc, phase, RH3T_ERR_SINGULAR, (unsigned)hg.gid[h]);
    } else if (!isfinite(out)) {
      raise_error(descs[sid].sc, phase, RH3T_ERR_GHOST, (unsigned)hg.gid[h]);
    }
  }
  if (wd > 0) {
    // Derived::from_state of the new ghost value (derived_from_state's
    // expressions), its first wd planes, spread over lanes 1..5 of the group:
    // lane f forms field f's quotient by rho (u, v, then s_3 / rho .. s_5 /
    // rho), shares it through LDS, then lanes 3..5 form e_k, p_k and the
    // temperature of their energy.  One quotient per lane instead of five
    // in sequence on one lane.
    double* const s_q = s_vk + 44;
    double q = 0.0;
    if (active && l >= 1 && l < 6) {
      const double num = l == 1 ? s_ug[1] * nx - s_ug[2] * ny : (l == 2 ? s_ug[1] * ny + s_ug[2] * nx : s_ug[l]);
      q = num / s_ug[0];
      s_q[l] = q;
    }
    __syncthreads();
    if (active && l >= 1 && l < 6) {
      const double rho = s_ug[0], u = s_q[1], v = s_q[2];
      const double w2 = sq(u) + sq(v);
      double* const Dp = reinterpret_cast<double*>(__double_as_longlong(s_sc[12]));
      const long plane = (long)__double_as_longlong(s_sc[11]), pad = (long)__double_as_longlong(s_sc[1]);
      if (l < 3) {
        Dp[(l == 1 ? D_U : D_V) * plane + pad] = q;
      } else {
        const int k = l - 3;
        const double e = q - w2 / 6.0;
        const double g = k == 0 ? ph.ge : (k == 1 ? ph.gi : ph.gr);
        Dp[(D_EE + k) * plane + pad] = e;
        if (D_PE + k < wd) Dp[(D_PE + k) * plane + pad] = g * (s_ug[l] - rho * w2 / 6.0);
        if (wd > 8) {
          double t;
          if (k < 2) {
            const double cv = k == 0 ? ph.cve : ph.cvi;
            t = s_ug[l] / (rho * cv) - w2 / (6.0 * cv);
          } else {
            t = pow(e * rho / ph.a_rad, 0.25);
          }
          Dp[(D_TE + k) * plane + pad] = t;
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_ghost_open(
    const SolveDesc* __restrict__ descs, HeavyGhosts hg, Phys ph, int buf, unsigned phase, int h0, int count,
    int wd) {
  ghost_grp<false>(descs, hg, ph, buf, phase, h0, count, wd);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_ghost_wall(
    const SolveDesc* __restrict__ descs, HeavyGhosts hg, Phys ph, int buf, unsigned phase, int h0, int count,
    int wd) {
  ghost_grp<true>(descs, hg, ph, buf, phase, h0, count, wd);
}

// ----------------------------------------------------------- derived

__global__ __launch_bounds__(256) void k_derived(const SolveDesc* __restrict__ descs, const int* __restrict__ tiles,
                                                 int nsolve, Phys ph, int buf, int nplanes) {
  const int s = find_solve(tiles, nsolve, blockIdx.x);
  const SolveDesc& sd = descs[s];
  const long p = (long)(blockIdx.x - tiles[s]) * blockDim.x + threadIdx.x;
  if (p >= sd.g.plane || glb(sd.pt)[p] == 0) return;
  const double* S = glb(sd.U[buf]);
  double* __restrict__ Dout = glb(sd.D);
  double st[6], d[D_N];
#pragma unroll
  for (int f = 0; f < 6; ++f) st[f] = S[f * sd.g.plane + p];
  derived_from_state(ph, st, d, nplanes > 8);
#pragma unroll
  for (int f = 0; f < D_N; ++f)
    if (f < nplanes) Dout[f * sd.g.plane + p] = d[f];
}

// ------------------------------------------------------- WENO flux

// Interface flux of one direction for every solve.  DIR 0: interface i-1/2
// of row j, i in [0, nx], stencil points i-3..i+2.  DIR 1: interface j-1/2
// of column i, j in [0, ny].  Output planes are i-major: x at i*ny + j, y at
// i*(ny+1) + j.  Workgroup shape (bx, by) is a launch parameter.
#ifndef RH3T_FLUX_WAVES
#define RH3T_FLUX_WAVES 2
#endif
// Flux of interface (i, j) of direction DIR (DIR 0: i - 1/2 at column j,
// DIR 1: j - 1/2 at row i) from buffer buf; 0.0 when neither neighbour is
// a fluid cell.  P receives the stencil (the caller may read it afterwards).
template <int DIR, bool FAST>
__device__ __forceinline__ void iface_flux(const SolveDesc& sd, const EigConst* __restrict__ Kp, double eps, int buf,
                                           int i, int j, double out[6], Stencil<FAST>& P) {
  const Geo g = sd.g;
  const double* __restrict__ S = glb(sd.U[buf]);
  const double* __restrict__ D = glb(sd.D);
  const uint8_t* __restrict__ pt = glb(sd.pt);
  const long c0 = pad_index(g, i, j);
  const long step = DIR == 0 ? g.PY : 1;
  // Both point types are read together and the flux is formed whether or
  // not the interface is active, then replaced by 0.0 if it is not: the
  // stencil loads need not wait for the point types (every stencil point
  // lies within the 4-point halo of the padded planes, so they are in
  // bounds either way), one memory round trip before the arithmetic instead
  // of three.  Values at points that are neither fluid nor ghost are never
  // used.
  const bool active = (pt[c0] == 1) | (pt[c0 - step] == 1);
  // 6-point stencil of U with each point's physical flux (EXACT,
  // flux_from_derived, src/state.rs:56-85) or its normal velocity and
  // pressures (FAST; see Stencil).  In x, each value is read as a workgroup-uniform
  // base (plane and stencil row, formed in scalar registers) plus this
  // lane's 32-bit byte offset: one per-lane address for all the loads.  In
  // y, plain indexing lets the compiler merge neighbouring points into
  // 16-byte loads.
  const unsigned boff = (unsigned)(c0 * (long)sizeof(double));
  auto at = [&](const double* plane, long q_off) {
    return DIR == 0 ? plane_at(glb(plane + q_off), boff) : plane[c0 + q_off];
  };
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const long qo = (q - 3) * step;
    double* Uq = P.U[q];
#pragma unroll
    for (int f = 0; f < 6; ++f) Uq[f] = at(S + f * g.plane, qo);
    const double u = at(D + D_U * g.plane, qo), v = at(D + D_V * g.plane, qo);
    const double pe = at(D + D_PE * g.plane, qo), pi = at(D + D_PI * g.plane, qo), pr = at(D + D_PR * g.plane, qo);
    const double ptot = pe + pi + pr;
    const double w = DIR == 0 ? u : v;
    if constexpr (FAST) {
      P.w[q] = w;
      P.pe[q] = pe;
      P.pi[q] = pi;
      P.pr[q] = pr;
      P.p[q] = ptot;
    } else {
      P.F[q][0] = Uq[0] * w;
      P.F[q][1] = DIR == 0 ? Uq[0] * u * u + ptot : Uq[0] * u * v;
      P.F[q][2] = DIR == 0 ? Uq[0] * u * v : Uq[0] * v * v + ptot;
      P.F[q][3] = (Uq[3] + pe) * w;
      P.F[q][4] = (Uq[4] + pi) * w;
      P.F[q][5] = (Uq[5] + pr) * w;
    }
  }
  const double e3[3] = {at(D + D_EE * g.plane, 0), at(D + D_EI * g.plane, 0), at(D + D_ER * g.plane, 0)};
  const double e2[3] = {at(D + D_EE * g.plane, -step), at(D + D_EI * g.plane, -step), at(D + D_ER * g.plane, -step)};
  const double u3 = at(D + D_U * g.plane, 0), v3 = at(D + D_V * g.plane, 0);
  const double u2 = at(D + D_U * g.plane, -step), v2 = at(D + D_V * g.plane, -step);
  interface_flux<DIR, FAST>(Kp, eps, P, u3, v3, e3, u2, v2, e2, out);
  if (!active) {
#pragma unroll
    for (int f = 0; f < 6; ++f) out[f] = 0.0;
  }
}

template <int DIR, bool FAST>
__global__ __launch_bounds__(256, RH3T_FLUX_WAVES) void k_flux(const SolveDesc* __restrict__ descs,
                                                               const int* __restrict__ tiles, int nsolve,
                                                               const EigConst* __restrict__ Kp, double eps,
                                                               int buf) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int s = find_solve(tiles, nsolve, t);
  const SolveDesc& sd = descs[s];
  const Geo g = sd.g;
  const int ni = DIR == 0 ? g.nx + 1 : g.nx;
  const int nj = DIR == 0 ? g.ny : g.ny + 1;
  const int tiles_j = (nj + blockDim.x - 1) / blockDim.x;
  const int lt = t - tiles[s];
  const int j = (lt % tiles_j) * blockDim.x + threadIdx.x;
  const int i = (lt / tiles_j) * blockDim.y + threadIdx.y;
  if (i >= ni || j >= nj) return;
  double* __restrict__ F = glb(DIR == 0 ? sd.FX : sd.FY);
  const long fpl = DIR == 0 ? (long)(g.nx + 1) * g.ny : (long)g.nx * (g.ny + 1);
  const long fidx = DIR == 0 ? (long)i * g.ny + j : (long)i * (g.ny + 1) + j;
  double out[6];
  Stencil<FAST> P;
  iface_flux<DIR, FAST>(sd, Kp, eps, buf, i, j, out, P);
#pragma unroll
  for (int f = 0; f < 6; ++f) F[f * fpl + fidx] = out[f];
}

// --------------------------------------------------------- assemble

// One cell (i, j) of solve sd: l()'s right-hand side and, for sa.mode >= 1,
// the stage update.  Returns get_local_dt of the updated state after stage 3
// (the dt rule of the next step), +inf otherwise.
// FAST: the noncon divisions as products (noncon_fast); EXACT: the
// reference's divisions.
// FYB / FYT: the y fluxes at j - 1/2 and j + 1/2, read from the FY planes
// (FROM_PLANES) or passed in by the fused y-flux kernel, which also passes
// what it already holds from the y stencil of interface j - 1/2: the state
// of the cell in buffer src (uc_in, used for a stage's base or added state
// when that is buffer src) and, in FAST, the y noncon terms (ncy_in).
template <bool FAST, bool FROM_PLANES>
__device__ __forceinline__ double assemble_core(const SolveDesc& sd, const Phys& ph, const StageSpec& sa, int i, int j,
                                                const double* fyb_in, const double* fyt_in,
                                                const double* ncy_in = nullptr, const double* uc_in = nullptr,
                                                int src = -1) {
  constexpr bool NCY_IN = FAST && !FROM_PLANES;
  const Geo g = sd.g;
  const long c0 = pad_index(g, i, j);
  const long lin = (long)i * g.ny + j;
  const long ncell = (long)g.nx * g.ny;
  double* __restrict__ RHS = glb(sd.RHS);
  if (glb(sd.pt)[c0] != 1) {
    if (sa.mode == 0)
#pragma unroll
      for (int f = 0; f < 6; ++f) RHS[f * ncell + lin] = 0.0;
    return INFINITY;
  }
  const double* __restrict__ D = glb(sd.D);
  const double* __restrict__ FX = glb(sd.FX);
  const long fxpl = (long)(g.nx + 1) * g.ny;
  const long ixl = (long)i * g.ny + j, ixr = ixl + g.ny;
  const double inv_dx = g.rdx, inv_dy = g.rdy;
  double r[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    double fyb, fyt;
    if (FROM_PLANES) {
      const double* __restrict__ FY = glb(sd.FY);
      const long fypl = (long)g.nx * (g.ny + 1);
      const long iyb = (long)i * (g.ny + 1) + j, iyt = iyb + 1;
      fyb = FY[f * fypl + iyb];
      fyt = FY[f * fypl + iyt];
    } else {
      fyb = fyb_in[f];
      fyt = fyt_in[f];
    }
    const double fxt = -(FX[f * fxpl + ixr] - FX[f * fxpl + ixl]) * inv_dx;
    const double fy_ = -(fyt - fyb) * inv_dy;
    r[f] = fxt + fy_;
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
    if (FAST)
      noncon_fast(vel, pe, pi, pr, inv_dx, ncx);
    else
      noncon(vel, pe, pi, pr, g.dx, g.rdx, g.r60dx, ncx);
    if constexpr (NCY_IN) {
#pragma unroll
      for (int e = 0; e < 3; ++e) ncy[e] = ncy_in[e];
    } else {
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const long p = c0 + (k - 4);
        vel[k] = D[D_V * g.plane + p];
        pe[k] = D[D_PE * g.plane + p];
        pi[k] = D[D_PI * g.plane + p];
        pr[k] = D[D_PR * g.plane + p];
      }
      if (FAST)
        noncon_fast(vel, pe, pi, pr, inv_dy, ncy);
      else
        noncon(vel, pe, pi, pr, g.dy, g.rdy, g.r60dy, ncy);
    }
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
    const double cx = g.cdx, cy = g.cdy;
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
    for (int f = 0; f < 6; ++f) RHS[f * ncell + lin] = r[f];
    return INFINITY;
  }
  // stage_update_rhs / stage_update_comb, src/main.rs:307-349
  const double dt = sd.sc->dt;
  const double* __restrict__ base = glb(sd.U[sa.base]);
  const double* __restrict__ add = glb(sd.U[sa.add]);
  double* dst = glb(sd.U[sa.dst]);
  double v[6], bv[6], av[6];
  if (!FROM_PLANES && sa.base == src) {
#pragma unroll
    for (int f = 0; f < 6; ++f) bv[f] = uc_in[f];
  } else {
#pragma unroll
    for (int f = 0; f < 6; ++f) bv[f] = base[f * g.plane + c0];
  }
  if (sa.mode >= 2) {
    if (!FROM_PLANES && sa.add == src) {
#pragma unroll
      for (int f = 0; f < 6; ++f) av[f] = uc_in[f];
    } else {
#pragma unroll
      for (int f = 0; f < 6; ++f) av[f] = add[f * g.plane + c0];
    }
  }
  if (sa.mode == 1) {
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = bv[f] + r[f] * dt;
  } else if (sa.mode == 2) {
    const double coef = dt / 4.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = bv[f] * 0.75 + av[f] * 0.25 + r[f] * coef;
  } else {
    const double coef = 2.0 * dt / 3.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) v[f] = bv[f] * (1.0 / 3.0) + av[f] * (2.0 / 3.0) + r[f] * coef;
  }
#pragma unroll
  for (int f = 0; f < 6; ++f) dst[f * g.plane + c0] = v[f];
  if (sa.wd > 0) {  // Derived of the new state, for the next stage's kernels
    double d[D_N];
    derived_from_state(ph, v, d, sa.wd > 8);
    double* __restrict__ Dn = glb(sd.Dn);
#pragma unroll
    for (int f = 0; f < D_N; ++f)
      if (f < sa.wd) Dn[f * g.plane + c0] = d[f];
  }
  if (!admissible(v)) raise_error(sd.sc, sa.phase, RH3T_ERR_NONPHYSICAL, (unsigned)lin);
  return sa.mode == 3 ? local_dt(ph, v, g.dx, g.dy) : INFINITY;
}

// fmin over the workgroup, written to partial[tile]: a wavefront reduction
// (lane shuffles) per wavefront, then over the <= 4 wavefronts.  Every lane
// of the workgroup takes part.  fmin is exact and ignores NaN like the
// serial f64::min of src/main.rs:402-403, so the result does not depend on
// the reduction order.
__device__ __forceinline__ void put_group_min(double m, double* __restrict__ partial, int t) {
  __shared__ double s_min[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
  const int tid = threadIdx.y * blockDim.x + threadIdx.x, nw = (blockDim.x * blockDim.y) >> 6;
  if ((tid & 63) == 0) s_min[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    double r = s_min[0];
    for (int w = 1; w < nw; ++w) r = fmin(r, s_min[w]);
    partial[t] = r;
  }
}

// Assembly + stage update of every solve; the last RK stage also leaves the
// per-workgroup minima of get_local_dt of the new state (src/dt.rs:5-34)
// for the next step's dt.
template <bool FAST>
__global__ __launch_bounds__(256) void k_assemble(const SolveDesc* __restrict__ descs, const int* __restrict__ tiles,
                                                  int nsolve, Phys ph, StageSpec sa, double* __restrict__ partial) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int s = find_solve(tiles, nsolve, t);
  const SolveDesc& sd = descs[s];
  const int tiles_j = (sd.g.ny + blockDim.x - 1) / blockDim.x;
  const int lt = t - tiles[s];
  const int j = (lt % tiles_j) * blockDim.x + threadIdx.x;
  const int i = (lt / tiles_j) * blockDim.y + threadIdx.y;
  const double m =
      (i < sd.g.nx && j < sd.g.ny) ? assemble_core<FAST, true>(sd, ph, sa, i, j, nullptr, nullptr) : INFINITY;
  if (sa.mode == 3) put_group_min(m, partial, t);
}

// The y fluxes and the assembly in one pass, so that the FY planes are
// neither written nor read back: a row of blockDim.x lanes (a power of two
// up to 64, so a row lies within one wavefront) takes the y interfaces
// j0 - 1/2 .. j0 + blockDim.x - 3/2 and assembles the blockDim.x - 1 cells
// j0 .. j0 + blockDim.x - 2 (the top flux of a cell is the bottom flux of
// the next lane, passed by a lane shuffle).  Same arithmetic as k_flux<1>
// followed by k_assemble.
template <bool FAST>
__global__ __launch_bounds__(256, RH3T_FLUX_WAVES) void k_fy_assemble(const SolveDesc* __restrict__ descs,
                                                                      const int* __restrict__ tiles, int nsolve,
                                                                      const EigConst* __restrict__ Kp, Phys ph,
                                                                      int src, StageSpec sa,
                                                                      double* __restrict__ partial) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int s = find_solve(tiles, nsolve, t);
  const SolveDesc& sd = descs[s];
  const int nx = sd.g.nx, ny = sd.g.ny;
  const int cw = blockDim.x - 1;
  const int tiles_j = (ny + cw - 1) / cw;
  const int lt = t - tiles[s];
  const int j = (lt % tiles_j) * cw + threadIdx.x;
  const int i = (lt / tiles_j) * blockDim.y + threadIdx.y;
  double fb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double uc[6], ncy[3];
  if (i < nx && j <= ny) {
    Stencil<FAST> P;
    iface_flux<1, FAST>(sd, Kp, ph.eps, src, i, j, fb, P);
#pragma unroll
    for (int f = 0; f < 6; ++f) uc[f] = P.U[3][f];  // point j of the stencil: the cell
    if constexpr (FAST) {
      // The y noncon line of cell j (points j-4 .. j+4, src/noncon.rs:288-332):
      // the stencil holds points j-3 .. j+2 of v, p_e, p_i, p_r; the other
      // three are read here (for j = ny, not a cell, they may fall past the
      // halo into the next padded row: in bounds, not used).
      const Geo g = sd.g;
      const double* __restrict__ D = glb(sd.D);
      const long c0 = pad_index(g, i, j);
      double vel[9], pe[9], pi[9], pr[9];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        vel[q + 1] = P.w[q];
        pe[q + 1] = P.pe[q];
        pi[q + 1] = P.pi[q];
        pr[q + 1] = P.pr[q];
      }
      vel[0] = vel[7] = vel[8] = 0.0;  // noncon_fast reads vel[3..5] only
      const int ko[3] = {0, 7, 8};
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const long p = c0 + (ko[e] - 4);
        pe[ko[e]] = D[D_PE * g.plane + p];
        pi[ko[e]] = D[D_PI * g.plane + p];
        pr[ko[e]] = D[D_PR * g.plane + p];
      }
      noncon_fast(vel, pe, pi, pr, g.rdy, ncy);
    }
  }
  double ft[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) ft[f] = __shfl_down(fb[f], 1, blockDim.x);
  const double m = ((int)threadIdx.x < cw && i < nx && j < ny)
                       ? assemble_core<FAST, false>(sd, ph, sa, i, j, fb, ft, ncy, uc, src)
                       : INFINITY;
  if (sa.mode == 3) put_group_min(m, partial, t);
}

// One thread per solve: copies the state of the current step's first
// failure, if its phase is in [lo, hi], into that solve's scalars.  Runs
// after the stages whose phases it covers (once per step in the time loop):
// no stage writes a buffer that an earlier phase of the same step failed in
// -- phase 2k (the ghosts of stage k) fails in its source buffer U[k - 1]
// (U[0] for an l()), phase 2k + 1 (stage k) in its destination, U[1], U[2]
// and U[3] (the alias of U[0], or the tape buffer the adjoint's forward
// points it at) -- so the failing values are still there.
__device__ __forceinline__ void capture_failure(const SolveDesc& sd, unsigned lo, unsigned hi) {
  Scalars* sc = sd.sc;
  unsigned long long key = sc->err;
  if (key == ~0ULL) return;
  const unsigned phase = (unsigned)((key >> 40) & 0xF);
  if (((key >> 44) & 0xFFFFF) != (sc->step & 0xFFFFF) || phase < lo || phase > hi || phase < 2 || phase > 7) return;
  const int buf = phase == 7 ? 3 : (int)(phase / 2) - (phase % 2 == 0 ? 1 : 0);
  unsigned idx = (unsigned)(key & 0xFFFFFFFFULL);
  long p = (phase % 2 == 0) ? sd.gpad[idx] : pad_index(sd.g, idx / sd.g.ny, idx % sd.g.ny);
  const double* b = sd.U[buf];
  for (int f = 0; f < 6; ++f) sc->err_state[f] = b[f * sd.g.plane + p];
}

__global__ void k_capture(const SolveDesc* __restrict__ descs, int nsolve, unsigned lo, unsigned hi) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsolve) capture_failure(descs[s], lo, hi);
}

// ------------------------------------------------------------- dt

// get_local_dt (src/dt.rs:5-34) of the state in U0 on every fluid cell,
// reduced per workgroup; same tiles and partial layout as the assembly
// kernel in use (cw cells per tile row: blockDim.x for k_assemble,
// blockDim.x - 1 for k_fy_assemble).  Used when the state did not come out
// of a stage 3 (first step, host writes).
__global__ __launch_bounds__(256) void k_dt_partial(const SolveDesc* __restrict__ descs,
                                                    const int* __restrict__ tiles, int nsolve, Phys ph,
                                                    double* __restrict__ partial, int cw) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int s = find_solve(tiles, nsolve, t);
  const SolveDesc& sd = descs[s];
  const int tiles_j = (sd.g.ny + cw - 1) / cw;
  const int lt = t - tiles[s];
  const int j = (lt % tiles_j) * cw + threadIdx.x;
  const int i = (lt / tiles_j) * blockDim.y + threadIdx.y;
  double m = INFINITY;
  if ((int)threadIdx.x < cw && i < sd.g.nx && j < sd.g.ny) {
    const long p = pad_index(sd.g, i, j);
    if (glb(sd.pt)[p] == 1) {
      const double* S = glb(sd.U[0]);
      double st[6];
#pragma unroll
      for (int f = 0; f < 6; ++f) st[f] = S[f * sd.g.plane + p];
      m = local_dt(ph, st, sd.g.dx, sd.g.dy);
    }
  }
  put_group_min(m, partial, t);
}

// One workgroup (1024 lanes) per solve over its workgroup minima:
// calc_global_dt (src/main.rs:375-408) and, when advancing,
// dt = cfl_factor * dt_cfl (src/main.rs:461) or the fixed dt, the step
// counter and Field.time.
// advance: a time-loop step begins (step, dt and time move on); it first
// captures the previous step's first failure, phases 2..7 (the previous
// step's buffers are intact until this step's stages run), which spares the
// time loop a capture launch per step.
__global__ __launch_bounds__(1024) void k_dt_final(const SolveDesc* __restrict__ descs,
                                                   const int* __restrict__ tiles,
                                                   const double* __restrict__ partial, int advance) {
  __shared__ double red[16];
  const int s = blockIdx.x;
  // sixteen independent loads in flight per thread, the last partial
  // round predicated rather than walked one load at a time (fmin is exact
  // and order independent: the same minimum in any grouping)
  const int k1 = tiles[s + 1];
  double mm[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) mm[u] = INFINITY;
  int k = tiles[s] + threadIdx.x;
  for (; k + 15 * 1024 < k1; k += 16 * 1024)
#pragma unroll
    for (int u = 0; u < 16; ++u) mm[u & 7] = fmin(mm[u & 7], partial[k + u * 1024]);
  // the last fewer than 16 of this thread's partials, loaded together
  if (k < k1) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = k + u * 1024 < k1 ? partial[k + u * 1024] : INFINITY;
#pragma unroll
    for (int u = 0; u < 16; ++u) mm[u & 7] = fmin(mm[u & 7], v[u]);
  }
  double m = mm[0];
#pragma unroll
  for (int u = 1; u < 8; ++u) m = fmin(m, mm[u]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double dt_cfl = red[0];
    for (int w = 1; w < 16; ++w) dt_cfl = fmin(dt_cfl, red[w]);
    Scalars* sc = descs[s].sc;
    sc->dt_cfl = dt_cfl;
    if (advance) {
      capture_failure(descs[s], 2, 7);
      sc->step = sc->step + 1;
      double dt = sc->dt_fixed > 0.0 ? sc->dt_fixed : sc->cfl_factor * dt_cfl;
      sc->dt = dt;
      if (!(dt > 0.0)) atomicMin(&sc->err, err_key(sc->step, 1, RH3T_ERR_DT, 0));
      sc->t_n = sc->time;
      sc->time = sc->time + dt;
    }
  }
}

__global__ void k_err_clean(const SolveDesc* __restrict__ descs, int nsolve) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsolve) descs[s].sc->err = ~0ULL;
}

// ----------------------------------------------------------- layout I/O

// host AoS (i*ny+j, 6 doubles) <-> padded SoA on fluid cells of one solve
__global__ __launch_bounds__(256) void k_aos_to_pad(const double* __restrict__ aos, SolveDesc sd) {
  long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lin >= (long)sd.g.nx * sd.g.ny) return;
  long p = pad_index(sd.g, lin / sd.g.ny, lin % sd.g.ny);
  if (sd.pt[p] != 1) return;
#pragma unroll
  for (int f = 0; f < 6; ++f) sd.U[0][f * sd.g.plane + p] = aos[6 * lin + f];
}

__global__ __launch_bounds__(256) void k_pad_to_aos(SolveDesc sd, double* __restrict__ aos) {
  long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lin >= (long)sd.g.nx * sd.g.ny) return;
  long p = pad_index(sd.g, lin / sd.g.ny, lin % sd.g.ny);
  if (sd.pt[p] != 1) return;
#pragma unroll
  for (int f = 0; f < 6; ++f) aos[6 * lin + f] = sd.U[0][f * sd.g.plane + p];
}

__global__ __launch_bounds__(256) void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos,
                                                    long n) {
  long lin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lin >= n) return;
#pragma unroll
  for (int f = 0; f < 6; ++f) aos[6 * lin + f] = soa[f * n + lin];
}

__global__ __launch_bounds__(256) void k_gather_ghosts(SolveDesc sd, int n, double* __restrict__ out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
#pragma unroll
  for (int f = 0; f < 6; ++f) out[6L * k + f] = sd.U[0][f * sd.g.plane + sd.gpad[k]];
}

}  // namespace

// ======================================================================
// Engine: device state of a batch of solves
// ======================================================================

namespace rh3t_eng {
const char* const kKernelNames[K_COUNT] = {"ghost_cheap", "ghost_heavy", "derived", "flux_x", "flux_y",
                                           "assemble", "dt_partial", "dt_final", "capture", "adj_iface",
                                           "adj_gather", "adj_ghost", "adj_param", "flux_y_assemble",
                                           "ghost_prep"};
}  // namespace rh3t_eng

namespace {

// The y fluxes are formed inside the assembly pass (k_fy_assemble) when a
// row of the workgroup is a power of two of lanes within one wavefront.
bool fused_fy(const Engine& E) { return E.bx == 16 || E.bx == 32 || E.bx == 64; }

// prefix tables for the current launch shape
int build_tables(Engine& E) {
  const int n = E.nsolve;
  E.t_fx.assign(n + 1, 0);
  E.t_fy.assign(n + 1, 0);
  E.t_as.assign(n + 1, 0);
  E.t_der.assign(n + 1, 0);
  E.t_fa.assign(n + 1, 0);
  for (int s = 0; s < n; ++s) {
    const Geo& g = E.solves[s].g;
    auto tiles = [&](int ni, int nj) { return ((nj + E.bx - 1) / E.bx) * ((ni + E.by - 1) / E.by); };
    const int xbx = E.fx_bx(), xby = E.fx_by();
    E.t_fx[s + 1] = E.t_fx[s] + ((g.ny + xbx - 1) / xbx) * ((g.nx + 1 + xby - 1) / xby);
    E.t_fy[s + 1] = E.t_fy[s] + tiles(g.nx, g.ny + 1);
    E.t_as[s + 1] = E.t_as[s] + tiles(g.nx, g.ny);
    E.t_der[s + 1] = E.t_der[s] + (int)((g.plane + 255) / 256);
    E.t_fa[s + 1] = E.t_fa[s] + ((g.ny + E.bx - 2) / (E.bx - 1)) * ((g.nx + E.by - 1) / E.by);
  }
  std::vector<int> all;
  for (auto* t : {&E.t_fx, &E.t_fy, &E.t_as, &E.t_der, &E.t_fa}) all.insert(all.end(), t->begin(), t->end());
  if (!E.d_tabs) {
    int st = dev_alloc(E, &E.d_tabs, all.size());
    if (st) return st;
  }
  HIP_TRY(hipMemcpy(E.d_tabs, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  // one dt partial per assembly workgroup
  const long need = (long)std::max(total(E.t_as), total(E.t_fa));
  if (need > E.partial_cap) {
    int st = dev_alloc(E, &E.d_partial, (size_t)need);
    if (st) return st;
    E.partial_cap = need;
  }
  E.dt_partials_valid = false;
  if (E.graph) {
    (void)hipGraphExecDestroy(E.graph);
    E.graph = nullptr;
  }
  return RH3T_OK;
}

}  // namespace

namespace rh3t_eng {

// The ghost update of buffer buf with descriptors desc; wd > 0: also the
// ghost slots' Derived planes (their first wd), formed by the kernel that
// writes each ghost value -- except for the heavy ghosts of a context with
// many of them, whose Derived values a pass of their own forms.  In the
// heavy-ghost kernels the Derived values take a serial tail of a few lanes
// of each two-ghost wavefront (~0.1 us per 1,000 ghosts and launch); the
// separate pass takes a launch (~6 us) plus ~0.055 us per 1,000 slots at
// full lane use: below ~130k heavy ghosts the tail costs less (a 121x481
// cylinder, 4,480 ghosts: the RK3 step 7% shorter), above it the pass (the
// 64-solve batch, 434k ghosts: 0.13 against 0.10 ms per step).
void ghosts_on(Engine& E, const SolveDesc* desc, int buf, unsigned phase, int gt, int wd) {
  const int n_open = E.heavy.n_open;
  const bool with_prep = E.heavy.n > 0 && E.cheap_with_prep;
  if (E.cheap.n > 0 && !with_prep) {
    Launcher L(E, K_GHOST_CHEAP);
    k_ghost_cheap<<<(E.cheap.n + 255) / 256, 256, 0, E.stream>>>(desc, E.cheap, buf, phase, gt, E.phys, wd);
  }
  if (E.heavy.n > 0) {
    // the frames and packed points, with the cheap ghosts riding along
    // when nothing the first two read is a ghost point
    Launcher L(E, K_GHOST_PREP);
    const int nbf = (E.frames.n + 63) / 64, nbp = nbf + (E.packed.n + 63) / 64;
    const int nbc = with_prep ? (E.cheap.n + 63) / 64 : 0;
    k_ghost_prep<<<nbp + nbc, 64, 0, E.stream>>>(desc, E.frames, E.packed, E.cheap, E.phys, buf, phase, gt, nbf,
                                                 nbp, wd);
  }
  const bool heavy_pass = E.heavy.n >= kHeavyDerivedPass;
  const int wdh = heavy_pass ? 0 : wd;
  if (E.heavy.n > 0) {
    Launcher L(E, K_GHOST_HEAVY);
    const int g = E.ghosts_per_block, n_wall = E.heavy.n - n_open;
    if (n_open > 0) {
      const int per = g * kOpenPerWave;  // ghosts per workgroup
      k_ghost_open<<<(n_open + per - 1) / per, dim3(64, g), (size_t)per * kOpenLds * sizeof(double), E.stream>>>(
          desc, E.heavy, E.phys, buf, phase, 0, n_open, wdh);
    }
    if (n_wall > 0) {
      const int per = g * kOpenPerWave;
      k_ghost_wall<<<(n_wall + per - 1) / per, dim3(64, g), (size_t)per * grp_lds(true) * sizeof(double),
                     E.stream>>>(desc, E.heavy, E.phys, buf, phase, n_open, n_wall, wdh);
    }
  }
  if (wd > 0 && heavy_pass) {
    Launcher L(E, K_DERIVED);
    k_derived_ghosts<<<(E.heavy.n + 255) / 256, 256, 0, E.stream>>>(desc, E.heavy, E.phys, buf, wd);
  }
}

void launch_ghosts(Engine& E, int buf, unsigned phase, int gt) { ghosts_on(E, E.d_desc, buf, phase, gt, 0); }

void launch_capture(Engine& E, unsigned lo, unsigned hi) {
  Launcher L(E, K_CAPTURE);
  k_capture<<<(E.nsolve + 63) / 64, 64, 0, E.stream>>>(E.d_desc, E.nsolve, lo, hi);
}

void derived_on(Engine& E, const SolveDesc* desc, int buf) {
  Launcher L(E, K_DERIVED);
  k_derived<<<total(E.t_der), 256, 0, E.stream>>>(desc, E.tab(3), E.nsolve, E.phys, buf, E.nplanes);
}

void launch_derived(Engine& E, int buf) {
  derived_on(E, E.d_desc, buf);
  E.dvalid = 0;  // the planes of U[0] now hold another state's (or U[0]'s ghosts only)
}

// Derived of every point of U[buf] into its own planes, if not current
void ensure_derived(Engine& E, int buf) {
  if (E.dvalid & (1u << buf)) return;
  derived_on(E, E.d_desc3 + (size_t)buf * E.nsolve, buf);
  E.dvalid |= 1u << buf;
}

void launch_stage(Engine& E, int src, const StageSpec& sa_in, bool kept) {
  StageSpec sa = sa_in;
  const SolveDesc* desc = E.d_desc;
  const unsigned phase = sa.mode == 0 ? 2u : (unsigned)(2 * sa.mode);
  if (kept) {
    desc = E.d_desc3 + (size_t)src * E.nsolve;
    if (E.dvalid & (1u << src)) {
      ghosts_on(E, desc, src, phase, sa.mode, E.nplanes);
    } else {
      ghosts_on(E, desc, src, phase, sa.mode, 0);
      derived_on(E, desc, src);
      E.dvalid |= 1u << src;
    }
    // the assembly fills the planes of the new state, U[(src + 1) % 3]
    if (sa.mode >= 1 && sa.dst == (src + 1) % 3) {
      sa.wd = E.nplanes;
      E.dvalid |= 1u << sa.dst;
    } else {
      sa.wd = 0;
      if (sa.mode >= 1) E.dvalid &= ~(1u << sa.dst);
    }
  } else {
    sa.wd = 0;
    ghosts_on(E, desc, src, phase, sa.mode, 0);
    derived_on(E, desc, src);
    E.dvalid = 0;
  }
  dim3 blk(E.bx, E.by);
  {
    Launcher L(E, K_FLUX_X);
    const dim3 xblk(E.fx_bx(), E.fx_by());
    if (E.fast)
      k_flux<0, true><<<total(E.t_fx), xblk, 0, E.stream>>>(desc, E.tab(0), E.nsolve, E.d_eig, E.phys.eps, src);
    else
      k_flux<0, false><<<total(E.t_fx), xblk, 0, E.stream>>>(desc, E.tab(0), E.nsolve, E.d_eig, E.phys.eps, src);
  }
  if (fused_fy(E)) {
    Launcher L(E, K_FY_ASSEMBLE);
    if (E.fast)
      k_fy_assemble<true><<<total(E.t_fa), blk, 0, E.stream>>>(desc, E.tab(4), E.nsolve, E.d_eig, E.phys, src,
                                                               sa, E.d_partial);
    else
      k_fy_assemble<false><<<total(E.t_fa), blk, 0, E.stream>>>(desc, E.tab(4), E.nsolve, E.d_eig, E.phys, src,
                                                                sa, E.d_partial);
  } else {
    {
      Launcher L(E, K_FLUX_Y);
      if (E.fast)
        k_flux<1, true><<<total(E.t_fy), blk, 0, E.stream>>>(desc, E.tab(1), E.nsolve, E.d_eig, E.phys.eps, src);
      else
        k_flux<1, false><<<total(E.t_fy), blk, 0, E.stream>>>(desc, E.tab(1), E.nsolve, E.d_eig, E.phys.eps,
                                                              src);
    }
    Launcher L(E, K_ASSEMBLE);
    if (E.fast)
      k_assemble<true><<<total(E.t_as), blk, 0, E.stream>>>(desc, E.tab(2), E.nsolve, E.phys, sa, E.d_partial);
    else
      k_assemble<false><<<total(E.t_as), blk, 0, E.stream>>>(desc, E.tab(2), E.nsolve, E.phys, sa, E.d_partial);
  }
}

// get_local_dt minima of U0 per assembly workgroup (the layout k_dt_final reads)
void launch_dt_partial(Engine& E) {
  Launcher L(E, K_DT_PARTIAL);
  const bool fu = fused_fy(E);
  k_dt_partial<<<total(fu ? E.t_fa : E.t_as), dim3(E.bx, E.by), 0, E.stream>>>(E.d_desc, E.tab(fu ? 4 : 2), E.nsolve,
                                                                             E.phys, E.d_partial,
                                                                             fu ? E.bx - 1 : E.bx);
}

// dt of every solve from the wavefront minima; fresh: recompute them from
// U0 first (otherwise the last stage 3 left them for the current state)
void launch_dt(Engine& E, int advance, bool fresh) {
  if (fresh) launch_dt_partial(E);
  Launcher L(E, K_DT_FINAL);
  k_dt_final<<<E.nsolve, 1024, 0, E.stream>>>(E.d_desc, E.tab(fused_fy(E) ? 4 : 2), E.d_partial, advance);
}

// rk3_ssp, src/main.rs:351-373, with the state kept in U0:
//   stage 1: U1 = U0 + dt L(U0)
//   stage 2: U2 = 3/4 U0 + 1/4 U1 + dt/4 L(U1)
//   stage 3: U0 = 1/3 U0 + 2/3 U2 + 2dt/3 L(U2)   (in place: each cell
//            reads its own U0 value before writing it)
void launch_rk3(Engine& E, bool capture) {
  launch_stage(E, 0, StageSpec{1, 0, 0, 1, 3}, true);
  launch_stage(E, 1, StageSpec{2, 0, 1, 2, 5}, true);
  launch_stage(E, 2, StageSpec{3, 0, 2, 0, 7}, true);
  if (capture) launch_capture(E, 2, 7);
}

// Reads every solve's first-failure key; returns the first failing status.
int check_errors(Engine& E) {
  std::vector<Scalars> sc(E.nsolve);
  HIP_TRY(hipMemcpyAsync(sc.data(), E.d_sc, sizeof(Scalars) * E.nsolve, hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  int first = RH3T_OK;
  E.last.assign(E.nsolve, rh3t_error{});
  for (int s = 0; s < E.nsolve; ++s) {
    if (sc[s].err == ~0ULL) continue;
    unsigned long long k = sc[s].err;
    unsigned phase = (unsigned)((k >> 40) & 0xF), kind = (unsigned)((k >> 32) & 0xFF);
    unsigned idx = (unsigned)(k & 0xFFFFFFFFULL);
    rh3t_error& e = E.last[s];
    e.kind = (int)kind;
    // ghost-update failures (even phases) report stage 0, as the reference's
    // GhostGrid panic names no RK stage; state failures report RK1..RK3
    e.stage = (phase >= 3 && phase % 2 == 1) ? (int)((phase - 3) / 2 + 1) : 0;
    e.ghost_id = -1;
    e.i = e.j = -1;
    std::memcpy(e.state, sc[s].err_state, sizeof(e.state));
    if (kind == RH3T_ERR_DT) {
      std::snprintf(e.message, sizeof(e.message), "solve %d: non-positive dt at step %u", s, (unsigned)(k >> 44));
    } else if (phase % 2 == 0) {
      e.ghost_id = (int)idx;
      std::snprintf(e.message, sizeof(e.message), "solve %d: NON-FINITE GHOST STATE ghost id %u", s, idx);
    } else {
      e.i = (int)(idx / E.solves[s].g.ny);
      e.j = (int)(idx % E.solves[s].g.ny);
      std::snprintf(e.message, sizeof(e.message), "solve %d: non-physical state at (%d, %d) in RK%d state", s, e.i,
                    e.j, e.stage);
    }
    if (first == RH3T_OK) first = (int)kind;
  }
  return first;
}

int reset_errors(Engine& E) {
  k_err_clean<<<(E.nsolve + 63) / 64, 64, 0, E.stream>>>(E.d_desc, E.nsolve);
  HIP_TRY(hipGetLastError());
  return RH3T_OK;
}

int set_arith(Engine& E, bool fast) {
  if (fast != E.fast && E.graph) {
    (void)hipGraphExecDestroy(E.graph);
    E.graph = nullptr;
  }
  E.fast = fast;
  return RH3T_OK;
}

}  // namespace rh3t_eng

namespace {

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
    K.rcp_six_g[k] = 1.0 / K.six_g[k];
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

}  // namespace

namespace rh3t_eng {

Phys physics_of(const rh3t_physics& p) { return make_phys(p); }

int apply_physics(Engine& E, const rh3t_physics& p) {
  E.src_phys = p;
  E.phys = make_phys(p);
  E.eig = make_eig(E.phys);
  HIP_TRY(hipMemcpy(E.d_eig, &E.eig, sizeof(EigConst), hipMemcpyHostToDevice));
  E.nplanes = (E.adj || E.phys.diffusion || E.phys.source) ? D_N : 8;
  // the stored get_local_dt minima and Derived planes were formed with the
  // old constants
  E.dt_partials_valid = false;
  E.dvalid = 0;
  if (E.graph) {
    (void)hipGraphExecDestroy(E.graph);
    E.graph = nullptr;
  }
  return RH3T_OK;
}

}  // namespace rh3t_eng

namespace {

void free_engine(Engine& E) {
  (void)hipSetDevice(E.device);
  // nothing queued on the stream may still use the graph or the buffers
  if (E.stream) (void)hipStreamSynchronize(E.stream);
  if (E.graph) (void)hipGraphExecDestroy(E.graph);
  for (auto& r : E.prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  free_adjoint(E);
  for (void* p : E.allocs) (void)hipFree(p);
  if (E.stream) (void)hipStreamDestroy(E.stream);
  E.allocs.clear();
}

// Appends the swizzled row block (kSwzLen doubles, layout at kSwzA in
// engine.hpp) of one heavy ghost's S (979) and w (225) tables.  S_r row a is
// at sum_{i<r} (i+1)^4 + a (r+1)^2, w_r row k at sum_{i<r} (i+1)^3 + k (r+1)^2.
static void swizzle_rows(const double* S, const double* W, std::vector<double>& out) {
  auto s_at = [&](int r, int a, int b) {
    const int m = (r + 1) * (r + 1);
    return S[r * (r + 1) * (2 * r + 1) * (3 * r * r + 3 * r - 1) / 30 + a * m + b];
  };
  auto w_at = [&](int r, int k, int b) {
    const int m = (r + 1) * (r + 1);
    return W[r * r * (r + 1) * (r + 1) / 4 + k * m + b];
  };
  const size_t base = out.size();
  out.resize(base + kSwzLen, 0.0);
  double* o = out.data() + base;
  for (int b = 0; b < kSwzA; ++b)
    for (int l = 0; l < 32; ++l) {
      double v = 0.0;
      if (l < 25)
        v = s_at(4, l, b);
      else if (l < 30)
        v = w_at(4, l - 25, b);
      else if (b == 0)
        v = l == 30 ? s_at(0, 0, 0) : w_at(0, 0, 0);
      o[(b / 2) * 64 + 2 * l + b % 2] = v;
    }
  for (int b = 0; b < kSwzB; ++b)
    for (int l = 0; l < 32; ++l) {
      double v = 0.0;
      if (l < 16)
        v = s_at(3, l, b);
      else if (l < 20)
        v = w_at(3, l - 16, b);
      else if (b < 9)
        v = l < 29 ? s_at(2, l - 20, b) : w_at(2, l - 29, b);
      else if (b < 13 && l < 26)
        v = l < 24 ? s_at(1, l - 20, b - 9) : w_at(1, l - 24, b - 9);
      o[(kSwzPairsA + b / 2) * 64 + 2 * l + b % 2] = v;
    }
}

int build_engine(Engine& E, const rh3t::Problem* const* P, int n, const rh3t_physics& phys, int device) {
  E.device = device;
  E.src_phys = phys;
  E.phys = make_phys(phys);
  E.eig = make_eig(E.phys);
  E.nsolve = n;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
  E.nplanes = (E.phys.diffusion || E.phys.source) ? D_N : 8;
  E.solves.resize(n);
  E.h_desc.resize(n);
  std::vector<std::array<double*, 3>> dplanes(n);  // Derived planes of U0, U1, U2 per solve
  E.last.assign(n, rh3t_error{});
  int st;
  if ((st = dev_alloc(E, &E.d_sc, (size_t)n))) return st;
  if ((st = dev_alloc(E, &E.d_eig, (size_t)1))) return st;
  HIP_TRY(hipMemcpy(E.d_eig, &E.eig, sizeof(EigConst), hipMemcpyHostToDevice));
  {
    std::vector<Scalars> sc(n);
    for (auto& s : sc) {
      std::memset(&s, 0, sizeof(s));
      s.err = ~0ULL;
      s.cfl_factor = 0.05;
    }
    HIP_TRY(hipMemcpy(E.d_sc, sc.data(), sizeof(Scalars) * n, hipMemcpyHostToDevice));
  }
  std::vector<int> c_sid, c_kind, c_gid, h_sid, h_kind, h_gid, h_tab;
  std::vector<std::pair<const void*, int>> table_base;  // shared heavy table -> first row
  std::vector<long> c_pad, c_src, h_pad, h_near, h_st;
  std::vector<double> c_n, c_state, h_n, h_dist, h_state, h_of, h_d5, h_S, h_W, h_T, c_p0x, c_state2, c_front;
  long max_cells = 0, max_ghost = 1;
  for (int s = 0; s < n; ++s) {
    const rh3t::Problem& pb = *P[s];
    SolveHost& sh = E.solves[s];
    Geo& g = sh.g;
    int row0 = -1;
    for (auto& tb : table_base)
      if (tb.first == (const void*)pb.heavy.get()) row0 = tb.second;
    const bool new_tables = row0 < 0;
    if (new_tables) {
      row0 = (int)(h_d5.size() / 5);
      table_base.push_back({(const void*)pb.heavy.get(), row0});
    }
    g.nx = pb.grid.nx;
    g.ny = pb.grid.ny;
    g.PY = ((g.ny + 8 + 15) / 16) * 16;
    g.plane = (long)(g.nx + 8) * g.PY;
    g.dx = pb.grid.dx;
    g.dy = pb.grid.dy;
    g.rdx = 1.0 / g.dx;
    g.rdy = 1.0 / g.dy;
    g.r60dx = 1.0 / (60.0 * g.dx);
    g.r60dy = 1.0 / (60.0 * g.dy);
    g.cdx = -1.0 / (g.dx * g.dx);
    g.cdy = -1.0 / (g.dy * g.dy);
    const long ncell = (long)g.nx * g.ny;
    max_cells = std::max(max_cells, ncell);
    sh.fluid = pb.fluid;
    sh.nonfluid.assign(6 * (size_t)ncell, 0.0);
    sh.n_ghost = (long)pb.ghosts.size();
    max_ghost = std::max(max_ghost, sh.n_ghost);
    std::vector<uint8_t> pt(g.plane, 0);
    for (long i = 0; i < g.nx; ++i)
      for (long j = 0; j < g.ny; ++j)
        if (pb.fluid[i * g.ny + j]) pt[pad_index(g, i, j)] = 1;
    std::vector<long> gpad(pb.ghosts.size());
    for (size_t k = 0; k < pb.ghosts.size(); ++k) {
      const auto& gh = pb.ghosts[k];
      if (gh.i < -4 || gh.i >= g.nx + 4 || gh.j < -4 || gh.j >= g.ny + 4) {
        std::snprintf(E.message, sizeof(E.message), "ghost outside the halo");
        return RH3T_ERR_GEOMETRY;
      }
      gpad[k] = pad_index(g, gh.i, gh.j);
      pt[gpad[k]] = 2;
      const auto& bc = pb.bc_list(gh.boundary)[gh.side_id];
      if (gh.heavy_id < 0) {
        c_sid.push_back(s);
        c_pad.push_back(gpad[k]);
        c_src.push_back(gh.src_i >= 0 ? pad_index(g, gh.src_i, gh.src_j) : 0);
        c_kind.push_back(gh.bc_kind);
        c_gid.push_back((int)k);
        c_n.push_back(gh.project.normal.x);
        c_n.push_back(gh.project.normal.y);
        for (int f = 0; f < 6; ++f) c_state.push_back(bc.state[f]);
        const bool td = gh.bc_kind == RH3T_BC_TIME_DEPENDENT;
        c_p0x.push_back(gh.project.point.x);
        for (int f = 0; f < 6; ++f) c_state2.push_back(td ? bc.state2[f] : 0.0);
        c_front.push_back(td ? bc.speed : 0.0);
        c_front.push_back(td ? bc.front0 : 0.0);
      } else {
        const auto& hb = (*pb.heavy)[gh.heavy_id];
        h_sid.push_back(s);
        h_tab.push_back(row0 + gh.heavy_id);
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
      }
    }
    if (new_tables) {
      for (const auto& hb : *pb.heavy) {
        for (int q = 0; q < 5; ++q) h_d5.push_back(hb.d[q]);
        for (int q = 0; q < 55; ++q) {
          uint32_t l = hb.stencil[q];
          h_st.push_back(pad_index(g, l / g.ny, l % g.ny));
        }
        h_S.insert(h_S.end(), hb.s, hb.s + 979);
        h_W.insert(h_W.end(), hb.w, hb.w + 225);
        swizzle_rows(hb.s, hb.w, h_T);
      }
    }
    SolveDesc& d = E.h_desc[s];
    d.g = g;
    uint8_t* dpt;
    if ((st = upload(E, &dpt, pt))) return st;
    d.pt = dpt;
    long* dg;
    if ((st = upload(E, &dg, gpad))) return st;
    d.gpad = dg;
    for (int b = 0; b < 3; ++b) {
      if ((st = dev_alloc(E, &d.U[b], 6 * (size_t)g.plane))) return st;
      HIP_TRY(hipMemset(d.U[b], 0, 6 * (size_t)g.plane * sizeof(double)));
    }
    d.U[3] = d.U[0];
    for (int b = 0; b < 3; ++b) {
      if ((st = dev_alloc(E, &dplanes[s][b], (size_t)D_N * g.plane))) return st;
      HIP_TRY(hipMemset(dplanes[s][b], 0, (size_t)D_N * g.plane * sizeof(double)));
    }
    d.D = dplanes[s][0];
    d.Dn = dplanes[s][1];
    if ((st = dev_alloc(E, &d.FX, 6 * (size_t)(g.nx + 1) * g.ny))) return st;
    if ((st = dev_alloc(E, &d.FY, 6 * (size_t)g.nx * (g.ny + 1)))) return st;
    if ((st = dev_alloc(E, &d.RHS, 6 * (size_t)ncell))) return st;
    d.sc = E.d_sc + s;
  }
  {
    std::vector<SolveDesc> d3;
    for (int b = 0; b < 3; ++b)
      for (int s = 0; s < n; ++s) {
        SolveDesc d = E.h_desc[s];
        d.D = dplanes[s][b];
        d.Dn = dplanes[s][(b + 1) % 3];
        d3.push_back(d);
      }
    if ((st = upload(E, &E.d_desc3, d3))) return st;
    E.d_desc = E.d_desc3;
    E.dvalid = 0;
  }
  if ((st = dev_alloc(E, &E.d_aos, 6 * (size_t)max_cells))) return st;
  if ((st = dev_alloc(E, &E.d_gvals, 6 * (size_t)max_ghost))) return st;

  E.h_cheap_sid = c_sid;
  E.h_cheap_kind = c_kind;
  E.h_cheap_pad = c_pad;
  E.h_cheap_src = c_src;
  E.h_cheap_normal = c_n;
  E.cheap.n = (int)c_pad.size();
  E.heavy.n = (int)h_pad.size();
  {
    // FarField / Outflow ghosts, then Wall ghosts; within each, ghosts that
    // share a table row (the same ghost of clones of one problem) next to
    // each other, so that with xcd_tile's mapping one XCD's L2 serves the
    // row to all of them
    std::vector<int> order(E.heavy.n);
    for (int k = 0; k < E.heavy.n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      const bool wa = h_kind[a] == RH3T_BC_WALL, wb = h_kind[b] == RH3T_BC_WALL;
      if (wa != wb) return wb;
      return h_tab[a] < h_tab[b];
    });
    E.heavy.n_open = 0;
    for (int k = 0; k < E.heavy.n; ++k) E.heavy.n_open += h_kind[k] != RH3T_BC_WALL;
    auto permute = [&](auto& v, int width) {
      auto src = v;
      for (size_t k = 0; k < order.size(); ++k)
        for (int q = 0; q < width; ++q) v[k * width + q] = src[(size_t)order[k] * width + q];
    };
    permute(h_sid, 1);
    permute(h_tab, 1);
    permute(h_pad, 1);
    permute(h_kind, 1);
    permute(h_gid, 1);
    permute(h_near, 1);
    permute(h_n, 2);
    permute(h_dist, 1);
    permute(h_state, 6);
    permute(h_of, 3);
  }
  // heavy ghosts of one solve that share a nearest point and a normal share
  // a frame.  Frames are numbered in (solve, nearest slot) order, so that
  // the lanes of a k_ghost_prep wavefront read neighbouring slots of one
  // solve (the ghost order puts the clones of a ghost, in other solves, next
  // to each other).
  std::vector<int> h_frame(E.heavy.n, 0), f_sid;
  std::vector<long> f_near;
  std::vector<double> f_n;
  {
    std::map<std::tuple<int, long, uint64_t, uint64_t>, int> fid;
    for (int k = 0; k < E.heavy.n; ++k) {
      uint64_t bx, by;
      std::memcpy(&bx, &h_n[2 * (size_t)k], sizeof(bx));
      std::memcpy(&by, &h_n[2 * (size_t)k + 1], sizeof(by));
      fid.emplace(std::make_tuple(h_sid[k], h_near[k], bx, by), 0);
    }
    for (auto& kv : fid) {  // key order
      kv.second = (int)f_sid.size();
      f_sid.push_back(std::get<0>(kv.first));
      f_near.push_back(std::get<1>(kv.first));
      double nx_, ny_;
      const uint64_t bx = std::get<2>(kv.first), by = std::get<3>(kv.first);
      std::memcpy(&nx_, &bx, sizeof(nx_));
      std::memcpy(&ny_, &by, sizeof(ny_));
      f_n.push_back(nx_);
      f_n.push_back(ny_);
    }
    for (int k = 0; k < E.heavy.n; ++k) {
      uint64_t bx, by;
      std::memcpy(&bx, &h_n[2 * (size_t)k], sizeof(bx));
      std::memcpy(&by, &h_n[2 * (size_t)k + 1], sizeof(by));
      h_frame[k] = fid.at(std::make_tuple(h_sid[k], h_near[k], bx, by));
    }
  }
  // the distinct E_r points of the heavy ghosts, numbered in (solve, slot)
  // order: k_ghost_prep's lanes copy neighbouring slots of one solve
  std::vector<int> h_gpk((size_t)E.heavy.n * 55), p_sid;
  std::vector<long> p_slot;
  {
    std::map<std::pair<int, long>, int> pid;
    for (int k = 0; k < E.heavy.n; ++k)
      for (int a = 0; a < 55; ++a) pid.emplace(std::make_pair(h_sid[k], h_st[55 * (size_t)h_tab[k] + a]), 0);
    for (auto& kv : pid) {  // key order
      kv.second = (int)p_sid.size();
      p_sid.push_back(kv.first.first);
      p_slot.push_back(kv.first.second);
    }
    for (int k = 0; k < E.heavy.n; ++k)
      for (int a = 0; a < 55; ++a)
        h_gpk[55 * (size_t)k + a] = pid.at(std::make_pair(h_sid[k], h_st[55 * (size_t)h_tab[k] + a]));
  }
  E.h_heavy_kind = h_kind;
  E.h_heavy_tab = h_tab;
  E.h_heavy_pad = h_pad;
  E.h_heavy_near = h_near;
  E.h_heavy_stencil = h_st;
  long* lp;
  int* ip;
  double* dp;
#define UP(vec, dst)                          \
  if ((st = upload(E, &dst##_tmp, vec))) return st; \
  dst = dst##_tmp;
  int* ip_tmp;
  long* lp_tmp;
  double* dp_tmp;
  (void)lp;
  (void)ip;
  (void)dp;
  UP(c_sid, ip) E.cheap.sid = ip;
  UP(c_pad, lp) E.cheap.pad = lp;
  UP(c_src, lp) E.cheap.src = lp;
  UP(c_kind, ip) E.cheap.kind = ip;
  UP(c_gid, ip) E.cheap.gid = ip;
  UP(c_n, dp) E.cheap.normal = dp;
  UP(c_state, dp) E.cheap.cstate = dp;
  UP(c_p0x, dp) E.cheap.p0x = dp;
  UP(c_state2, dp) E.cheap.cstate2 = dp;
  UP(c_front, dp) E.cheap.front = dp;
  UP(h_sid, ip) E.heavy.sid = ip;
  UP(h_pad, lp) E.heavy.pad = lp;
  UP(h_kind, ip) E.heavy.kind = ip;
  UP(h_gid, ip) E.heavy.gid = ip;
  UP(h_near, lp) E.heavy.nearest = lp;
  UP(h_n, dp) E.heavy.normal = dp;
  UP(h_dist, dp) E.heavy.dist = dp;
  UP(h_state, dp) E.heavy.bcstate = dp;
  UP(h_of, dp) E.heavy.outflow = dp;
  UP(h_tab, ip) E.heavy.tab = ip;
  UP(h_d5, dp) E.heavy.d5 = dp;
  UP(h_st, lp) E.heavy.stencil = lp;
  UP(h_S, dp) E.heavy.S = dp;
  UP(h_W, dp) E.heavy.W = dp;
  UP(h_T, dp) E.heavy.T = dp;
  UP(h_frame, ip) E.heavy.frame = ip;
  E.frames.n = (int)f_sid.size();
  UP(f_sid, ip) E.frames.sid = ip;
  UP(f_near, lp) E.frames.nearest = lp;
  UP(f_n, dp) E.frames.normal = dp;
  if ((st = dev_alloc(E, &E.frames.out, (size_t)kFrameLen * (f_sid.empty() ? 1 : f_sid.size())))) return st;
  E.heavy.F = E.frames.out;
  UP(h_gpk, ip) E.heavy.gpk = ip;
  {
    std::vector<int> dord(E.heavy.n);
    for (int k = 0; k < E.heavy.n; ++k) dord[k] = k;
    std::sort(dord.begin(), dord.end(), [&](int a, int b) {
      return h_sid[a] != h_sid[b] ? h_sid[a] < h_sid[b] : h_pad[a] < h_pad[b];
    });
    UP(dord, ip) E.heavy.dord = ip;
  }
  E.packed.n = (int)p_sid.size();
  UP(p_sid, ip) E.packed.sid = ip;
  UP(p_slot, lp) E.packed.slot = lp;
  if ((st = dev_alloc(E, &E.packed.out, 6 * (p_sid.empty() ? 1 : p_sid.size())))) return st;
  E.heavy.P = E.packed.out;
  // the cheap ghosts may share the launch of the frames and packed points
  // when none of those reads a ghost point
  E.cheap_with_prep = true;
  auto fluid_slot = [&](int s, long q) {
    const Geo& g = E.solves[s].g;
    const long i = q / g.PY - 4, j = q % g.PY - 4;
    return i >= 0 && i < g.nx && j >= 0 && j < g.ny && E.solves[s].fluid[i * g.ny + j];
  };
  for (size_t e = 0; e < p_sid.size() && E.cheap_with_prep; ++e) E.cheap_with_prep = fluid_slot(p_sid[e], p_slot[e]);
  for (size_t k = 0; k < f_sid.size() && E.cheap_with_prep; ++k) E.cheap_with_prep = fluid_slot(f_sid[k], f_near[k]);
#undef UP
  if ((st = build_tables(E))) return st;
  return RH3T_OK;
}

int set_state(Engine& E, int s, const double* u, double time) {
  if (s < 0 || s >= E.nsolve || !u) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(E.device));
  SolveHost& sh = E.solves[s];
  const long ncell = (long)sh.g.nx * sh.g.ny;
  for (long l = 0; l < ncell; ++l)
    if (!sh.fluid[l])
      for (int f = 0; f < 6; ++f) sh.nonfluid[6 * l + f] = u[6 * l + f];
  HIP_TRY(hipMemcpyAsync(E.d_aos, u, 6 * ncell * sizeof(double), hipMemcpyHostToDevice, E.stream));
  k_aos_to_pad<<<(unsigned)((ncell + 255) / 256), 256, 0, E.stream>>>(E.d_aos, E.h_desc[s]);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&E.d_sc[s].time, &time, sizeof(double), hipMemcpyHostToDevice, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  E.dt_partials_valid = false;
  E.dvalid = 0;
  return RH3T_OK;
}

int get_state(Engine& E, int s, double* u, double* time) {
  if (s < 0 || s >= E.nsolve || !u) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(E.device));
  SolveHost& sh = E.solves[s];
  const long ncell = (long)sh.g.nx * sh.g.ny;
  k_pad_to_aos<<<(unsigned)((ncell + 255) / 256), 256, 0, E.stream>>>(E.h_desc[s], E.d_aos);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(u, E.d_aos, 6 * ncell * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  double t = 0.0;
  HIP_TRY(hipMemcpyAsync(&t, &E.d_sc[s].time, sizeof(double), hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  for (long l = 0; l < ncell; ++l)
    if (!sh.fluid[l])
      for (int f = 0; f < 6; ++f) u[6 * l + f] = sh.nonfluid[6 * l + f];
  if (time) *time = t;
  return RH3T_OK;
}

int advance(Engine& E, int nsteps, double cfl_factor, double dt_fixed) {
  if (nsteps < 0) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(E.device));
  // read-modify-write of the scalars on the solver's stream: CFL rule, fixed
  // dt, and a clean first-failure key
  std::vector<Scalars> sc(E.nsolve);
  HIP_TRY(hipMemcpyAsync(sc.data(), E.d_sc, sizeof(Scalars) * E.nsolve, hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  for (auto& s : sc) {
    s.cfl_factor = cfl_factor;
    s.dt_fixed = dt_fixed;
    s.err = ~0ULL;
  }
  HIP_TRY(hipMemcpyAsync(E.d_sc, sc.data(), sizeof(Scalars) * E.nsolve, hipMemcpyHostToDevice, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  if (!E.dt_partials_valid) launch_dt_partial(E);
  // the captured step starts from current Derived planes of U0 (as every
  // step leaves them), so that its first stage forms only the ghosts'
  ensure_derived(E, 0);
  // each step's dt kernel captures the previous step's failure; the last
  // step's is captured after the loop
  if (E.profiling) {
    for (int s = 0; s < nsteps; ++s) {
      launch_dt(E, 1, false);
      launch_rk3(E, false);
    }
    HIP_TRY(hipGetLastError());
  } else {
    if (!E.graph) {
      hipGraph_t gr = nullptr;
      HIP_TRY(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
      launch_dt(E, 1, false);
      launch_rk3(E, false);
      // end the capture whatever the launches reported, so that the stream
      // never stays in capture mode
      hipError_t le = hipGetLastError();
      hipError_t ce = hipStreamEndCapture(E.stream, &gr);
      hipError_t e = le != hipSuccess ? le : ce;
      if (e == hipSuccess) e = hipGraphInstantiate(&E.graph, gr, nullptr, nullptr, 0);
      if (gr) (void)hipGraphDestroy(gr);
      if (e != hipSuccess) {
        E.graph = nullptr;
        std::snprintf(E.message, sizeof(E.message), "step graph capture: %s", hipGetErrorString(e));
        return RH3T_ERR_HIP;
      }
    }
    for (int s = 0; s < nsteps; ++s) HIP_TRY(hipGraphLaunch(E.graph, E.stream));
  }
  launch_capture(E, 2, 7);
  HIP_TRY(hipGetLastError());
  // the last stage 3 left the wavefront minima of the new state
  E.dt_partials_valid = nsteps > 0 || E.dt_partials_valid;
  return check_errors(E);
}

int profile_enable(Engine& E, int enable) {
  HIP_TRY(hipSetDevice(E.device));
  HIP_TRY(hipStreamSynchronize(E.stream));
  for (auto& r : E.prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  E.prof.clear();
  for (int k = 0; k < K_COUNT; ++k) {
    E.prof_ms[k] = 0;
    E.prof_n[k] = 0;
  }
  E.profiling = enable != 0;
  return RH3T_OK;
}

int profile_read(Engine& E, int max_kernels, char* names, int names_cap, double* avg_ms, int64_t* launches,
                 int32_t* n_kernels) {
  if (!n_kernels) return RH3T_ERR_ARG;
  HIP_TRY(hipSetDevice(E.device));
  HIP_TRY(hipStreamSynchronize(E.stream));
  for (auto& r : E.prof) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
    E.prof_ms[r.kernel] += ms;
    E.prof_n[r.kernel] += 1;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  E.prof.clear();
  int n = std::min<int>(K_COUNT, max_kernels);
  std::string s;
  for (int k = 0; k < n; ++k) {
    s += kKernelNames[k];
    s.push_back('\0');
    if (avg_ms) avg_ms[k] = E.prof_n[k] ? E.prof_ms[k] / (double)E.prof_n[k] : 0.0;
    if (launches) launches[k] = E.prof_n[k];
  }
  if (names && names_cap > 0) std::memcpy(names, s.data(), std::min<int>((int)s.size(), names_cap));
  *n_kernels = n;
  return RH3T_OK;
}

bool valid_shape(int bx, int by) {
  return bx > 0 && by > 0 && bx * by <= 256 && bx % 16 == 0 && (bx * by) % 64 == 0;
}

int set_launch(Engine& E, int bx, int by, int gpb) {
  if (!valid_shape(bx, by) || gpb < 1 || gpb > kMaxGhostsPerBlock) return RH3T_ERR_ARG;
  E.bx = bx;
  E.by = by;
  E.ghosts_per_block = gpb;
  return build_tables(E);
}

// (0, 0): the x-flux kernel follows the assembly shape again
int set_launch_x(Engine& E, int bx, int by) {
  if (!(bx == 0 && by == 0) && !valid_shape(bx, by)) return RH3T_ERR_ARG;
  E.xbx = bx;
  E.xby = by;
  return build_tables(E);
}

rh3t_physics phys_or_default(const rh3t_physics* p) {
  rh3t_physics q;
  if (p)
    q = *p;
  else
    rh3t_physics_default(&q);
  return q;
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

static int check_device(int32_t device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RH3T_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return RH3T_ERR_ARG;
  return RH3T_OK;
}

int rh3t_ctx_create(const rh3t_problem* pb, const rh3t_physics* phys, int32_t device, rh3t_ctx** out) {
  if (!pb || !out || !pb->p.ghosts_built) return RH3T_ERR_ARG;
  *out = nullptr;
  int st = check_device(device);
  if (st) return st;
  auto* ctx = new rh3t_ctx();
  const rh3t::Problem* P = &pb->p;
  st = build_engine(ctx->e, &P, 1, phys_or_default(phys), device);
  if (st != RH3T_OK) {
    free_engine(ctx->e);
    delete ctx;
    return st;
  }
  *out = ctx;
  return RH3T_OK;
}

void rh3t_ctx_destroy(rh3t_ctx* ctx) {
  if (!ctx) return;
  free_engine(ctx->e);
  delete ctx;
}

int rh3t_set_state(rh3t_ctx* ctx, const double* u, double time) {
  if (!ctx) return RH3T_ERR_ARG;
  return set_state(ctx->e, 0, u, time);
}

int rh3t_get_state(rh3t_ctx* ctx, double* u, double* time) {
  if (!ctx) return RH3T_ERR_ARG;
  return get_state(ctx->e, 0, u, time);
}

int rh3t_l(rh3t_ctx* ctx, double* rhs_aos, double* ghost_vals) {
  if (!ctx || !rhs_aos) return RH3T_ERR_ARG;
  Engine& E = ctx->e;
  HIP_TRY(hipSetDevice(E.device));
  int st = reset_errors(E);
  if (st) return st;
  launch_stage(E, 0, StageSpec{0, 0, 0, 0, 2}, true);
  launch_capture(E, 2, 2);
  HIP_TRY(hipGetLastError());
  const SolveDesc& d = E.h_desc[0];
  const long ncell = (long)d.g.nx * d.g.ny;
  k_soa_to_aos<<<(unsigned)((ncell + 255) / 256), 256, 0, E.stream>>>(d.RHS, E.d_aos, ncell);
  HIP_TRY(hipMemcpyAsync(rhs_aos, E.d_aos, 6 * ncell * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  const long ng = E.solves[0].n_ghost;
  if (ghost_vals && ng > 0) {
    k_gather_ghosts<<<(unsigned)((ng + 255) / 256), 256, 0, E.stream>>>(d, (int)ng, E.d_gvals);
    HIP_TRY(hipMemcpyAsync(ghost_vals, E.d_gvals, 6 * ng * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  }
  return check_errors(E);
}

int rh3t_rk3_step(rh3t_ctx* ctx, double dt) {
  if (!ctx) return RH3T_ERR_ARG;
  if (!(dt > 0.0)) return RH3T_ERR_DT;
  Engine& E = ctx->e;
  HIP_TRY(hipSetDevice(E.device));
  int st = reset_errors(E);
  if (st) return st;
  HIP_TRY(hipMemcpyAsync(&E.d_sc[0].dt, &dt, sizeof(double), hipMemcpyHostToDevice, E.stream));
  // the step starts at Field.time (stage times t, t + dt, t + dt/2)
  HIP_TRY(hipMemcpyAsync(&E.d_sc[0].t_n, &E.d_sc[0].time, sizeof(double), hipMemcpyDeviceToDevice, E.stream));
  launch_rk3(E);
  HIP_TRY(hipGetLastError());
  E.dt_partials_valid = true;
  // Field.time advances by dt (u3.time = u.time + dt, src/main.rs:370)
  double t = 0.0;
  HIP_TRY(hipMemcpyAsync(&t, &E.d_sc[0].time, sizeof(double), hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  t = t + dt;
  HIP_TRY(hipMemcpyAsync(&E.d_sc[0].time, &t, sizeof(double), hipMemcpyHostToDevice, E.stream));
  return check_errors(E);
}

int rh3t_calc_global_dt(rh3t_ctx* ctx, double* dt_cfl) {
  if (!ctx || !dt_cfl) return RH3T_ERR_ARG;
  Engine& E = ctx->e;
  HIP_TRY(hipSetDevice(E.device));
  launch_dt(E, 0, true);
  E.dt_partials_valid = true;
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(dt_cfl, &E.d_sc[0].dt_cfl, sizeof(double), hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  return RH3T_OK;
}

int rh3t_advance(rh3t_ctx* ctx, int32_t nsteps, double cfl_factor, double dt_fixed, double* time) {
  if (!ctx) return RH3T_ERR_ARG;
  Engine& E = ctx->e;
  int st = advance(E, nsteps, cfl_factor, dt_fixed);
  if (time) {
    HIP_TRY(hipMemcpy(time, &E.d_sc[0].time, sizeof(double), hipMemcpyDeviceToHost));
  }
  return st;
}

int rh3t_ctx_set_arith(rh3t_ctx* ctx, int32_t mode) {
  if (!ctx || (mode != RH3T_ARITH_EXACT && mode != RH3T_ARITH_FAST)) return RH3T_ERR_ARG;
  return set_arith(ctx->e, mode == RH3T_ARITH_FAST);
}

int rh3t_ctx_set_physics(rh3t_ctx* ctx, const rh3t_physics* phys) {
  if (!ctx || !phys) return RH3T_ERR_ARG;
  Engine& E = ctx->e;
  HIP_TRY(hipSetDevice(E.device));
  HIP_TRY(hipStreamSynchronize(E.stream));
  return apply_physics(E, *phys);
}

int rh3t_ctx_set_launch(rh3t_ctx* ctx, int32_t bx, int32_t by, int32_t ghosts_per_block) {
  if (!ctx) return RH3T_ERR_ARG;
  return set_launch(ctx->e, bx, by, ghosts_per_block);
}

int rh3t_ctx_set_launch_x(rh3t_ctx* ctx, int32_t bx, int32_t by) {
  if (!ctx) return RH3T_ERR_ARG;
  return set_launch_x(ctx->e, bx, by);
}

int rh3t_last_error(const rh3t_ctx* ctx, rh3t_error* err) {
  if (!ctx || !err) return RH3T_ERR_ARG;
  *err = ctx->e.last.empty() ? rh3t_error{} : ctx->e.last[0];
  if (err->kind == 0 && ctx->e.message[0]) std::memcpy(err->message, ctx->e.message, sizeof(err->message));
  return RH3T_OK;
}

int rh3t_profile(rh3t_ctx* ctx, int32_t enable) {
  if (!ctx) return RH3T_ERR_ARG;
  return profile_enable(ctx->e, enable);
}

int rh3t_profile_read(rh3t_ctx* ctx, int32_t max_kernels, char* names, int32_t names_cap, double* avg_ms,
                      int64_t* launches, int32_t* n_kernels) {
  if (!ctx) return RH3T_ERR_ARG;
  return profile_read(ctx->e, max_kernels, names, names_cap, avg_ms, launches, n_kernels);
}

// ------------------------------------------------------------- batch

int rh3t_batch_create(const rh3t_problem* const* problems, int32_t n, const rh3t_physics* phys, int32_t device,
                      rh3t_batch** out) {
  if (!problems || n <= 0 || !out) return RH3T_ERR_ARG;
  for (int s = 0; s < n; ++s)
    if (!problems[s] || !problems[s]->p.ghosts_built) return RH3T_ERR_ARG;
  *out = nullptr;
  int st = check_device(device);
  if (st) return st;
  std::vector<const rh3t::Problem*> P(n);
  for (int s = 0; s < n; ++s) P[s] = &problems[s]->p;
  auto* b = new rh3t_batch();
  st = build_engine(b->e, P.data(), n, phys_or_default(phys), device);
  if (st != RH3T_OK) {
    free_engine(b->e);
    delete b;
    return st;
  }
  *out = b;
  return RH3T_OK;
}

void rh3t_batch_destroy(rh3t_batch* b) {
  if (!b) return;
  free_engine(b->e);
  delete b;
}

int rh3t_batch_set_state(rh3t_batch* b, int32_t solve, const double* u, double time) {
  if (!b) return RH3T_ERR_ARG;
  return set_state(b->e, solve, u, time);
}

int rh3t_batch_get_state(rh3t_batch* b, int32_t solve, double* u, double* time) {
  if (!b) return RH3T_ERR_ARG;
  return get_state(b->e, solve, u, time);
}

int rh3t_batch_advance(rh3t_batch* b, int32_t nsteps, double cfl_factor, double* times) {
  if (!b) return RH3T_ERR_ARG;
  Engine& E = b->e;
  int st = advance(E, nsteps, cfl_factor, 0.0);
  if (times) {
    std::vector<Scalars> sc(E.nsolve);
    HIP_TRY(hipMemcpy(sc.data(), E.d_sc, sizeof(Scalars) * E.nsolve, hipMemcpyDeviceToHost));
    for (int s = 0; s < E.nsolve; ++s) times[s] = sc[s].time;
  }
  return st;
}

int rh3t_batch_set_arith(rh3t_batch* b, int32_t mode) {
  if (!b || (mode != RH3T_ARITH_EXACT && mode != RH3T_ARITH_FAST)) return RH3T_ERR_ARG;
  return set_arith(b->e, mode == RH3T_ARITH_FAST);
}

int rh3t_batch_set_launch(rh3t_batch* b, int32_t bx, int32_t by, int32_t ghosts_per_block) {
  if (!b) return RH3T_ERR_ARG;
  return set_launch(b->e, bx, by, ghosts_per_block);
}

int rh3t_batch_set_launch_x(rh3t_batch* b, int32_t bx, int32_t by) {
  if (!b) return RH3T_ERR_ARG;
  return set_launch_x(b->e, bx, by);
}

int rh3t_batch_last_error(const rh3t_batch* b, int32_t solve, rh3t_error* err) {
  if (!b || !err || solve < 0 || solve >= b->e.nsolve) return RH3T_ERR_ARG;
  *err = b->e.last.empty() ? rh3t_error{} : b->e.last[solve];
  return RH3T_OK;
}

int rh3t_batch_profile(rh3t_batch* b, int32_t enable) {
  if (!b) return RH3T_ERR_ARG;
  return profile_enable(b->e, enable);
}

int rh3t_batch_profile_read(rh3t_batch* b, int32_t max_kernels, char* names, int32_t names_cap, double* avg_ms,
                            int64_t* launches, int32_t* n_kernels) {
  if (!b) return RH3T_ERR_ARG;
  return profile_read(b->e, max_kernels, names, names_cap, avg_ms, launches, n_kernels);
}

}  // extern "C"
