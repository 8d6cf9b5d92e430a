// Device-side layout shared by the HIP kernels and the context code.
//
// Every per-point field lives in a halo-padded, structure-of-arrays plane:
// grid point (i, j), i in [-4, nx+4), j in [-4, ny+4), sits at
//     pad(i, j) = (i + 4) * PY + (j + 4),     PY = pitch >= ny + 8
// with y (j) fastest, like the reference's l = i*ny + j.  Fluid cells and
// cached ghosts share one plane: the reference's access-table target of a
// stencil point depends only on the point (fluid index or ghost id), so the
// 17-slot table of src/ghost.rs:68-107 collapses to plain address
// arithmetic on this layout.
#pragma once

#include <cstdint>

#include "device_math.hpp"

namespace rh3t_dev {

struct Geo {
  int nx, ny;
  int PY;        // pitch of a padded row
  long plane;    // (nx + 8) * PY
  double dx, dy;
  // host-rounded reciprocals (the same IEEE quotients a kernel would form):
  // 1/dx, 1/dy, 1/(60 dx), 1/(60 dy), -1/(dx dx), -1/(dy dy)
  double rdx, rdy, r60dx, r60dy, cdx, cdy;
};

__host__ __device__ inline long pad_index(const Geo& g, long i, long j) {
  return (i + 4) * (long)g.PY + (j + 4);
}

// Element at byte offset boff of a plane in global memory: a uniform
// (scalar) plane base plus a 32-bit per-lane offset, so a stencil's addresses
// cost one VGPR per point instead of a 64-bit flat address per value (planes
// are far below 4 GiB).
typedef __attribute__((address_space(1))) const char gchar;
typedef __attribute__((address_space(1))) const double gdouble;
__device__ __forceinline__ double plane_at(const double* base, unsigned boff) {
  return *(gdouble*)((gchar*)base + boff);
}

// Error key: step (20 bits) | phase (4) | kind (8) | index (32); the
// smallest key of a call is its first failure in program order.
// Phases: 1 dt, 2/3 stage-1 ghosts/update, 4/5 stage 2, 6/7 stage 3.
__host__ __device__ inline unsigned long long err_key(unsigned step, unsigned phase, unsigned kind,
                                                      unsigned idx) {
  return ((unsigned long long)step << 44) | ((unsigned long long)phase << 40) |
         ((unsigned long long)kind << 32) | (unsigned long long)idx;
}

struct Scalars {     // device-resident time-loop scalars
  double dt;         // dt of the current step
  double dt_cfl;     // calc_global_dt of the current step
  double time;       // Field.time
  double t_n;        // Field.time at the start of the current RK3 step (the
                     // stage times are t_n, t_n + dt, t_n + dt/2)
  double cfl_factor; // dt = cfl_factor * dt_cfl  (0.05 in main.rs:461)
  double dt_fixed;   // if > 0, used instead of the CFL rule
  unsigned step;     // step counter (error ordering)
  unsigned pad_;
  unsigned long long err;  // first-failure key, ~0 when clean
  double err_state[6];
};

}  // namespace rh3t_dev
