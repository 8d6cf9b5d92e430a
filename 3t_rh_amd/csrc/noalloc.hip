// calc_global_dt (src/main.rs:375-408, src/dt.rs:5-34) of a state that the
// caller already holds in device memory, with no allocation: the per-
// workgroup minima live in a fixed file-scope device array instead of a
// context's buffer.  For callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops).
//
// Two kernels on the caller's stream: k_na_partial writes one fmin of
// get_local_dt per workgroup (kNaCellsPerGroup cells) into the scratch array,
// k_na_final reduces those into the caller's output.  Every call on a GPU
// shares that GPU's array; two calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "engine.hpp"

using namespace rh3t_dev;

namespace {

constexpr int kNaLanes = 256;
constexpr int kNaCellsPerLane = 16;
constexpr long kNaCellsPerGroup = (long)kNaLanes * kNaCellsPerLane;
constexpr int kNaMaxGroups = RH3T_NOALLOC_MAX_CELLS / (int)kNaCellsPerGroup;
static_assert((long)kNaMaxGroups * kNaCellsPerGroup == RH3T_NOALLOC_MAX_CELLS, "capacity");

// one partial minimum per workgroup of k_na_partial
static __device__ __align__(256) double scratch[kNaMaxGroups];

// Device address of `scratch` on each device, looked up once per device
// with that device current (the array's name is a host-side stand-in).
constexpr int kNaMaxDevices = 64;
double* g_scratch[kNaMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

int scratch_on(int device, double** out) {
  if (device < 0 || device >= kNaMaxDevices) return RH3T_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RH3T_ERR_HIP;
    g_scratch[device] = (double*)p;
  }
  *out = g_scratch[device];
  return RH3T_OK;
}

__device__ __forceinline__ double wave_min(double m) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
  return m;
}

// get_local_dt of every fluid cell of workgroup b's cells, reduced with
// fmin (exact and NaN-ignoring like f64::min, so the order is free).
__global__ __launch_bounds__(kNaLanes) void k_na_partial(const double* __restrict__ u, const uint8_t* __restrict__ fluid,
                                                         long ncell, Phys ph, double dx, double dy,
                                                         double* __restrict__ partial) {
  __shared__ double s_min[kNaLanes / 64];
  double m = INFINITY;
  const long base = (long)blockIdx.x * kNaCellsPerGroup + threadIdx.x;
#pragma unroll 4
  for (int k = 0; k < kNaCellsPerLane; ++k) {
    const long l = base + (long)k * kNaLanes;
    if (l < ncell && fluid[l]) {
      double st[6];
#pragma unroll
      for (int f = 0; f < 6; ++f) st[f] = u[6 * l + f];
      m = fmin(m, local_dt(ph, st, dx, dy));
    }
  }
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s_min[0];
    for (int w = 1; w < kNaLanes / 64; ++w) r = fmin(r, s_min[w]);
    partial[blockIdx.x] = r;
  }
}

// fmin of the first n partials into *out (one workgroup)
__global__ __launch_bounds__(1024) void k_na_final(const double* __restrict__ partial, int n,
                                                   double* __restrict__ out) {
  __shared__ double red[16];
  double m = INFINITY;
  for (int k = threadIdx.x; k < n; k += 1024) m = fmin(m, partial[k]);
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < 16; ++w) r = fmin(r, red[w]);
    *out = r;
  }
}

}  // namespace

extern "C" int rh3t_calc_global_dt_noalloc(const double* u_aos, const uint8_t* fluid, const rh3t_grid* grid,
                                           const rh3t_physics* phys, double* dt_cfl, void* stream) {
  if (!u_aos || !fluid || !grid || !dt_cfl || grid->nx <= 0 || grid->ny <= 0) return RH3T_ERR_ARG;
  const long ncell = (long)grid->nx * grid->ny;
  if (ncell > RH3T_NOALLOC_MAX_CELLS) return RH3T_ERR_SIZE;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return RH3T_ERR_NO_DEVICE;
  double* partial = nullptr;
  int st = scratch_on(device, &partial);
  if (st) return st;
  rh3t_physics p;
  if (phys)
    p = *phys;
  else
    rh3t_physics_default(&p);
  const Phys ph = rh3t_eng::physics_of(p);
  const int groups = (int)((ncell + kNaCellsPerGroup - 1) / kNaCellsPerGroup);
  hipStream_t s = (hipStream_t)stream;
  k_na_partial<<<groups, kNaLanes, 0, s>>>(u_aos, fluid, ncell, ph, grid->dx, grid->dy, partial);
  k_na_final<<<1, 1024, 0, s>>>(partial, groups, dt_cfl);
  return hipGetLastError() == hipSuccess ? RH3T_OK : RH3T_ERR_HIP;
}
