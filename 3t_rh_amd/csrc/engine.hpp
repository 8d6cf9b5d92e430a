// Device context shared by the forward engine (engine.hip) and the adjoint
// (adjoint.hip): per-solve descriptors, the ghost work lists, the Engine
// that owns a batch's device memory and HIP stream, and the launch helpers
// of one l() evaluation / RK3 stage.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "flux.hpp"
#include "kernels.hpp"
#include "rh3t.h"

namespace rh3t_eng {

using namespace rh3t_dev;

struct SolveDesc {
  Geo g;
  double* U[4];         // U0, U1, U2; U[3]: stage-3 destination (U0 itself
                        // except where a step writes a new buffer)
  double* D;            // Derived planes of the state a kernel reads (on an
                        // adjoint tape every step and stage state has its
                        // own descriptor and planes)
  double* Dn;           // Derived planes of the state a stage writes, filled
                        // by its assembly when StageSpec::wd > 0
  double* FX;
  double* FY;
  double* RHS;          // compact SoA nx*ny (l() output)
  const uint8_t* pt;    // 1 fluid, 2 ghost
  Scalars* sc;
  const long* gpad;     // ghost slots by sorted ghost id
};

struct CheapGhosts {
  int n;
  const int* sid;       // solve
  const long* pad;      // destination slot
  const long* src;      // source slot (reflective / zeroth / periodic)
  const int* kind;      // rh3t_bc_kind
  const double* normal; // 2 per ghost
  const double* cstate; // 6 per ghost (Constant; TimeDependent behind the front)
  const int* gid;       // sorted ghost id within its solve
  // TimeDependent (RH3T_TD_MOVING_FRONT) per ghost: P0.x, the state ahead of
  // the front (6), front speed and position at t = 0 (2); zero otherwise
  const double* p0x;
  const double* cstate2;
  const double* front;
};

// Which time a ghost update evaluates TimeDependent ghosts at: the state's
// own Field.time (an l() call), or stage 1..3 of the current RK3 step.
enum GhostTime { GT_FIELD = 0, GT_STAGE1 = 1, GT_STAGE2 = 2, GT_STAGE3 = 3 };

struct HeavyGhosts {
  int n;
  int n_open;            // FarField / Outflow ghosts first, Wall ghosts after
  const int* sid;        // solve
  const long* pad;       // destination slot
  const int* kind;       // Wall / Outflow / FarField
  const int* gid;        // sorted ghost id within its solve
  const long* nearest;   // slot of GhostBC::nearest
  const double* normal;  // 2 per ghost
  const double* dist;    // signed distance D
  const double* bcstate; // 6 per ghost (FarField u_inf)
  const double* outflow; // 3 per ghost (p_inf, sigma, l_domain)
  const int* tab;        // GhostBC table row of the ghost (shared by clones)
  const double* d5;      // 5 per table row
  const long* stencil;   // 55 slots per table row
  const double* S;       // 979 per table row
  const double* W;       // 225 per table row
  const double* T;       // kSwzLen per table row: S and W swizzled for k_ghost_open
  const int* frame;      // FarField / Outflow: the ghost's frame (GhostFrames)
  const double* F;       // kFrameLen per frame, formed by k_ghost_prep each update
  const int* gpk;        // FarField / Outflow: 55 per ghost, PackedPoints index of each E_r point
  const double* P;       // PackedPoints values, formed by k_ghost_prep each update
  const int* dord;       // the ghosts in (solve, slot) order, for k_derived_ghosts
};

// The E_r points read by the FarField / Outflow ghosts, each distinct
// (solve, slot) once, their 6 state values contiguous (48 bytes, one cache
// line at most two) instead of in 6 planes: k_ghost_prep copies them out of
// the state buffer before k_ghost_open gathers them.
struct PackedPoints {
  int n;
  const int* sid;
  const long* slot;
  double* out;  // 6 per point
};

// The frames of the FarField / Outflow ghosts: ghosts of one solve that
// share a nearest point and a normal share the eigensystem of
// build_local_plain (src/bc1.rs:294-312), which k_ghost_prep forms once
// per frame: rows of L (36), rows of R (36), eigenvalues (6).
constexpr int kFrameLen = 78;
struct GhostFrames {
  int n;
  const int* sid;
  const long* nearest;   // slot of the nearest point
  const double* normal;  // 2 per frame
  double* out;           // kFrameLen per frame
};

// The swizzled row block of a heavy-ghost table (FarField / Outflow kernel):
// entry (s, l) is what lane l of a ghost's lane group multiplies at step s,
// 0.0 where it has none.  Steps go in pairs, pair-major, lane-minor, the two
// steps of a pair adjacent ((s / 2) * 64 + 2 l + s % 2): a lane reads two
// steps with one 16-byte load, the group 512 consecutive bytes.
//   steps 0..24 (segment A, padded to 26):  lanes 0..24 row a = l of S_4,
//                 25..29 row k = l - 25 of w_4 (entry b = s); lane 30 the one
//                 entry of S_0, 31 of w_0 (s = 0)
//   steps 26..41 (segment B), b = s - 26: lanes 0..15 row l of S_3, 16..19
//                 row l - 16 of w_3 (b < 16); lanes 20..28 row l - 20 of S_2,
//                 29..31 row l - 29 of w_2 (b < 9); then lanes 20..23 row
//                 l - 20 of S_1, 24..25 row l - 24 of w_1 (entry b - 9,
//                 9 <= b < 13)
constexpr int kSwzA = 25, kSwzB = 16;
constexpr int kSwzPairsA = (kSwzA + 1) / 2, kSwzPairsB = kSwzB / 2;
constexpr int kSwzLen = (kSwzPairsA + kSwzPairsB) * 64;

struct StageSpec {
  int mode;        // 0: write rhs, 1..3: SSP-RK3 stage
  int base, add, dst;  // buffer indices into SolveDesc::U
  unsigned phase;
  int wd = 0;      // Derived planes the assembly writes for the new state (0: none)
};

enum KernelId {
  K_GHOST_CHEAP, K_GHOST_HEAVY, K_DERIVED, K_FLUX_X, K_FLUX_Y, K_ASSEMBLE, K_DT_PARTIAL, K_DT_FINAL, K_CAPTURE,
  K_ADJ_IFACE, K_ADJ_GATHER, K_ADJ_GHOST, K_ADJ_PARAM, K_FY_ASSEMBLE, K_GHOST_PREP, K_COUNT
};
extern const char* const kKernelNames[K_COUNT];

struct ProfRec {
  int kernel;
  hipEvent_t a, b;
};

struct SolveHost {
  Geo g;
  long n_ghost = 0;
  std::vector<uint8_t> fluid;
  std::vector<double> nonfluid;  // AoS values of non-fluid cells (kept as given)
};

struct AdjWork;  // adjoint workspace (adjoint.hip)

struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  rh3t_physics src_phys{};  // the constants as the caller gave them
  Phys phys{};
  EigConst eig{};
  EigConst* d_eig = nullptr;  // device copy for kernels that read it through a pointer
  bool fast = false;  // flux arithmetic: EXACT unless rh3t_*_set_arith selects FAST
  int nplanes = 8;
  int nsolve = 0;
  std::vector<SolveHost> solves;
  std::vector<SolveDesc> h_desc;
  SolveDesc* d_desc = nullptr;
  // Derived planes kept per state buffer (the plain and batched forward):
  // descriptor array k (d_desc3 + k * nsolve) reads the planes of U[k] and
  // its stage writes those of U[(k + 1) % 3], the next RK3 stage state;
  // array 0 is d_desc.  Bit k of dvalid: the fluid cells' planes of U[k]
  // hold Derived of its current state.
  SolveDesc* d_desc3 = nullptr;
  unsigned dvalid = 0;
  Scalars* d_sc = nullptr;
  CheapGhosts cheap{};
  HeavyGhosts heavy{};
  GhostFrames frames{};
  PackedPoints packed{};
  // host copies of the cheap-ghost lists (the adjoint transposes them)
  std::vector<int> h_cheap_sid, h_cheap_kind;
  std::vector<long> h_cheap_pad, h_cheap_src;
  std::vector<double> h_cheap_normal;
  // host copies of the heavy-ghost lists in launch order, and the stencil
  // slots of every table row (the adjoint transposes the heavy ghosts)
  std::vector<int> h_heavy_kind, h_heavy_tab;
  std::vector<long> h_heavy_pad, h_heavy_near, h_heavy_stencil;

  // launch shapes (autotuned): flux / assemble workgroup (bx, by), heavy
  // ghosts per workgroup
  // (64, 1) by default: a one-wavefront workgroup starts as soon as one SIMD
  // has a free slot.  Measured against (64, 4): 2048^2 box FAST 2.171 /
  // 2.336 ms per step, EXACT 3.675 / 3.658; cylinder 121x481 EXACT 0.1825 /
  // 0.1871; double Mach h = 1/320 FAST 0.681 / 0.727, EXACT 1.050 / 1.060.
  int bx = 64, by = 1, ghosts_per_block = 1;
  // the x-flux kernel's own workgroup shape (0: the same as (bx, by))
  int xbx = 0, xby = 0;
  int fx_bx() const { return xbx ? xbx : bx; }
  int fx_by() const { return xby ? xby : by; }
  std::vector<int> t_fx, t_fy, t_as, t_der, t_fa;  // host prefix tables
  int* d_tabs = nullptr;  // device copy: fx | fy | as | der, each nsolve+1
  double* d_partial = nullptr;  // get_local_dt minimum per assembly workgroup
  long partial_cap = 0;
  bool dt_partials_valid = false;  // d_partial holds the minima of the current U0
  bool cheap_with_prep = false;    // every frame and packed point is a fluid cell
  double* d_aos = nullptr;   // staging (largest solve)
  double* d_gvals = nullptr;
  std::vector<void*> allocs;

  hipGraphExec_t graph = nullptr;
  bool profiling = false;
  std::vector<ProfRec> prof;
  double prof_ms[K_COUNT] = {0};
  long long prof_n[K_COUNT] = {0};

  std::vector<rh3t_error> last;  // per solve
  char message[256] = {0};

  AdjWork* adj = nullptr;

  const int* tab(int k) const { return d_tabs + k * (nsolve + 1); }
};

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      std::snprintf(E.message, sizeof(E.message), "%s: %s", #expr, hipGetErrorString(e_)); \
      return RH3T_ERR_HIP;                                                                \
    }                                                                                     \
  } while (0)

template <class T>
int dev_alloc(Engine& E, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) {
    std::snprintf(E.message, sizeof(E.message), "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
    return RH3T_ERR_HIP;
  }
  E.allocs.push_back(q);
  *p = (T*)q;
  return RH3T_OK;
}

template <class T>
int upload(Engine& E, T** p, const std::vector<T>& v) {
  int st = dev_alloc(E, p, v.size());
  if (st) return st;
  if (!v.empty()) HIP_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return RH3T_OK;
}

// Brackets one launch with events on the engine's stream when profiling.
struct Launcher {
  Engine& E;
  int kid;
  hipEvent_t a = nullptr, b = nullptr;
  Launcher(Engine& e, int k) : E(e), kid(k) {
    if (E.profiling) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, E.stream);
    }
  }
  ~Launcher() {
    if (E.profiling) {
      (void)hipEventRecord(b, E.stream);
      E.prof.push_back(ProfRec{kid, a, b});
    }
  }
};

inline int total(const std::vector<int>& t) { return t.back(); }

// ghosts of every solve on buffer buf (a failure raises the solve's
// first-failure key), with TimeDependent ghosts at time gt
void launch_ghosts(Engine& E, int buf, unsigned phase, int gt);
// Derived::from_state on every fluid cell and ghost of buffer buf
void launch_derived(Engine& E, int buf);
// ghosts -> derived -> x/y fluxes -> assemble (one l() + stage) on buffer src,
// with the descriptors in E.d_desc (the adjoint swaps in its own).  kept:
// the engine's per-buffer Derived planes (d_desc3): Derived is formed only
// where the planes of U[src] are not already current (the ghost slots, or
// every point after a host write), and a stage's assembly fills the planes
// of its new state.
void launch_stage(Engine& E, int src, const StageSpec& sa, bool kept = false);
// the failing point's state of the current step's first failure, if its
// phase is in [lo, hi], into the solves' scalars (after the stages of those
// phases; a recompute of stages that already ran without error skips it)
void launch_capture(Engine& E, unsigned lo, unsigned hi);
// one SSP-RK3 step with the dt held in the solves' scalars, then (capture)
// the capture
void launch_rk3(Engine& E, bool capture = true);
int check_errors(Engine& E);
int reset_errors(Engine& E);
// releases the adjoint workspace (host side; device memory is in allocs)
void free_adjoint(Engine& E);
// replaces the physics constants (host structs, the device copy of the
// eigensystem constants, the Derived plane count); drops the step graph and
// the dt minima of the old constants.  The stream must be idle.
int apply_physics(Engine& E, const rh3t_physics& p);
// the device constants (gamma - 1, gates, ...) of a caller's physics struct
Phys physics_of(const rh3t_physics& p);
// selects the flux / assembly arithmetic (drops the step graph on change)
int set_arith(Engine& E, bool fast);

}  // namespace rh3t_eng

struct rh3t_ctx {
  rh3t_eng::Engine e;
};

struct rh3t_batch {
  rh3t_eng::Engine e;
};
