// Host-side mirror of the reference's domain model: GridInfo / Field
// (src/field1.rs), Polygon / CircularArc / ArcOverride (src/geometry.rs),
// BCType / GhostBC (src/bc1.rs) and GhostGrid (src/ghost.rs).  Everything
// here is build-time bookkeeping; its integer decisions (fluid mask, ghost
// set, ghost order, side choice, stencil indices) must match the reference
// bit for bit, so every floating-point expression keeps the reference's
// operation order and the file is compiled with -ffp-contract=off.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "rh3t.h"

namespace rh3t {

struct Point { double x, y; };
struct Vec2 { double x, y; };

struct Projection {
  Point point;    // P0
  Vec2 normal;    // fluid-domain outward normal
  double distance;  // D = (P - P0) . n
};

enum class FluidSide { Inside, Outside };

struct Polygon {
  std::vector<Point> points;
  FluidSide fluid = FluidSide::Inside;

  double signed_area2() const;
  bool is_fluid(Point p) const;
  Point closest_point(Point p) const;
  Vec2 normal(Point p0) const;
  Vec2 outward_normal_of_side(size_t side) const;
};

struct CircularArc {
  Point center;
  double radius = 0, theta_start = 0, sweep = 0;
  FluidSide fluid = FluidSide::Inside;

  static CircularArc from_three_points(Point s, Point m, Point e, FluidSide f);
  bool contains_angle(double theta) const;
  Point start_point() const;
  Point end_point() const;
  Point closest_point(Point p) const;
  Vec2 outward_normal(Point p0) const;
  Projection project(Point p) const;
};

struct ArcOverride {
  size_t side_start, side_end;
  CircularArc arc;
  bool contains_side(size_t s) const { return s >= side_start && s <= side_end; }
};

Projection project(const Polygon& poly, Point p);
std::vector<size_t> find_boundary_sides(Point p, const Polygon& poly, double eps);

// Exception type for build-time failures; converted to rh3t_status at the
// C boundary (the reference panics in the same places).
struct BuildError {
  int status;
  std::string message;
};

struct GhostBC {  // src/bc1.rs:49-57
  int32_t nearest_i, nearest_j;
  Vec2 normal;
  double d[5];
  uint32_t stencil[55];
  double s[979];
  double w[225];
};

struct GhostInfo {  // src/ghost.rs:18-28 (+ static cheap-BC source)
  int32_t i, j;
  Projection project;
  int32_t nearest_i, nearest_j;
  int32_t boundary;  // 0 outer, 1 inner
  int32_t side_id;
  int32_t bc_kind;
  int32_t heavy_id;  // index into heavy[], -1 if none
  int32_t src_i, src_j;
};

constexpr int kStencilOff[6] = {0, 1, 5, 14, 30, 55};
constexpr int kSOff[6] = {0, 1, 17, 98, 354, 979};
constexpr int kWOff[6] = {0, 1, 9, 36, 100, 225};

struct Problem {
  rh3t_grid grid{};
  Polygon outer, inner;
  std::vector<rh3t_bc> bc_outer, bc_inner;
  std::vector<ArcOverride> outer_arcs, inner_arcs;
  std::vector<uint8_t> fluid;  // i*ny + j

  bool ghosts_built = false;
  std::vector<GhostInfo> ghosts;  // sorted by (i, j)
  // Heavy-ghost tables depend only on geometry and grid, never on BC
  // states: clones (e.g. the same cylinder at another Mach number) share
  // them, on the host and on the device.
  std::shared_ptr<std::vector<GhostBC>> heavy = std::make_shared<std::vector<GhostBC>>();

  // GridInfo helpers (src/field1.rs:43-73)
  bool grid_in_domain(int64_t i, int64_t j) const {
    return i >= 0 && i < grid.nx && j >= 0 && j < grid.ny;
  }
  double x(int64_t i) const { return grid.x0 + (double)i * grid.dx; }
  double y(int64_t j) const { return grid.y0 + (double)j * grid.dy; }
  void coord2idx(Point p, int64_t* i, int64_t* j) const;
  // Field::is_in_domain (src/field1.rs:179-192)
  bool is_in_domain(int64_t i, int64_t j) const;
  const std::vector<rh3t_bc>& bc_list(int boundary) const {
    return boundary == 0 ? bc_outer : bc_inner;
  }
  const CircularArc* arc_for_side(int boundary, size_t side) const;

  void build_fluid_mask();
  void build_ghosts();
};

// Exposed for tests through the C ABI and for the device context.
void discover_ghost_indices(const Problem& pb,
                            std::vector<std::pair<int64_t, int64_t>>* out);

}  // namespace rh3t

struct rh3t_problem {
  rh3t::Problem p;
};
