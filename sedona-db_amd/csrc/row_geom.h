// row_geom.h -- device access to geometry rows and the reference's
// point-to-geometry distance, shared by the distance refine and the KNN
// join.
//
// Distance semantics: geo's Euclidean distance from a point to a point,
// (multi)linestring or (multi)polygon (rust/sedona-geo-generic-alg/src/
// algorithm/line_measures/metric_spaces/euclidean/utils.rs:96-271): zero on
// a line or inside a polygon (and not in a hole) by exact orientation
// signs, else the least distance_coord_to_line over the segments (holes
// included), minimum over the parts.
#pragma once

#include "geom_math.h"
#include "sjhip.h"

namespace sjhip {

// One ring (or line run) of a geometry column.
struct RingRef {
  const double2* xy;  // ring points (double)
  int64_t v0;         // global index of the first point
  int64_t npts;
};

__device__ __forceinline__ RingRef ring_at(const GeomView& g, int64_t ring) {
  RingRef r;
  r.v0 = g.pt_off[ring];
  r.npts = g.pt_off[ring + 1] - r.v0;
  r.xy = g.xy + r.v0;
  return r;
}

// Exact point position w.r.t. a closed ring: 2 boundary, 1 inside, 0 outside.
__device__ inline int exact_ring_position(const RingRef& r, double px, double py) {
  int inside = 0;
  for (int64_t i = 0; i < r.npts; i++) {
    int64_t j = i + 1 == r.npts ? 0 : i + 1;
    double2 a = r.xy[i], b = r.xy[j];
    if (line_intersects_coord(a.x, a.y, b.x, b.y, px, py)) return 2;
    if ((a.y > py) != (b.y > py)) {
      int o = orient2d_sign(a.x, a.y, b.x, b.y, px, py);
      if ((b.y > a.y) ? (o > 0) : (o < 0)) inside ^= 1;
    }
  }
  return inside;
}

__device__ inline double point_to_run(const RingRef& r, double px, double py) {
  // utils.rs:184-210 (open run of points)
  if (r.npts < 2) return 0.0;
  for (int64_t i = 0; i + 1 < r.npts; i++) {
    double2 a = r.xy[i], b = r.xy[i + 1];
    if (line_intersects_coord(a.x, a.y, b.x, b.y, px, py)) return 0.0;
  }
  double2 a = r.xy[0], b = r.xy[1];
  double m = coord_to_line(px, py, a.x, a.y, b.x, b.y);
  for (int64_t i = 1; i + 1 < r.npts; i++) {
    a = r.xy[i];
    b = r.xy[i + 1];
    m = rmin(m, coord_to_line(px, py, a.x, a.y, b.x, b.y));
  }
  return m;
}

__device__ inline double point_to_polygon(const GeomView& g, int64_t part, double px, double py) {
  int64_t r0 = g.ring_off[part], r1 = g.ring_off[part + 1];
  if (r0 == r1) return 0.0;
  RingRef ext = ring_at(g, r0);
  int st = exact_ring_position(ext, px, py);
  if (st == 2) return 0.0;
  if (st == 1) {
    bool in_hole = false;
    for (int64_t h = r0 + 1; h < r1; h++) {
      int hs = exact_ring_position(ring_at(g, h), px, py);
      if (hs == 2) return 0.0;
      if (hs == 1) in_hole = true;
    }
    if (!in_hole) return 0.0;
  }
  double m = INFINITY;
  for (int64_t k = r0; k < r1; k++) {
    RingRef r = ring_at(g, k);
    for (int64_t i = 0; i + 1 < r.npts; i++) {
      double2 a = r.xy[i], b = r.xy[i + 1];
      m = rmin(m, coord_to_line(px, py, a.x, a.y, b.x, b.y));
    }
  }
  return m;
}

__device__ inline double distance_to_point(const GeomView& g, int64_t row, double px, double py) {
  int t = g.type[row];
  if (g.points_only) {
    double2 q = g.xy[row];
    return hypot_ref(q.x - px, q.y - py);
  }
  double m = INFINITY;
  for (int64_t q = g.part_off[row]; q < g.part_off[row + 1]; q++) {
    int64_t r0 = g.ring_off[q];
    RingRef r = ring_at(g, r0);
    double d;
    if (t == kPoint || t == kMultiPoint) d = hypot_ref(r.xy[0].x - px, r.xy[0].y - py);
    else if (t == kLine || t == kMultiLine) d = point_to_run(r, px, py);
    else d = point_to_polygon(g, q, px, py);
    m = rmin(m, d);
  }
  return m;
}

}  // namespace sjhip
