// refine.hip -- refine kernels over candidate pairs.
//
// Relation predicates follow the reference CPU refine (TG):
//   rust/sedona-spatial-join/src/refine/tg.rs:322-350 evaluates
//   Op(build, probe) with tg_geom_{intersects,contains,within,covers,
//   coveredby}; for (multi)polygon x point these reduce to
//   poly_contains_point (c/sedona-tg/src/tg/tg.c:3909-3955) over
//   tg_ring_contains_point (2809-2822), i.e. the ray classification of
//   every ring segment (raycast, 1199-1275): any ON segment decides
//   allow_on_edge, otherwise the parity of IN segments decides.
//   Polygon x polygon intersects follows tg.c:4100-4136 / 3166-3190 /
//   3089-3111.
// The distance predicate follows the reference's Geo refiner
//   (rust/sedona-spatial-join/src/refine/geo.rs:312-335) over
//   rust/sedona-geo-generic-alg/.../euclidean/utils.rs:127-271.
//
// Two precisions for the point-in-polygon refine:
//   precision 1: every segment in FP64 with the reference's exact
//                expression sequence (the plain kernel);
//   precision 0: every segment first in FP32 ring-local coordinates with a
//                forward error bound; only segments whose outcome the bound
//                cannot certify are re-evaluated in FP64 exactly as above.
// Both give the reference's decisions bit for bit.
//
// Work split: candidates whose polygon has at most kLongRing segments are
// refined one per lane; longer ones one per wavefront, with the ring's
// segments split across the 64 lanes and the ON / IN-parity results
// combined with wavefront ballots.
#include <hipcub/hipcub.hpp>

#include "geom_math.h"
#include "lbvh_filter.h"
#include "sjhip.h"
#include "row_geom.h"
#include "wave_out.h"
#include "ring_walk.h"

namespace sjhip {

namespace {

constexpr int kBlock = 256;
constexpr int kLongRing = 192;

enum PredCode : int {
  kEquals = 0, kDisjoint = 1, kTouches = 2, kContains = 3, kCovers = 4, kIntersects = 5, kWithin = 6,
  kCoveredBy = 7, kDWithin = 100
};

__device__ __forceinline__ bool is_polyish(int t) { return t == kPolygon || t == kMultiPolygon; }

__device__ __forceinline__ int64_t n_parts(const GeomView& g, int64_t row) {
  return g.part_off[row + 1] - g.part_off[row];
}


__device__ __forceinline__ double4 ring_rect_scan(const RingRef& r) {
  double4 b = make_double4(INFINITY, INFINITY, -INFINITY, -INFINITY);
  for (int64_t i = 0; i < r.npts; i++) {
    double2 p = r.xy[i];
    b.x = fmin0(b.x, p.x); b.y = fmin0(b.y, p.y);
    b.z = fmax0(b.z, p.x); b.w = fmax0(b.w, p.y);
  }
  return b;
}

struct Ctx {
  GeomView B;         // build column
  RingIndexView R;    // build ring index
  GeomView P;         // probe column
  RingIndexView PR;   // probe ring index (polygon probes of polygon x polygon; else empty)
};

// The same over a build ring (offsets into the build column and its
// ring index; toff < 0: no segment tree).
template <int PREC>
__device__ __forceinline__ bool ring_contains_at(const Ctx& c, const double4& rb, int64_t v0, int64_t npts, double2 o,
                                                 float E, int64_t toff, double px, double py, bool allow_on_edge,
                                                 uint64_t& tests, uint64_t& fallbacks) {
  return ring_contains_run<PREC>(c.B.xy + v0, c.R.xy32 + v0, toff >= 0 ? c.R.seg_tree + toff : nullptr,
                                 c.R.seg_leaf_shift, npts, rb, o, E, px, py, allow_on_edge, tests, fallbacks);
}

template <int PREC>
__device__ __forceinline__ bool build_ring_contains(const Ctx& c, int64_t ring, double px, double py,
                                                    bool allow_on_edge, uint64_t& tests, uint64_t& fallbacks) {
  double4 rb = c.R.ring_rect[ring];
  if (!rect_covers(rb, px, py)) return false;
  int64_t v0 = c.B.pt_off[ring];
  int64_t toff = c.R.seg_tree_off ? c.R.seg_tree_off[ring] : -1;
  return ring_contains_at<PREC>(c, rb, v0, c.B.pt_off[ring + 1] - v0, c.R.origin[ring], c.R.err[ring], toff, px, py,
                                allow_on_edge, tests, fallbacks);
}

// tg_ring_contains_point for any ring (probe side: rect scanned on the fly).
__device__ bool any_ring_contains(const GeomView& g, int64_t ring, double px, double py, bool allow_on_edge) {
  RingRef r = ring_at(g, ring);
  double4 rb = ring_rect_scan(r);
  if (!rect_covers(rb, px, py)) return false;
  int64_t ns = r.npts >= 3 ? r.npts - 1 : 0;
  int on = 0, in = 0;
  uint64_t t = 0;
  ring_segments_f64(r, 0, ns, 1, px, py, on, in, t);
  return on ? allow_on_edge : (in != 0);
}

// poly_contains_point (tg.c:3909-3932) for part `part` of a build row.
template <int PREC>
__device__ __forceinline__ bool build_poly_contains(const Ctx& c, int64_t part, double px, double py,
                                                    bool allow_on_edge, uint64_t& tests, uint64_t& fallbacks) {
  int64_t r0 = c.B.ring_off[part], r1 = c.B.ring_off[part + 1];
  if (r0 == r1) return false;
  if (r1 - r0 == 1) return build_ring_contains<PREC>(c, r0, px, py, allow_on_edge, tests, fallbacks);
  if (c.B.pt_off[r0 + 1] - c.B.pt_off[r0] < 3) return false;  // tg_poly_empty
  if (!build_ring_contains<PREC>(c, r0, px, py, allow_on_edge, tests, fallbacks)) return false;
  for (int64_t h = r0 + 1; h < r1; h++)
    if (build_ring_contains<PREC>(c, h, px, py, !allow_on_edge, tests, fallbacks)) return false;
  return true;
}

__device__ bool any_poly_contains(const GeomView& g, int64_t part, double px, double py, bool allow_on_edge) {
  int64_t r0 = g.ring_off[part], r1 = g.ring_off[part + 1];
  if (r0 == r1) return false;
  if (r1 - r0 == 1) return any_ring_contains(g, r0, px, py, allow_on_edge);
  if (g.pt_off[r0 + 1] - g.pt_off[r0] < 3) return false;
  if (!any_ring_contains(g, r0, px, py, allow_on_edge)) return false;
  for (int64_t h = r0 + 1; h < r1; h++)
    if (any_ring_contains(g, h, px, py, !allow_on_edge)) return false;
  return true;
}

__device__ __forceinline__ bool point_of(const GeomView& g, int64_t row, double& x, double& y) {
  if (g.points_only) {
    // (a point column gathered into another order keeps no type array: its
    // rows are valid points or NaN)
    if (g.type && g.type[row] != kPoint) return false;
    double2 p = g.xy[row];
    x = p.x; y = p.y;
    return p.x == p.x || p.y == p.y;
  }
  if (g.type[row] != kPoint) return false;
  if (g.part_off[row] == g.part_off[row + 1]) return false;
  double2 p = g.xy[g.pt_off[g.ring_off[g.part_off[row]]]];
  x = p.x; y = p.y;
  // an empty point (NaN, NaN) that entered through a points-only batch
  return p.x == p.x || p.y == p.y;
}

// tg_point_touches_poly (tg.c:1171-1186) for one polygon part: the point
// lies on a segment of the exterior ring or of a hole (point_on_segment,
// tg.c:1192-1197: the segment's rect covers the point and the three points
// are collinear).
__device__ bool part_boundary_covers(const GeomView& g, int64_t part, double px, double py) {
  for (int64_t r = g.ring_off[part]; r < g.ring_off[part + 1]; r++) {
    const RingRef ring = ring_at(g, r);
    const int64_t ns = ring.npts >= 3 ? ring.npts - 1 : 0;
    for (int64_t j = 0; j < ns; j++) {
      const double2 a = ring.xy[j], b = ring.xy[j + 1];
      if (px < fmin0(a.x, b.x) || py < fmin0(a.y, b.y) || px > fmax0(a.x, b.x) || py > fmax0(a.y, b.y)) continue;
      if (collinear(a.x, a.y, b.x, b.y, px, py)) return true;
    }
  }
  return false;
}

// tg_geom_touches(polygon or multipolygon row, point): a polygon touches
// the point on its boundary (tg.c:6682-6690, 3957-3959); a multipolygon
// when some part does and no part that does not holds the point
// (tg.c:6735-6763, where a non-touching part that intersects ends it).
__device__ bool polyish_touches_point(const GeomView& g, int64_t row, double px, double py) {
  if (g.type[row] == kPolygon) return part_boundary_covers(g, g.part_off[row], px, py);
  bool touches = false;
  for (int64_t q = g.part_off[row]; q < g.part_off[row + 1]; q++) {
    if (part_boundary_covers(g, q, px, py)) touches = true;
    else if (any_poly_contains(g, q, px, py, true)) return false;
  }
  return touches;
}

// tg_geom_touches(point, polygon or multipolygon row): the point touches
// some part (tg.c:6578-6604; the other parts' interiors are not tested).
__device__ bool point_touches_polyish(const GeomView& g, int64_t row, double px, double py) {
  for (int64_t q = g.part_off[row]; q < g.part_off[row + 1]; q++)
    if (part_boundary_covers(g, q, px, py)) return true;
  return false;
}

// tg_point_covers_poly (tg.c:1150-1154) for every part of a polyish row.
__device__ bool point_covers_polyish(const GeomView& g, int64_t row, double px, double py) {
  int64_t p0 = g.part_off[row], p1 = g.part_off[row + 1];
  if (p0 == p1) return false;
  for (int64_t q = p0; q < p1; q++) {
    int64_t r0 = g.ring_off[q];
    if (g.ring_off[q + 1] == r0) return false;
    RingRef r = ring_at(g, r0);
    if (r.npts < 3) return false;
    double4 b = ring_rect_scan(r);
    if (!(feq(b.x, px) && feq(b.y, py) && feq(b.z, px) && feq(b.w, py))) return false;
  }
  return true;
}

// ---- polygon x polygon intersects (tg.c:3089-3111, 3166-3190, 4100-4136)

__device__ bool ring_covers_pt(const RingRef& r, const double4& rb, double px, double py) {
  if (!rect_covers(rb, px, py)) return false;
  int64_t ns = r.npts >= 3 ? r.npts - 1 : 0;
  int on = 0, in = 0;
  uint64_t t = 0;
  ring_segments_f64(r, 0, ns, 1, px, py, on, in, t);
  return on ? true : (in != 0);
}

__device__ bool ring_intersects_segment(const RingRef& ring, const double4& rb, double ax, double ay, double bx,
                                        double by) {
  double sx0 = fmin0(ax, bx), sy0 = fmin0(ay, by), sx1 = fmax0(ax, bx), sy1 = fmax0(ay, by);
  if (rb.x > sx1 || rb.z < sx0 || rb.y > sy1 || rb.w < sy0) return false;
  if (ring_covers_pt(ring, rb, ax, ay) || ring_covers_pt(ring, rb, bx, by)) return true;
  int count = 0;
  int64_t ns = ring.npts >= 3 ? ring.npts - 1 : 0;
  for (int64_t i = 0; i < ns && count < 2; i++) {
    double2 c = ring.xy[i], d = ring.xy[i + 1];
    if (segment_intersects(ax, ay, bx, by, c.x, c.y, d.x, d.y)) count++;
  }
  return count >= 2;
}

__device__ bool ring_intersects_ring(RingRef ring, RingRef other) {
  if (ring.npts < 3 || other.npts < 3) return false;
  double4 a = ring_rect_scan(ring), b = ring_rect_scan(other);
  if (b.x > a.z || b.z < a.x || b.y > a.w || b.w < a.y) return false;
  if ((b.z - b.x) * (b.w - b.y) > (a.z - a.x) * (a.w - a.y)) {
    RingRef t = ring; ring = other; other = t;
    double4 tb = a; a = b; b = tb;
  }
  int64_t ns = other.npts - 1;
  for (int64_t i = 0; i < ns; i++) {
    double2 p = other.xy[i], q = other.xy[i + 1];
    if (ring_intersects_segment(ring, a, p.x, p.y, q.x, q.y)) return true;
  }
  return false;
}

// A ring of a column with a ring index: the vertex runs, segment tree and
// FP32 frame of tg_ring_intersects_ring's operands.
struct IRing {
  const double2* xy;
  const float2* q32;
  const float4* tree;  // nullptr: none
  int ls;              // segments per tree leaf: 1 << ls
  int64_t npts;
  double4 rect;
  double2 o;
  float E;
};

__device__ __forceinline__ IRing iring(const GeomView& g, const RingIndexView& R, int64_t ring) {
  IRing r;
  const int64_t v0 = g.pt_off[ring];
  r.xy = g.xy + v0;
  r.q32 = R.xy32 + v0;
  r.npts = g.pt_off[ring + 1] - v0;
  const int64_t toff = R.seg_tree_off ? R.seg_tree_off[ring] : -1;
  r.tree = toff >= 0 ? R.seg_tree + toff : nullptr;
  r.ls = R.seg_leaf_shift;
  r.rect = R.ring_rect[ring];
  r.o = R.origin[ring];
  r.E = R.err[ring];
  return r;
}

// Whether at least two segments of r intersect segment (a, b) in the
// reference's segment x segment semantics (its exact bounding-box test
// comes first, so the tree's boxes, within 2E of the exact ones, prune
// conservatively with a margin of 4E).
__device__ __forceinline__ bool ring_two_crossings(const double2* xy, int64_t npts, const float4* tree, int ls,
                                                   float E, double2 o, double ax, double ay, double bx, double by) {
  const int64_t ns = npts >= 3 ? npts - 1 : 0;
  int count = 0;
  if (!tree || !(E < INFINITY)) {
    for (int64_t i = 0; i < ns && count < 2; i++) {
      double2 c = xy[i], d = xy[i + 1];
      if (segment_intersects(ax, ay, bx, by, c.x, c.y, d.x, d.y)) count++;
    }
    return count >= 2;
  }
  const float m = 4.0f * E;
  const float x0 = (float)(fmin(ax, bx) - o.x) - m, x1 = (float)(fmax(ax, bx) - o.x) + m;
  const float y0 = (float)(fmin(ay, by) - o.y) - m, y1 = (float)(fmax(ay, by) - o.y) + m;
  const int nsi = (int)ns;
  auto pass = [=](const float4& b) { return b.x <= x1 && b.z >= x0 && b.y <= y1 && b.w >= y0; };
  auto leaf = [&count, xy, nsi, ls, ax, ay, bx, by](int lf) {
    const int s0 = lf << ls, s1 = min(s0 + (1 << ls), nsi);
    for (int j = s0; j < s1; j++) {
      double2 c = xy[j], d = xy[j + 1];
      if (segment_intersects(ax, ay, bx, by, c.x, c.y, d.x, d.y)) {
        count++;
        if (count >= 2) return true;
      }
    }
    return false;
  };
  return walk_any(tree, seg_tree_shape(nsi, ls), pass, leaf);
}

// tg_ring_intersects_ring (tg.c:3166-3190) over indexed rings: some
// segment of the smaller-rect ring has an endpoint covered by the other
// ring or crosses it at least twice (tg.c:3089-3111).  The point tests walk
// the segment tree (certified FP32, FP64 for uncertain segments) and the
// crossing counts query it, instead of scanning every segment; the
// decision is the same.
__device__ bool iring_intersects_iring(IRing ring, IRing other) {
  if (ring.npts < 3 || other.npts < 3) return false;
  double4 a = ring.rect, b = other.rect;
  if (b.x > a.z || b.z < a.x || b.y > a.w || b.w < a.y) return false;
  if ((b.z - b.x) * (b.w - b.y) > (a.z - a.x) * (a.w - a.y)) {
    IRing t = ring; ring = other; other = t;
    a = ring.rect;
  }
  uint64_t tests = 0, fallbacks = 0;
  const int64_t ns = other.npts - 1;
  for (int64_t i = 0; i < ns; i++) {
    double2 p = other.xy[i], q = other.xy[i + 1];
    double sx0 = fmin0(p.x, q.x), sy0 = fmin0(p.y, q.y), sx1 = fmax0(p.x, q.x), sy1 = fmax0(p.y, q.y);
    if (a.x > sx1 || a.z < sx0 || a.y > sy1 || a.w < sy0) continue;
    if (ring_contains_run<0>(ring.xy, ring.q32, ring.tree, ring.ls, ring.npts, a, ring.o, ring.E, p.x, p.y, true, tests,
                             fallbacks) ||
        ring_contains_run<0>(ring.xy, ring.q32, ring.tree, ring.ls, ring.npts, a, ring.o, ring.E, q.x, q.y, true, tests,
                             fallbacks))
      return true;
    if (ring_two_crossings(ring.xy, ring.npts, ring.tree, ring.ls, ring.E, ring.o, p.x, p.y, q.x, q.y)) return true;
  }
  return false;
}

// tg.c:1910-2120 (process_points): the convex flag TG keeps per ring.
// Every point takes part, the closing one included, with its two successors
// taken modulo npoints; the first turn's sign sets the direction (a zero
// turn counts as counter-clockwise), later zero turns are ignored, and the
// first turn against the direction makes the ring concave.
__device__ bool ring_convex_tg(const RingRef& r) {
  const int64_t n = r.npts;
  if (n < 2) return false;
  int dir = 0;
  for (int64_t i = 0; i < n; i++) {
    const int64_t j = i + 1 < n ? i + 1 : i + 1 - n;
    const int64_t k = i + 2 < n ? i + 2 : i + 2 - n;
    const double2 a = r.xy[i], b = r.xy[j], c = r.xy[k];
    const double z = (b.x - a.x) * (c.y - b.y) - (b.y - a.y) * (c.x - b.x);
    if (dir == 0) {
      dir = z < 0 ? -1 : 1;
    } else if (z < 0) {
      if (dir == 1) return false;
    } else if (z > 0) {
      if (dir == -1) return false;
    }
  }
  return true;
}

// Point strictly inside ring a (tg_ring_contains_point(a, p, false).hit).
__device__ __forceinline__ bool ring_holds_point_strict(const RingRef& a, const double4& ra, double px, double py) {
  if (!rect_covers(ra, px, py)) return false;
  int on = 0, in = 0;
  uint64_t t = 0;
  ring_segments_f64(a, 0, a.npts - 1, 1, px, py, on, in, t);
  return !on && in;
}

// tg_ring_contains_ring(hole, ring, false) as poly x poly intersects uses
// it (tg.c:3119-3151): a convex hole contains the ring when every point of
// the ring lies strictly inside it; a concave one when every segment does
// (tg_ring_contains_segment(.., false), tg.c:2910-3016: the hole's rect
// covers the segment's, both endpoints strictly inside, a zero-length
// segment needing one, and no hole segment meets it).
__device__ bool ring_contains_ring_strict(const RingRef& a, const RingRef& b) {
  if (a.npts < 3 || b.npts < 3) return false;
  double4 ra = ring_rect_scan(a), rb = ring_rect_scan(b);
  if (rb.x < ra.x || rb.z > ra.z || rb.y < ra.y || rb.w > ra.w) return false;
  if (ring_convex_tg(a)) {
    for (int64_t i = 0; i < b.npts; i++) {
      const double2 p = b.xy[i];
      if (!ring_holds_point_strict(a, ra, p.x, p.y)) return false;
    }
    return true;
  }
  for (int64_t i = 0; i + 1 < b.npts; i++) {
    const double2 p = b.xy[i], q = b.xy[i + 1];
    const double sx0 = fmin0(p.x, q.x), sy0 = fmin0(p.y, q.y), sx1 = fmax0(p.x, q.x), sy1 = fmax0(p.y, q.y);
    if (sx0 < ra.x || sx1 > ra.z || sy0 < ra.y || sy1 > ra.w) return false;
    if (!ring_holds_point_strict(a, ra, p.x, p.y)) return false;
    if (feq(q.x, p.x) && feq(q.y, p.y)) continue;
    if (!ring_holds_point_strict(a, ra, q.x, q.y)) return false;
    for (int64_t j = 0; j + 1 < a.npts; j++) {
      const double2 c = a.xy[j], d = a.xy[j + 1];
      if (segment_intersects(p.x, p.y, q.x, q.y, c.x, c.y, d.x, d.y)) return false;
    }
  }
  return true;
}

__device__ bool poly_intersects_poly(const GeomView& A, int64_t pa, const GeomView& Bv, int64_t pb,
                                     const RingIndexView* RA = nullptr, const RingIndexView* RB = nullptr) {
  int64_t a0 = A.ring_off[pa], a1 = A.ring_off[pa + 1];
  int64_t b0 = Bv.ring_off[pb], b1 = Bv.ring_off[pb + 1];
  if (a0 == a1 || b0 == b1) return false;
  RingRef ae = ring_at(A, a0), be = ring_at(Bv, b0);
  const bool indexed = RA && RB && RA->ring_rect && RB->ring_rect;
  if (a1 - a0 == 1 && b1 - b0 == 1)
    return indexed ? iring_intersects_iring(iring(A, *RA, a0), iring(Bv, *RB, b0)) : ring_intersects_ring(ae, be);
  if (ae.npts < 3 || be.npts < 3) return false;
  if (!(indexed ? iring_intersects_iring(iring(Bv, *RB, b0), iring(A, *RA, a0)) : ring_intersects_ring(be, ae)))
    return false;
  for (int64_t h = a0 + 1; h < a1; h++)
    if (ring_contains_ring_strict(ring_at(A, h), be)) return false;
  for (int64_t h = b0 + 1; h < b1; h++)
    if (ring_contains_ring_strict(ring_at(Bv, h), ae)) return false;
  return true;
}

// ---------------------------------------------------------------------------
// TG's relations with boundaries between polygons, and intersects with
// lines (tg.c:2669-3275, 3595-3611, 4016-4249, 5925-6800, 15299-15330), in
// FP64 with the reference's expression sequences, one lane per candidate.
// Rings are closed (the decoder rejects open ones): npts - 1 segments.

__device__ __forceinline__ bool pteq2(double ax, double ay, double bx, double by) { return feq(ax, bx) && feq(ay, by); }

// tg.c:1910-2120 (process_points): the clockwise flag, the shoelace sum
// over every point (closing one included, successor modulo npoints) in
// index order, clockwise when positive.
__device__ bool ring_clockwise_tg(const RingRef& r) {
  const int64_t n = r.npts;
  if (n < 2) return false;
  double cwc = 0;
  for (int64_t i = 0; i < n; i++) {
    const int64_t j = i + 1 < n ? i + 1 : i + 1 - n;
    const double2 a = r.xy[i], b = r.xy[j];
    cwc += (b.x - a.x) * (b.y + a.y);
  }
  return cwc > 0;
}

// tg.c:2669-2822: point in ring with the index of the last segment (in
// index order) the point is on, -1 when none.  A segment the point is on
// sets the result to allow_on_edge; IN toggles count only before the first
// such segment.
__device__ bool ring_contains_point_ix(const RingRef& r, const double4& rr, double px, double py, bool allow_on_edge,
                                       int64_t& idx) {
  idx = -1;
  if (!rect_covers(rr, px, py)) return false;
  bool in = false;
  const int64_t ns = r.npts - 1;
  for (int64_t i = 0; i < ns; i++) {
    const double2 a = r.xy[i], b = r.xy[i + 1];
    if (py < fmin0(a.y, b.y) || py > fmax0(a.y, b.y)) continue;  // OUT in raycast's first test
    const int rc = raycast(a.x, a.y, b.x, b.y, px, py);
    if (rc == kRayOn) {
      in = allow_on_edge;
      idx = i;
    } else if (rc == kRayIn && idx == -1) {
      in = !in;
    }
  }
  return in;
}

__device__ __forceinline__ bool rect_covers_rect4(const double4& a, double x0, double y0, double x1, double y1) {
  return !(x0 < a.x || x1 > a.z || y0 < a.y || y1 > a.w);
}
__device__ __forceinline__ bool rect_isect4(const double4& a, double x0, double y0, double x1, double y1) {
  return !(x0 > a.z || x1 < a.x || y0 > a.w || y1 < a.y);
}

// tg.c:2910-3016 tg_ring_contains_segment (the iterators' own segment test
// rejects disjoint rects, so visiting every segment decides as the rect
// search does).
__device__ bool ring_contains_segment_tg(const RingRef& r, const double4& rr, bool convex, double sax, double say,
                                         double sbx, double sby, bool allow_on_edge) {
  if (!rect_covers_rect4(rr, fmin0(sax, sbx), fmin0(say, sby), fmax0(sax, sbx), fmax0(say, sby))) return false;
  int64_t ia, ib;
  if (!ring_contains_point_ix(r, rr, sax, say, allow_on_edge, ia)) return false;
  if (pteq2(sbx, sby, sax, say)) return true;
  if (!ring_contains_point_ix(r, rr, sbx, sby, allow_on_edge, ib)) return false;
  if (convex) return true;
  const int64_t ns = r.npts - 1;
  int mode = 3;
  if (allow_on_edge) {
    if (ia != -1 && ib != -1) {
      if (ib == ia) return true;
      double2 a0 = r.xy[ia], a1 = r.xy[ia + 1], b0 = r.xy[ib], b1 = r.xy[ib + 1];
      if (pteq2(a0.x, a0.y, sax, say) || pteq2(a1.x, a1.y, sax, say) || pteq2(b0.x, b0.y, sax, say) ||
          pteq2(b1.x, b1.y, sax, say) || pteq2(a0.x, a0.y, sbx, sby) || pteq2(a1.x, a1.y, sbx, sby) ||
          pteq2(b0.x, b0.y, sbx, sby) || pteq2(b1.x, b1.y, sbx, sby))
        return true;
      if (ib < ia) {
        double2 t0 = a0, t1 = a1;
        a0 = b0; a1 = b1; b0 = t0; b1 = t1;
      }
      const double2 pts[5] = {a0, a1, b0, b1, a0};
      double cwc = 0.0;
      for (int i = 0; i < 4; i++) cwc += (pts[i + 1].x - pts[i].x) * (pts[i + 1].y + pts[i].y);
      if ((cwc > 0) != ring_clockwise_tg(r)) return false;
      mode = 1;
    } else if (ia != -1) {
      mode = 4;
    } else if (ib != -1) {
      mode = 2;
    } else {
      mode = 5;
    }
  }
  for (int64_t i = 0; i < ns; i++) {
    const double2 c = r.xy[i], d = r.xy[i + 1];
    if (!segment_intersects(sax, say, sbx, sby, c.x, c.y, d.x, d.y)) continue;
    bool hit;
    switch (mode) {
      case 1: hit = raycast(c.x, c.y, d.x, d.y, sax, say) != kRayOn && raycast(c.x, c.y, d.x, d.y, sbx, sby) != kRayOn; break;
      case 4: hit = raycast(c.x, c.y, d.x, d.y, sax, say) != kRayOn; break;
      case 2: hit = raycast(c.x, c.y, d.x, d.y, sbx, sby) != kRayOn; break;
      case 5: hit = raycast(sax, say, sbx, sby, c.x, c.y) != kRayOn && raycast(sax, say, sbx, sby, d.x, d.y) != kRayOn; break;
      default: hit = true;
    }
    if (hit) return false;
  }
  return true;
}

// tg.c:3119-3151 tg_ring_contains_ring with allow_on_edge.
__device__ bool ring_contains_ring_tg(const RingRef& a, const RingRef& b, bool allow_on_edge) {
  if (a.npts < 3 || b.npts < 3) return false;
  const double4 ra = ring_rect_scan(a), rb = ring_rect_scan(b);
  if (!rect_covers_rect4(ra, rb.x, rb.y, rb.z, rb.w)) return false;
  int64_t idx;
  if (ring_convex_tg(a)) {
    for (int64_t i = 0; i < b.npts; i++) {
      const double2 p = b.xy[i];
      if (!ring_contains_point_ix(a, ra, p.x, p.y, allow_on_edge, idx)) return false;
    }
    return true;
  }
  for (int64_t i = 0; i + 1 < b.npts; i++) {
    const double2 p = b.xy[i], q = b.xy[i + 1];
    if (!ring_contains_segment_tg(a, ra, false, p.x, p.y, q.x, q.y, allow_on_edge)) return false;
  }
  return true;
}

// tg.c:3019-3111 tg_ring_intersects_segment with intersegiter: segments in
// index order, the count reset by a collinear overlap, stopped at two.
__device__ bool ring_intersects_segment_tg(const RingRef& r, const double4& rr, double sax, double say, double sbx,
                                           double sby, bool allow_on_edge) {
  const double x0 = fmin0(sax, sbx), y0 = fmin0(say, sby), x1 = fmax0(sax, sbx), y1 = fmax0(say, sby);
  if (!rect_isect4(rr, x0, y0, x1, y1)) return false;
  int64_t idx;
  if (ring_contains_point_ix(r, rr, sax, say, allow_on_edge, idx) ||
      ring_contains_point_ix(r, rr, sbx, sby, allow_on_edge, idx))
    return true;
  int count = 0;
  bool a_on = false, b_on = false;
  const int64_t ns = r.npts - 1;
  for (int64_t i = 0; i < ns; i++) {
    const double2 c = r.xy[i], d = r.xy[i + 1];
    if (!segment_intersects(sax, say, sbx, sby, c.x, c.y, d.x, d.y)) continue;
    if (allow_on_edge) {
      if (++count >= 2) break;
      continue;
    }
    const bool ccol = collinear(sax, say, sbx, sby, c.x, c.y);
    const bool dcol = collinear(sax, say, sbx, sby, d.x, d.y);
    if (ccol && dcol) {
      count = 0;
    } else {
      if (!a_on && (pteq2(sax, say, c.x, c.y) || pteq2(sax, say, d.x, d.y))) {
        a_on = true;
        continue;
      }
      if (!b_on && (pteq2(sbx, sby, c.x, c.y) || pteq2(sbx, sby, d.x, d.y))) {
        b_on = true;
        continue;
      }
      count++;
    }
    if (count >= 2) break;
  }
  return count >= 2;
}

// tg.c:3166-3190 tg_ring_intersects_ring with allow_on_edge.
__device__ bool ring_intersects_ring_tg(RingRef ring, RingRef other, bool allow_on_edge) {
  if (ring.npts < 3 || other.npts < 3) return false;
  double4 a = ring_rect_scan(ring), b = ring_rect_scan(other);
  if (!rect_isect4(a, b.x, b.y, b.z, b.w)) return false;
  if ((b.z - b.x) * (b.w - b.y) > (a.z - a.x) * (a.w - a.y)) {
    RingRef t = ring; ring = other; other = t;
    a = b;
  }
  for (int64_t i = 0; i + 1 < other.npts; i++) {
    const double2 p = other.xy[i], q = other.xy[i + 1];
    if (ring_intersects_segment_tg(ring, a, p.x, p.y, q.x, q.y, allow_on_edge)) return true;
  }
  return false;
}

// tg.c:3436-3443 tg_line_covers_point over a run of segments.
__device__ bool run_covers_point(const RingRef& r, double px, double py) {
  for (int64_t i = 0; i + 1 < r.npts; i++) {
    const double2 a = r.xy[i], b = r.xy[i + 1];
    if (segment_covers_point(a.x, a.y, b.x, b.y, px, py)) return true;
  }
  return false;
}

// tg.c:4040-4092 tg_poly_covers_poly (also tg_poly_contains_poly,
// tg.c:4094-4096); polygons are parts pa of A and pb of B.
__device__ bool poly_covers_poly_tg(const GeomView& A, int64_t pa, const GeomView& B, int64_t pb) {
  const int64_t a0 = A.ring_off[pa], a1 = A.ring_off[pa + 1], b0 = B.ring_off[pb], b1 = B.ring_off[pb + 1];
  if (a0 == a1 || b0 == b1) return false;
  const RingRef ae = ring_at(A, a0), be = ring_at(B, b0);
  if (ae.npts < 3 || be.npts < 3) return false;
  if (!ring_contains_ring_tg(ae, be, true)) return false;
  for (int64_t h = a0 + 1; h < a1; h++) {
    const RingRef hr = ring_at(A, h);
    if (ring_contains_ring_tg(hr, be, true) || ring_intersects_ring_tg(hr, be, false)) {
      bool covers = false;
      for (int64_t j = b0 + 1; j < b1 && !covers; j++) covers = ring_contains_ring_tg(ring_at(B, j), hr, true);
      if (!covers) return false;
    }
  }
  return true;
}

// tg.c:4184-4249 tg_poly_touches_poly.
__device__ bool poly_touches_poly_tg(const GeomView& A, int64_t pa, const GeomView& B, int64_t pb) {
  const int64_t a0 = A.ring_off[pa], a1 = A.ring_off[pa + 1], b0 = B.ring_off[pb], b1 = B.ring_off[pb + 1];
  if (a0 == a1 || b0 == b1) return false;
  const RingRef aext = ring_at(A, a0), bext = ring_at(B, b0);
  const double4 ra = ring_rect_scan(aext), rb = ring_rect_scan(bext);
  if (!rect_isect4(ra, rb.x, rb.y, rb.z, rb.w)) return false;
  for (int ii = 0; ii < 2; ii++) {
    const GeomView& P = ii == 0 ? A : B;
    const int64_t h0 = (ii == 0 ? a0 : b0) + 1, h1 = ii == 0 ? a1 : b1;
    const RingRef& ring = ii == 0 ? bext : aext;
    for (int64_t h = h0; h < h1; h++) {
      const RingRef hole = ring_at(P, h);
      if (ring_contains_ring_tg(hole, ring, true)) {
        for (int64_t j = 0; j < ring.npts; j++)
          if (run_covers_point(hole, ring.xy[j].x, ring.xy[j].y)) return true;
        return false;
      }
    }
  }
  const int64_t ansegs = aext.npts - 1, bnsegs = bext.npts - 1;
  int64_t atouches = 0, btouches = 0;
  for (int ii = 0; ii < 2; ii++) {
    const RingRef& x = ii == 0 ? aext : bext;
    const RingRef& y = ii == 0 ? bext : aext;
    const double4 rx = ii == 0 ? ra : rb;
    int64_t touches = 0;
    for (int64_t i = 0; i + 1 < y.npts; i++) {
      const double2 p = y.xy[i], q = y.xy[i + 1];
      const bool isects0 = ring_intersects_segment_tg(x, rx, p.x, p.y, q.x, q.y, true);
      const bool isects1 = ring_intersects_segment_tg(x, rx, p.x, p.y, q.x, q.y, false);
      if (isects0 && !isects1) touches++;
      else if (isects0 || isects1) return false;
    }
    if (ii == 0) btouches = touches;
    else atouches = touches;
  }
  if (atouches > 0 || btouches > 0) return !(atouches == ansegs && btouches == bnsegs);
  return false;
}

__device__ __forceinline__ bool part_empty(const GeomView& g, int64_t q) {
  const int64_t r0 = g.ring_off[q], r1 = g.ring_off[q + 1];
  return r0 == r1 || g.pt_off[r0 + 1] - g.pt_off[r0] < 3;
}

// tg_geom_covers(A row a, B row b) for (multi)polygon rows (tg.c:5925-6045).
__device__ bool polyish_covers_polyish(const GeomView& A, int64_t a, const GeomView& B, int64_t b) {
  const int64_t a0 = A.part_off[a], a1 = A.part_off[a + 1], b0 = B.part_off[b], b1 = B.part_off[b + 1];
  if (a0 == a1 || b0 == b1) return false;
  if (A.type[a] == kPolygon) {
    for (int64_t q = b0; q < b1; q++)
      if (!poly_covers_poly_tg(A, a0, B, q)) return false;
    return true;
  }
  bool any = false;
  for (int64_t q = b0; q < b1; q++) {
    if (part_empty(B, q)) continue;
    bool found = false;
    for (int64_t r = a0; r < a1 && !found; r++) found = poly_covers_poly_tg(A, r, B, q);
    if (!found) return false;
    any = true;
  }
  return any;
}

// tg_geom_contains for (multi)polygon rows (tg.c:6060-6106, 6418-6520).
__device__ bool polyish_contains_polyish(const GeomView& A, int64_t a, const GeomView& B, int64_t b) {
  if (A.type[a] == kPolygon) return polyish_covers_polyish(A, a, B, b);
  if (!polyish_covers_polyish(A, a, B, b)) return false;
  for (int64_t q = B.part_off[b]; q < B.part_off[b + 1]; q++) {
    if (part_empty(B, q)) continue;
    for (int64_t r = A.part_off[a]; r < A.part_off[a + 1]; r++) {
      if (part_empty(A, r)) continue;
      if (poly_intersects_poly(A, r, B, q) && !poly_touches_poly_tg(A, r, B, q)) return true;
    }
  }
  return false;
}

// poly_touches_geom(part pa of A, B row b) (tg.c:6660-6700).
__device__ bool part_touches_polyish(const GeomView& A, int64_t pa, const GeomView& B, int64_t b) {
  const int64_t b0 = B.part_off[b], b1 = B.part_off[b + 1];
  if (b0 == b1) return false;
  if (B.type[b] == kPolygon) return poly_touches_poly_tg(A, pa, B, b0);
  bool touches = false;
  for (int64_t q = b0; q < b1; q++) {
    if (poly_touches_poly_tg(A, pa, B, q)) touches = true;
    else if (poly_intersects_poly(A, pa, B, q)) return false;
  }
  return touches;
}

// tg_geom_touches for (multi)polygon rows (tg.c:6660-6770).
__device__ bool polyish_touches_polyish(const GeomView& A, int64_t a, const GeomView& B, int64_t b) {
  const int64_t a0 = A.part_off[a], a1 = A.part_off[a + 1];
  if (a0 == a1) return false;
  if (A.type[a] == kPolygon) return part_touches_polyish(A, a0, B, b);
  bool touches = false;
  for (int64_t r = a0; r < a1; r++) {
    if (part_touches_polyish(A, r, B, b)) {
      touches = true;
    } else {
      for (int64_t q = B.part_off[b]; q < B.part_off[b + 1]; q++)
        if (poly_intersects_poly(A, r, B, q)) return false;
    }
  }
  return touches;
}

// Lines are open runs (tg.c:2126-2135): npts - 1 segments.

// tg.c:3252-3275 tg_ring_intersects_line.
__device__ bool ring_intersects_run_tg(const RingRef& r, const RingRef& l, bool allow_on_edge) {
  if (r.npts < 3 || l.npts < 2) return false;
  const double4 rr = ring_rect_scan(r), lr = ring_rect_scan(l);
  if (!rect_isect4(rr, lr.x, lr.y, lr.z, lr.w)) return false;
  for (int64_t i = 0; i + 1 < l.npts; i++) {
    const double2 p = l.xy[i], q = l.xy[i + 1];
    if (ring_intersects_segment_tg(r, rr, p.x, p.y, q.x, q.y, allow_on_edge)) return true;
  }
  return false;
}

// tg.c:3194-3248 tg_ring_contains_line.
__device__ bool ring_contains_run_tg(const RingRef& a, const RingRef& l, bool allow_on_edge, bool respect_boundaries) {
  if (a.npts < 3 || l.npts < 2) return false;
  const double4 ra = ring_rect_scan(a), lr = ring_rect_scan(l);
  if (!rect_covers_rect4(ra, lr.x, lr.y, lr.z, lr.w)) return false;
  const bool convex = ring_convex_tg(a);
  for (int64_t i = 0; i + 1 < l.npts; i++) {
    const double2 p = l.xy[i], q = l.xy[i + 1];
    if (!allow_on_edge && respect_boundaries) {
      if (!ring_contains_segment_tg(a, ra, convex, p.x, p.y, q.x, q.y, true)) return false;
      if (!ring_intersects_segment_tg(a, ra, p.x, p.y, q.x, q.y, false)) return false;
    } else if (!ring_contains_segment_tg(a, ra, convex, p.x, p.y, q.x, q.y, allow_on_edge)) {
      return false;
    }
  }
  return true;
}

// tg.c:4016-4038 tg_poly_intersects_line (part pa of A, line run l).
__device__ bool poly_intersects_run_tg(const GeomView& A, int64_t pa, const RingRef& l) {
  const int64_t a0 = A.ring_off[pa], a1 = A.ring_off[pa + 1];
  if (a0 == a1) return false;
  const RingRef e = ring_at(A, a0);
  if (e.npts < 3 || l.npts < 2) return false;
  if (!ring_intersects_run_tg(e, l, true)) return false;
  for (int64_t h = a0 + 1; h < a1; h++)
    if (ring_contains_run_tg(ring_at(A, h), l, false, false)) return false;
  return true;
}

// tg.c:3595-3611 (SI_INTERSECTS) through tg_ring_ring_search
// (tg.c:2596-2647): rects meeting and some pair of segments intersecting.
__device__ bool run_intersects_run_tg(const RingRef& a, const RingRef& b) {
  if (a.npts < 2 || b.npts < 2) return false;
  const double4 ra = ring_rect_scan(a), rb = ring_rect_scan(b);
  if (!rect_isect4(ra, rb.x, rb.y, rb.z, rb.w)) return false;
  for (int64_t i = 0; i + 1 < a.npts; i++) {
    const double2 p = a.xy[i], q = a.xy[i + 1];
    for (int64_t j = 0; j + 1 < b.npts; j++) {
      const double2 c = b.xy[j], d = b.xy[j + 1];
      if (segment_intersects(p.x, p.y, q.x, q.y, c.x, c.y, d.x, d.y)) return true;
    }
  }
  return false;
}

__device__ __forceinline__ bool is_lineish(int t) { return t == kLine || t == kMultiLine; }

// tg_geom_intersects(L row l, O row o) where L is a (multi)linestring and O a
// (multi)linestring or (multi)polygon row (tg.c:5544-5740: some pair of
// parts intersects).
__device__ bool lineish_intersects(const GeomView& L, int64_t l, const GeomView& O, int64_t o) {
  const int to = O.type[o];
  for (int64_t q = L.part_off[l]; q < L.part_off[l + 1]; q++) {
    const RingRef ln = ring_at(L, L.ring_off[q]);
    for (int64_t r = O.part_off[o]; r < O.part_off[o + 1]; r++) {
      bool hit;
      if (is_lineish(to)) {
        hit = run_intersects_run_tg(ln, ring_at(O, O.ring_off[r]));
      } else {
        hit = poly_intersects_run_tg(O, r, ln);
      }
      if (hit) return true;
    }
  }
  return false;
}

// Evaluates pred(build b, probe p).  Returns 0/1, or -1 when unsupported.
__device__ int eval_relation(const Ctx& c, int64_t b, int64_t p, int pred, int precision, uint64_t& tests,
                             uint64_t& fallbacks) {
  int tb = c.B.type[b], tp = c.P.type[p];
  if (tb == kNull || tp == kNull) return 0;
  double px, py, bx, by;
  if (is_polyish(tb) && tp == kPoint) {
    if (!point_of(c.P, p, px, py)) return 0;
    if (n_parts(c.B, b) == 0) return 0;
    bool allow;
    switch (pred) {
      case kIntersects: case kCovers: allow = true; break;
      case kContains: allow = false; break;
      case kWithin: return 0;
      case kCoveredBy: return point_covers_polyish(c.B, b, px, py) ? 1 : 0;
      case kTouches: return polyish_touches_point(c.B, b, px, py) ? 1 : 0;
      case kEquals: return 0;  // within && contains (tg.c:15299-15301): within(polygon, point) is false
      default: return -1;
    }
    for (int64_t q = c.B.part_off[b]; q < c.B.part_off[b + 1]; q++)
      if (precision == 0 ? build_poly_contains<0>(c, q, px, py, allow, tests, fallbacks)
                         : build_poly_contains<1>(c, q, px, py, allow, tests, fallbacks))
        return 1;
    return 0;
  }
  if (tb == kPoint && is_polyish(tp)) {
    if (!point_of(c.B, b, bx, by)) return 0;
    if (n_parts(c.P, p) == 0) return 0;
    bool allow;
    switch (pred) {
      case kIntersects: case kCoveredBy: allow = true; break;
      case kWithin: allow = false; break;
      case kContains: return 0;
      case kCovers: return point_covers_polyish(c.P, p, bx, by) ? 1 : 0;
      case kTouches: return point_touches_polyish(c.P, p, bx, by) ? 1 : 0;
      case kEquals: return 0;  // contains(point, polygon) is false
      default: return -1;
    }
    for (int64_t q = c.P.part_off[p]; q < c.P.part_off[p + 1]; q++)
      if (any_poly_contains(c.P, q, bx, by, allow)) return 1;
    return 0;
  }
  if (tb == kPoint && tp == kPoint) {
    if (!point_of(c.B, b, bx, by) || !point_of(c.P, p, px, py)) return 0;
    switch (pred) {
      case kIntersects: case kContains: case kWithin: case kCovers: case kCoveredBy: case kEquals:
        return (feq(bx, px) && feq(by, py)) ? 1 : 0;
      case kTouches: return 0;  // points have no boundary (tg.c:1102-1106)
      default: return -1;
    }
  }
  if (is_polyish(tb) && is_polyish(tp) && pred == kIntersects) {
    for (int64_t qa = c.B.part_off[b]; qa < c.B.part_off[b + 1]; qa++)
      for (int64_t qb = c.P.part_off[p]; qb < c.P.part_off[p + 1]; qb++)
        if (poly_intersects_poly(c.B, qa, c.P, qb, &c.R, &c.PR)) return 1;
    return 0;
  }
  return -1;
}

// The relations eval_relation leaves out: (multi)polygon x (multi)polygon
// under predicates other than Intersects, and Intersects with
// (multi)linestrings.  Returns 0/1, -1 when unsupported, -2 when the pair
// is eval_relation's.
__device__ int eval_relate_extended(const Ctx& c, int64_t b, int64_t p, int pred) {
  const int tb = c.B.type[b], tp = c.P.type[p];
  if (tb == kNull || tp == kNull) return 0;
  if (is_polyish(tb) && is_polyish(tp) && pred != kIntersects) {
    switch (pred) {
      case kCovers: return polyish_covers_polyish(c.B, b, c.P, p) ? 1 : 0;
      case kCoveredBy: return polyish_covers_polyish(c.P, p, c.B, b) ? 1 : 0;
      case kContains: return polyish_contains_polyish(c.B, b, c.P, p) ? 1 : 0;
      case kWithin: return polyish_contains_polyish(c.P, p, c.B, b) ? 1 : 0;
      case kTouches: return polyish_touches_polyish(c.B, b, c.P, p) ? 1 : 0;
      // within(a, b) && contains(a, b) (tg.c:15299-15301)
      case kEquals:
        return (polyish_contains_polyish(c.P, p, c.B, b) && polyish_contains_polyish(c.B, b, c.P, p)) ? 1 : 0;
      default: return -1;
    }
  }
  if (pred == kIntersects && (is_lineish(tb) || is_lineish(tp))) {
    if (tb == kPoint || tp == kPoint) {
      // a point row (points-only column or not) and a line row:
      // tg_line_covers_point over the line's parts (tg.c:3436-3462)
      const bool bpt = tb == kPoint;
      double x, y;
      if (!point_of(bpt ? c.B : c.P, bpt ? b : p, x, y)) return 0;
      const GeomView& L = bpt ? c.P : c.B;
      const int64_t l = bpt ? p : b;
      for (int64_t q = L.part_off[l]; q < L.part_off[l + 1]; q++)
        if (run_covers_point(ring_at(L, L.ring_off[q]), x, y)) return 1;
      return 0;
    }
    if (is_lineish(tb)) return lineish_intersects(c.B, b, c.P, p) ? 1 : 0;
    return lineish_intersects(c.P, p, c.B, b) ? 1 : 0;
  }
  return -2;
}

// Whether candidate i is handled by the wavefront-cooperative kernel.
__device__ __forceinline__ bool is_long(const Ctx& c, int64_t b, int64_t p) {
  if (c.B.type[b] != kPolygon || c.P.type[p] != kPoint) return false;
  int64_t part = c.B.part_off[b];
  if (c.B.ring_off[part + 1] - c.B.ring_off[part] != 1) return false;
  int64_t ring = c.B.ring_off[part];
  if (c.R.seg_tree_off && c.R.seg_tree_off[ring] >= 0) return false;  // the tree walk handles it
  return c.B.pt_off[ring + 1] - c.B.pt_off[ring] > kLongRing;
}

// Per-wavefront LDS output buffer of a kernel (see wave_out.h).
#define SJ_WAVE_OUT(wo)                                                                   \
  __shared__ uint32_t wo##_b[kBlock / 64][kWaveOutCap], wo##_p[kBlock / 64][kWaveOutCap]; \
  WaveOut wo { wo##_b[threadIdx.x >> 6], wo##_p[threadIdx.x >> 6], 0 }

// Wavefront compaction: lanes with keep=1 append (b, p) to the output.
__device__ __forceinline__ void emit(WaveOut& wo, bool keep, uint32_t b, uint32_t p, uint32_t* ob, uint32_t* op,
                                     uint64_t cap, unsigned long long* counter) {
  wave_append(wo, keep, b, p, ob, op, cap, counter);
}

// Point-in-polygon refine (build polygons / multipolygons x probe points,
// the point-probe case of Intersects, Covers and Contains): one lane per
// candidate, precision fixed at compile time, nothing but the ring walk
// inlined.
//
// With PRECISION == kPrecDeferred a candidate that met an uncertain FP32
// test is not decided here: its index goes to `redo` (count in *n_redo)
// and the FP64 instance refines it from that list (`list`: candidate
// indices to refine instead of 0..n-1).
#ifndef SJ_PIP_WAVES
#define SJ_PIP_WAVES 8
#endif
// The walk is latency-bound: more resident waves per SIMD beat the few
// registers spilled to reach them.
template <int PRECISION, int WAVES = SJ_PIP_WAVES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_refine_pip(Ctx c, int allow_on_edge, const uint32_t* cb,
                                                       const uint32_t* cp, uint64_t n, const uint32_t* list,
                                                       uint32_t* ob, uint32_t* op, uint64_t cap,
                                                       unsigned long long* counter, unsigned long long* stats,
                                                       const uint32_t* probe_ids, uint32_t* redo,
                                                       unsigned long long* n_redo, const unsigned long long* n_dev) {
  SJ_WAVE_OUT(wo);
  if (n_dev) n = min(n, (uint64_t)*n_dev);  // candidate count produced on the device (n: its bound)
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t tests = 0, fallbacks = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    uint64_t i = base + threadIdx.x;
    bool keep = false;
    uint32_t b = 0, p = 0;
    if (i < n) {
      if (list) i = list[i];
      b = cb[i];
      p = cp[i];
      const uint64_t fb0 = fallbacks;
      // the probe is a point column: null and empty rows hold NaN
      double2 pt = c.P.xy[p];
      if (pt.x == pt.x || pt.y == pt.y) {
        const PolyDesc d = c.R.desc[b];
        // cell raster: inside / outside away from the boundary, no walk
        int cls = 0;
        if (d.raster_off >= 0 && rect_covers(d.rect, pt.x, pt.y)) {
          const double cw = (d.rect.z - d.rect.x) / kRasterCells, ch = (d.rect.w - d.rect.y) / kRasterCells;
          const int ci = min(max((int)floor((pt.x - d.rect.x) / cw), 0), kRasterCells - 1);
          const int cj = min(max((int)floor((pt.y - d.rect.y) / ch), 0), kRasterCells - 1);
          const int cell = cj * kRasterCells + ci;
          cls = (c.R.raster[d.raster_off + (cell >> 4)] >> (2 * (cell & 15))) & 3;
        }
        if (cls) {
          keep = cls == 1;
        } else if (d.npts >= 0) {
          keep = ring_contains_at<PRECISION>(c, d.rect, d.v0, d.npts, ring_origin(d.rect), d.err, d.tree_off, pt.x,
                                             pt.y, allow_on_edge != 0, tests, fallbacks);
        } else if (c.B.type[b] != kNull) {
          int64_t q1 = c.B.part_off[b + 1];
          for (int64_t q = c.B.part_off[b]; q < q1 && !keep; q++)
            keep = build_poly_contains<PRECISION>(c, q, pt.x, pt.y, allow_on_edge != 0, tests, fallbacks);
        }
      }
      if constexpr (PRECISION == kPrecDeferred) {
        if (fallbacks != fb0) {
          keep = false;
          redo[atomicAdd(n_redo, 1ull)] = (uint32_t)i;
        }
      }
      if (keep && probe_ids) p = probe_ids[p];
    }
    emit(wo, keep, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
  for (int o = 32; o > 0; o >>= 1) {
    tests += __shfl_xor(tests, o);
    fallbacks += __shfl_xor(fallbacks, o);
  }
  // (wavefronts with nothing to add skip the atomics: a launch over an empty
  // list would otherwise send one pair per wavefront to the same address)
  if ((threadIdx.x & 63) == 0 && stats && (tests | fallbacks)) {
    atomicAdd(&stats[0], (unsigned long long)tests);
    atomicAdd(&stats[1], (unsigned long long)fallbacks);
  }
}

// The FP64 pass over the candidates the certified walks deferred (list[0,
// *n_list) of candidate indices), one wavefront per candidate: for a row
// of one ring the lanes take the ring's segments with stride 64 (or its
// tree's leaves, skipping those a +x ray cannot meet, as the per-lane
// FP64 walk prunes them) and vote (any ON: allow_on_edge, else the parity
// of IN): tg_ring_contains_point along +x (tg.c:2809-2822).  Other rows
// are decided by lane 0 as in k_refine_pip<1>.  The deferred rays were long or met a vertex, so a lane
// per candidate left most of its wavefront idle.
__global__ __launch_bounds__(kBlock) void k_pip_redo_wave(Ctx c, int allow_on_edge, const uint32_t* cb,
                                                          const uint32_t* cp, const uint32_t* list,
                                                          const unsigned long long* n_list, const uint32_t* probe_ids,
                                                          uint32_t* ob, uint32_t* op, uint64_t cap,
                                                          unsigned long long* counter) {
  SJ_WAVE_OUT(wo);
  const int lane = threadIdx.x & 63;
  const uint64_t nl = *n_list;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t rounds = (nl + nwaves - 1) / nwaves;  // (every wavefront runs the same rounds: emit is wave-wide)
  for (uint64_t t = 0; t < rounds; t++) {
    const uint64_t k = wave + t * nwaves;
    bool keep = false;
    uint32_t b = 0, p = 0;
    if (k < nl) {
      const uint32_t i = list[k];
      b = cb[i];
      p = cp[i];
      const double2 pt = c.P.xy[p];
      if (pt.x == pt.x || pt.y == pt.y) {
        const PolyDesc d = c.R.desc[b];
        int cls = 0;
        if (d.raster_off >= 0 && rect_covers(d.rect, pt.x, pt.y)) {
          const double cw = (d.rect.z - d.rect.x) / kRasterCells, ch = (d.rect.w - d.rect.y) / kRasterCells;
          const int ci = min(max((int)floor((pt.x - d.rect.x) / cw), 0), kRasterCells - 1);
          const int cj = min(max((int)floor((pt.y - d.rect.y) / ch), 0), kRasterCells - 1);
          const int cell = cj * kRasterCells + ci;
          cls = (c.R.raster[d.raster_off + (cell >> 4)] >> (2 * (cell & 15))) & 3;
        }
        if (cls) {
          keep = cls == 1;
        } else if (d.npts >= 0) {
          if (rect_covers(d.rect, pt.x, pt.y)) {
            RingRef r;
            r.xy = c.B.xy + d.v0;
            r.v0 = 0;
            r.npts = d.npts;
            const int64_t ns = d.npts >= 3 ? d.npts - 1 : 0;
            int on = 0, in = 0;
            uint64_t tests = 0;
            if (d.tree_off >= 0) {
              // lanes take the ring tree's leaves (level 0: one FP32
              // ring-local box per 2^ls segments) and scan those a +x ray
              // can meet, with the tree walk's margin of 2E
              const float4* lv0 = c.R.seg_tree + d.tree_off;
              const int ls = c.R.seg_leaf_shift;
              const int64_t nleaf = (ns + (1 << ls) - 1) >> ls;
              const double2 o = ring_origin(d.rect);
              const float pxl = (float)(pt.x - o.x), pyl = (float)(pt.y - o.y), m = 2.0f * d.err;
              for (int64_t lf = lane; lf < nleaf; lf += 64) {
                const float4 bx = lv0[lf];
                if (pyl >= bx.y - m && pyl <= bx.w + m && pxl <= bx.z + m) {
                  const int64_t s0 = lf << ls;
                  ring_segments_f64(r, s0, min(s0 + ((int64_t)1 << ls), ns), 1, pt.x, pt.y, on, in, tests);
                }
              }
            } else {
              ring_segments_f64(r, lane, ns, 64, pt.x, pt.y, on, in, tests);
            }
            const bool any_on = __ballot(on != 0) != 0;
            keep = any_on ? allow_on_edge != 0 : (__popcll(__ballot(in != 0)) & 1) != 0;
          }
        } else if (c.B.type[b] != kNull) {
          uint64_t tests = 0, fallbacks = 0;
          bool k0 = false;
          if (lane == 0) {
            const int64_t q1 = c.B.part_off[b + 1];
            for (int64_t q = c.B.part_off[b]; q < q1 && !k0; q++)
              k0 = build_poly_contains<1>(c, q, pt.x, pt.y, allow_on_edge != 0, tests, fallbacks);
          }
          keep = __shfl((int)k0, 0) != 0;
        }
      }
      if (keep && probe_ids) p = probe_ids[p];
    }
    emit(wo, lane == 0 && keep, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
}

// tg_ring_intersects_ring for one candidate with the whole wavefront: the
// lanes take the smaller-rect ring's segments with stride 64 and the
// wavefront votes.  Every lane returns the same value.
__device__ bool wave_iring_intersects(IRing ring, IRing other, int lane) {
  if (ring.npts < 3 || other.npts < 3) return false;
  double4 a = ring.rect, b = other.rect;
  if (b.x > a.z || b.z < a.x || b.y > a.w || b.w < a.y) return false;
  if ((b.z - b.x) * (b.w - b.y) > (a.z - a.x) * (a.w - a.y)) {
    IRing t = ring; ring = other; other = t;
    a = ring.rect;
  }
  uint64_t tests = 0, fallbacks = 0;
  const int64_t ns = other.npts - 1;
  for (int64_t base = 0; base < ns; base += 64) {
    const int64_t i = base + lane;
    bool hit = false;
    if (i < ns) {
      double2 p = other.xy[i], q = other.xy[i + 1];
      double sx0 = fmin0(p.x, q.x), sy0 = fmin0(p.y, q.y), sx1 = fmax0(p.x, q.x), sy1 = fmax0(p.y, q.y);
      if (!(a.x > sx1 || a.z < sx0 || a.y > sy1 || a.w < sy0))
        hit = ring_contains_run<0>(ring.xy, ring.q32, ring.tree, ring.ls, ring.npts, a, ring.o, ring.E, p.x, p.y, true, tests,
                                   fallbacks) ||
              ring_contains_run<0>(ring.xy, ring.q32, ring.tree, ring.ls, ring.npts, a, ring.o, ring.E, q.x, q.y, true, tests,
                                   fallbacks) ||
              ring_two_crossings(ring.xy, ring.npts, ring.tree, ring.ls, ring.E, ring.o, p.x, p.y, q.x, q.y);
    }
    if (__ballot(hit)) return true;
  }
  return false;
}

// Certified FP32 form of ring_two_crossings: 1 when two segments of the
// ring certainly intersect segment p-q (segment_intersects), 0 when fewer
// than two can, -1 otherwise.  The segment is taken into the ring's FP32
// frame (its rounding Ep added to the ring's bound E); tree boxes farther
// than the test's own bounding-box margin are pruned.
__device__ __forceinline__ int ring_two_crossings_cert(IRing ring, double px, double py, double qx,
                                                       double qy) {
  if (!(ring.E < INFINITY)) return -1;
  const int ns = ring.npts >= 3 ? (int)(ring.npts - 1) : 0;
  const float pxl = (float)(px - ring.o.x), pyl = (float)(py - ring.o.y);
  const float qxl = (float)(qx - ring.o.x), qyl = (float)(qy - ring.o.y);
  const float Ep = fmaxf(fmaxf(fabsf(pxl), fabsf(pyl)), fmaxf(fabsf(qxl), fabsf(qyl))) * 5.9604645e-08f * 1.0001f;
  const float e = ring.E + Ep;
  int yes = 0, unc = 0;
  const float2* q32 = ring.q32;
  if (!ring.tree) {
    for (int j = 0; j < ns && yes < 2; j++) {
      const float2 c = q32[j], d = q32[j + 1];
      const int r = segments_cross_f32_certified(pxl, pyl, qxl, qyl, c.x, c.y, d.x, d.y, e);
      yes += r == 1 ? 1 : 0;
      unc += r < 0 ? 1 : 0;
    }
  } else {
    // depth-first over the tree (byte l of `masks`: the children of level
    // l still to visit), the level geometry recomputed rather than held
    const float m = 2.0f * e + 4.0f * ring.E;
    const float x0 = fminf(pxl, qxl) - m, x1 = fmaxf(pxl, qxl) + m;
    const float y0 = fminf(pyl, qyl) - m, y1 = fmaxf(pyl, qyl) + m;
    const int ls = ring.ls;
    auto hits = [&](const float4* b, int n) {
      uint32_t mask = 0;
      for (int k = 0; k < n; k++) {
        const float4 bx = b[k];
        if (bx.x <= x1 && bx.z >= x0 && bx.y <= y1 && bx.w >= y0) mask |= 1u << k;
      }
      return mask;
    };
    const int top = seg_tree_shape(ns, ls).levels - 1;
    int otop, ntop;
    level_geom(ns, ls, top, otop, ntop);
    bool done = false;
    for (int c0 = 0; c0 < ntop && !done; c0 += kSegFanout) {
      uint64_t masks = (uint64_t)hits(ring.tree + otop + c0, min(kSegFanout, ntop - c0)) << (8 * top);
      int lvl = top, base = c0;
      while (masks && !done) {
        const uint32_t mk = (uint32_t)(masks >> (8 * lvl)) & 0xffu;
        if (mk == 0) {
          lvl++;
          base = (base >> 3) & ~(kSegFanout - 1);
          continue;
        }
        const int k = __ffs(mk) - 1;
        masks &= ~(1ull << (8 * lvl + k));
        const int node = base + k;
        if (lvl == 0) {
          const int s0 = node << ls, s1 = min(s0 + (1 << ls), ns);
          for (int j = s0; j < s1 && !done; j++) {
            const float2 c = q32[j], d = q32[j + 1];
            const int r = segments_cross_f32_certified(pxl, pyl, qxl, qyl, c.x, c.y, d.x, d.y, e);
            yes += r == 1 ? 1 : 0;
            unc += r < 0 ? 1 : 0;
            done = yes >= 2;
          }
          continue;
        }
        const int cb = node * kSegFanout;
        int loff, lcnt;
        level_geom(ns, ls, lvl - 1, loff, lcnt);
        const uint32_t cm = hits(ring.tree + loff + cb, min(kSegFanout, lcnt - cb));
        if (cm) {
          lvl--;
          base = cb;
          masks |= (uint64_t)cm << (8 * lvl);
        }
      }
    }
  }
  return yes >= 2 ? 1 : (yes + unc >= 2 ? -1 : 0);
}

// tg_ring_intersects_ring as wave_iring_intersects decides it, in two
// phases run by two kernels, with certified FP32 tests only.  The answer
// is an OR over the segments of the smaller-rect ring of two tests, so the
// order of evaluation does not change it:
// phase 1, one vertex of that ring per lane (every segment endpoint is a
// vertex of the closed ring) covered by the other ring;
// phase 2, only when no vertex is covered: the two-crossings count of each
// segment.

// The operands of a candidate: 1 when both rows are one polygon with one
// ring whose rects meet (ring: the larger-rect ring, other: the smaller),
// 0 when the pair certainly does not intersect (not both polygons, an empty
// part, disjoint rects, a ring of fewer than 3 points), 2 when the rows
// have several parts or rings (left to k_refine_polypoly).
__device__ __forceinline__ int pp_operands(const Ctx& c, uint32_t b, uint32_t p, IRing& ring, IRing& other) {
  if (!(is_polyish(c.B.type[b]) && is_polyish(c.P.type[p]))) return 0;
  if (c.B.part_off[b + 1] - c.B.part_off[b] != 1 || c.P.part_off[p + 1] - c.P.part_off[p] != 1) return 2;
  const int64_t qa = c.B.part_off[b], qb = c.P.part_off[p];
  const int64_t a0 = c.B.ring_off[qa], a1 = c.B.ring_off[qa + 1];
  const int64_t b0 = c.P.ring_off[qb], b1 = c.P.ring_off[qb + 1];
  if (a0 == a1 || b0 == b1) return 0;  // an empty part meets nothing
  if (a1 - a0 != 1 || b1 - b0 != 1) return 2;
  ring = iring(c.B, c.R, a0);
  other = iring(c.P, c.PR, b0);
  if (ring.npts < 3 || other.npts < 3) return 0;
  const double4 a = ring.rect, r = other.rect;
  if (r.x > a.z || r.z < a.x || r.y > a.w || r.w < a.y) return 0;
  if ((r.z - r.x) * (r.w - r.y) > (a.z - a.x) * (a.w - a.y)) {
    IRing t = ring; ring = other; other = t;
  }
  return 1;
}

// Phase 1: 1 a vertex is certainly covered, 0 none is, -1 none certainly
// is but some test was uncertain.
__device__ int wave_vertices_covered(const IRing& ring, const IRing& other, int lane) {
  const double4 a = ring.rect;
  const int64_t ns = other.npts - 1;
  bool unsure = false;
  for (int64_t base = 0; base < ns; base += 64) {
    const int64_t i = base + lane;
    bool hit = false, uns = false;
    if (i < ns) {
      const double2 p = other.xy[i];
      if (rect_covers(a, p.x, p.y)) {
        uint64_t tests = 0, fb = 0;
        const bool in = ring_contains_run<kPrecDeferred>(ring.xy, ring.q32, ring.tree, ring.ls, ring.npts, a, ring.o,
                                                         ring.E, p.x, p.y, true, tests, fb);
        hit = fb == 0 && in;
        uns = fb != 0;
      }
    }
    if (__ballot(hit)) return 1;
    unsure |= __ballot(uns) != 0;
  }
  return unsure ? -1 : 0;
}

// Phase 2: 1 some segment certainly crosses the ring twice, 0 none can, -1
// undecided.
__device__ int wave_two_crossings(const IRing& ring, const IRing& other, int lane) {
  const double4 a = ring.rect;
  const int64_t ns = other.npts - 1;
  bool unsure = false;
  for (int64_t base = 0; base < ns; base += 64) {
    const int64_t i = base + lane;
    bool hit = false, uns = false;
    if (i < ns) {
      const double2 p = other.xy[i], q = other.xy[i + 1];
      const double sx0 = fmin0(p.x, q.x), sy0 = fmin0(p.y, q.y), sx1 = fmax0(p.x, q.x), sy1 = fmax0(p.y, q.y);
      if (!(a.x > sx1 || a.z < sx0 || a.y > sy1 || a.w < sy0)) {
        const int r = ring_two_crossings_cert(ring, p.x, p.y, q.x, q.y);
        hit = r == 1;
        uns = r < 0;
      }
    }
    if (__ballot(hit)) return 1;
    unsure |= __ballot(uns) != 0;
  }
  return unsure ? -1 : 0;
}

// One pair appended to a list by lane 0 of the wavefront: list (lb, lp)
// with counter *cnt, at the front (k) or, with from_back, at cap - 1 - k.
__device__ __forceinline__ void pp_push(int lane, uint32_t b, uint32_t p, uint32_t* lb, uint32_t* lp,
                                        unsigned long long* cnt, uint64_t cap, bool from_back) {
  if (lane != 0) return;
  const unsigned long long k = atomicAdd(cnt, 1ull);
  const uint64_t at = from_back ? cap - 1 - k : k;
  lb[at] = b;
  lp[at] = p;
}

// Polygon x polygon ST_Intersects, certified form, phase 1 (one wavefront
// per candidate).  Pairs without a covered vertex go to (lb, lp): from the
// front (count cnt[0]) when every test was certain, from the back (count
// cnt[1]) when one was not; rows of several parts or rings go to (db, dp)
// (count cnt[2]).  Capacities: n each.
__global__ __launch_bounds__(kBlock) void k_pp_vertices(Ctx c, const uint32_t* cb, const uint32_t* cp, uint64_t n,
                                                        const unsigned long long* n_dev, uint32_t* ob, uint32_t* op,
                                                        uint64_t cap, unsigned long long* counter, uint32_t* lb,
                                                        uint32_t* lp, uint32_t* db, uint32_t* dp,
                                                        unsigned long long* cnt) {
  SJ_WAVE_OUT(wo);
  const uint64_t n0 = n;
  if (n_dev) n = min(n, (uint64_t)*n_dev);  // candidate count produced on the device (n: its bound)
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wave; i < n; i += nwaves) {
    const uint32_t b = cb[i], p = cp[i];
    IRing ring, other;
    const int k = pp_operands(c, b, p, ring, other);
    int r = 0;
    if (k == 1) {
      r = wave_vertices_covered(ring, other, lane);
      if (r == 0) pp_push(lane, b, p, lb, lp, &cnt[0], n0, false);
      if (r < 0) pp_push(lane, b, p, lb, lp, &cnt[1], n0, true);
    } else if (k == 2) {
      pp_push(lane, b, p, db, dp, &cnt[2], n0, false);
    }
    emit(wo, lane == 0 && k == 1 && r == 1, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
}

// Phase 2 over the pairs phase 1 listed (UNSURE: those with an uncertain
// vertex test, read from the back of the list): kept when a segment
// certainly crosses twice; else to (db, dp) when a test was uncertain in
// either phase, else dropped.
template <bool UNSURE>
__global__ __launch_bounds__(kBlock) void k_pp_crossings(Ctx c, const uint32_t* lb, const uint32_t* lp,
                                                         uint64_t cap_l, const unsigned long long* n_dev,
                                                         uint32_t* ob, uint32_t* op, uint64_t cap,
                                                         unsigned long long* counter, uint32_t* db, uint32_t* dp,
                                                         unsigned long long* n_defer, uint64_t cap_d) {
  SJ_WAVE_OUT(wo);
  const uint64_t n = min(cap_l, (uint64_t)*n_dev);
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wave; i < n; i += nwaves) {
    const uint64_t at = UNSURE ? cap_l - 1 - i : i;
    const uint32_t b = lb[at], p = lp[at];
    IRing ring, other;
    int r = 0;
    if (pp_operands(c, b, p, ring, other) == 1) r = wave_two_crossings(ring, other, lane);
    if (r < 0 || (UNSURE && r == 0)) pp_push(lane, b, p, db, dp, n_defer, cap_d, false);
    emit(wo, lane == 0 && r == 1, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
}

// Polygon x polygon ST_Intersects with rings indexed on both sides, one
// wavefront per candidate (single-ring parts voted by the wavefront;
// parts with holes decided by every lane alike).
__global__ __launch_bounds__(kBlock) void k_refine_polypoly(Ctx c, const uint32_t* cb, const uint32_t* cp,
                                                            uint64_t n, uint32_t* ob, uint32_t* op, uint64_t cap,
                                                            unsigned long long* counter,
                                                            const unsigned long long* n_dev) {
  SJ_WAVE_OUT(wo);
  if (n_dev) n = min(n, (uint64_t)*n_dev);  // candidate count produced on the device (n: its bound)
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wave; i < n; i += nwaves) {
    const uint32_t b = cb[i], p = cp[i];
    bool keep = false;
    if (is_polyish(c.B.type[b]) && is_polyish(c.P.type[p])) {
      for (int64_t qa = c.B.part_off[b]; qa < c.B.part_off[b + 1] && !keep; qa++)
        for (int64_t qb = c.P.part_off[p]; qb < c.P.part_off[p + 1] && !keep; qb++) {
          const int64_t a0 = c.B.ring_off[qa], a1 = c.B.ring_off[qa + 1];
          const int64_t b0 = c.P.ring_off[qb], b1 = c.P.ring_off[qb + 1];
          if (a0 == a1 || b0 == b1) continue;
          if (a1 - a0 == 1 && b1 - b0 == 1)
            keep = wave_iring_intersects(iring(c.B, c.R, a0), iring(c.P, c.PR, b0), lane);
          else
            keep = poly_intersects_poly(c.B, qa, c.P, qb, &c.R, &c.PR);
        }
    }
    emit(wo, lane == 0 && keep, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
}

// Whether the raster shows no point of the polygon (record d) in rect o,
// a part of its rect: every cell that o meets, widened by one cell, is
// clean outside (such a cell holds no boundary and lies outside).  Windows
// of more than 64 cells at 32 x 32 (the same share of the rect at other
// resolutions) are not read (false).
__device__ __forceinline__ bool raster_outside(const RingIndexView& R, const PolyDesc& d, double4 o) {
  if (d.raster_off < 0) return false;
  const double cw = (d.rect.z - d.rect.x) / kRasterCells, ch = (d.rect.w - d.rect.y) / kRasterCells;
  const int i0 = (int)fmax(floor((o.x - d.rect.x) / cw) - 1.0, 0.0);
  const int i1 = (int)fmin(floor((o.z - d.rect.x) / cw) + 1.0, kRasterCells - 1.0);
  const int j0 = (int)fmax(floor((o.y - d.rect.y) / ch) - 1.0, 0.0);
  const int j1 = (int)fmin(floor((o.w - d.rect.y) / ch) + 1.0, kRasterCells - 1.0);
  constexpr int kMaxWindow = 64 * (kRasterCells / 32) * (kRasterCells / 32);
  if (i0 > i1 || j0 > j1 || (i1 - i0 + 1) * (j1 - j0 + 1) > kMaxWindow) return false;
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1; i++) {
      const int cell = j * kRasterCells + i;
      if (((R.raster[d.raster_off + (cell >> 4)] >> (2 * (cell & 15))) & 3) != 2) return false;
    }
  return true;
}

// The class of a candidate's point for build row b.
__device__ __forceinline__ int pip_raster_class(const Ctx& c, uint32_t b, double2 pt) {
  return raster_class(c.R, c.R.desc[b], pt);
}

// Cell-raster pass of the polygon x polygon refine (ST_Intersects): a
// candidate is kept when it finds a point of both polygons in clean inside
// cells of both rasters (the centre of the rects' overlap, or one of four
// spread vertices of either ring inside the other), dropped when the
// rings' rects do not meet or one raster shows no point of its polygon in
// the rects' overlap, and else marked (as in k_raster_pass) for
// k_refine_polypoly.
__global__ __launch_bounds__(kBlock) void k_pp_raster_pass(Ctx c, const uint32_t* cb, const uint32_t* cp, uint64_t n,
                                                           uint32_t* ob, uint32_t* op, uint64_t cap,
                                                           unsigned long long* counter, uint64_t* mask,
                                                           unsigned long long* tile_cnt);

// Cell-raster pass of the point-in-polygon refine: candidates whose point
// falls in a clean cell of a single-ring polygon are decided here (kept or
// dropped); the others are marked for the walk in per-wavefront bit masks
// (mask[i / 64], bit i % 64) and counted per tile of kRasterTile
// consecutive candidates (tile_cnt).  k_open_scatter then copies them in
// candidate order, so that the walk keeps the polygon-major order of its
// input, reads its candidates with coalesced loads, and its wavefronts are
// not held back by candidates that need no walk.
constexpr int kRasterTile = 4096;

__global__ __launch_bounds__(kBlock) void k_raster_pass(Ctx c, const uint32_t* cb, const uint32_t* cp, uint64_t n,
                                                        const uint32_t* probe_ids, uint32_t* ob, uint32_t* op,
                                                        uint64_t cap, unsigned long long* counter, uint64_t* mask,
                                                        unsigned long long* tile_cnt) {
  SJ_WAVE_OUT(wo);
  __shared__ unsigned int block_cnt;
  const uint64_t n_tiles = (n + kRasterTile - 1) / kRasterTile;
  const int lane = threadIdx.x & 63;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    if (threadIdx.x == 0) block_cnt = 0;
    __syncthreads();
    unsigned int wave_cnt = 0;
    for (int it = 0; it < kRasterTile / kBlock; it++) {
      const uint64_t i = t * kRasterTile + (uint64_t)it * kBlock + threadIdx.x;
      bool keep = false, walk = false;
      uint32_t b = 0, p = 0;
      if (i < n) {
        b = cb[i];
        p = cp[i];
        const double2 pt = c.P.xy[p];
        if (pt.x == pt.x || pt.y == pt.y) {
          const int cls = pip_raster_class(c, b, pt);
          keep = cls == 1;
          walk = cls == 0;
        }
        if (keep && probe_ids) p = probe_ids[p];
      }
      emit(wo, keep, b, p, ob, op, cap, counter);
      // i - lane is a multiple of 64: one mask word per wavefront step
      const uint64_t m = __ballot(walk);
      if (lane == 0 && i < n) mask[i >> 6] = m;
      wave_cnt += __popcll(m);
    }
    if (lane == 0 && wave_cnt) atomicAdd(&block_cnt, wave_cnt);
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[t] = block_cnt;
  }
  wave_flush(wo, ob, op, cap, counter);
}

__global__ __launch_bounds__(kBlock) void k_pp_raster_pass(Ctx c, const uint32_t* cb, const uint32_t* cp, uint64_t n,
                                                           uint32_t* ob, uint32_t* op, uint64_t cap,
                                                           unsigned long long* counter, uint64_t* mask,
                                                           unsigned long long* tile_cnt) {
  SJ_WAVE_OUT(wo);
  __shared__ unsigned int block_cnt;
  const uint64_t n_tiles = (n + kRasterTile - 1) / kRasterTile;
  const int lane = threadIdx.x & 63;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    if (threadIdx.x == 0) block_cnt = 0;
    __syncthreads();
    unsigned int wave_cnt = 0;
    for (int it = 0; it < kRasterTile / kBlock; it++) {
      const uint64_t i = t * kRasterTile + (uint64_t)it * kBlock + threadIdx.x;
      bool keep = false, open = false;
      uint32_t b = 0, p = 0;
      if (i < n) {
        b = cb[i];
        p = cp[i];
        open = true;
        const PolyDesc db = c.R.desc[b], dp = c.PR.desc[p];
        if (db.npts >= 3 && dp.npts >= 3) {  // one ring each
          const double4 ra = db.rect, rb = dp.rect;
          if (rb.x > ra.z || rb.z < ra.x || rb.y > ra.w || rb.w < ra.y) {
            open = false;
          } else {
            // a point of both: the centre of the rects' overlap inside both,
            // or one of four spread vertices of either ring inside the other
            const double2 m = make_double2(0.5 * (fmax(ra.x, rb.x) + fmin(ra.z, rb.z)),
                                           0.5 * (fmax(ra.y, rb.y) + fmin(ra.w, rb.w)));
            keep = raster_class(c.R, db, m) == 1 && raster_class(c.PR, dp, m) == 1;
            for (int k = 0; k < 4 && !keep; k++) {
              keep = raster_class(c.R, db, c.P.xy[dp.v0 + (int64_t)k * (dp.npts - 1) / 4]) == 1 ||
                     raster_class(c.PR, dp, c.B.xy[db.v0 + (int64_t)k * (db.npts - 1) / 4]) == 1;
            }
            // no point of both: one polygon has none in the rects' overlap
            const double4 ov = make_double4(fmax(ra.x, rb.x), fmax(ra.y, rb.y), fmin(ra.z, rb.z), fmin(ra.w, rb.w));
            open = !keep && !raster_outside(c.R, db, ov) && !raster_outside(c.PR, dp, ov);
          }
        }
      }
      emit(wo, keep, b, p, ob, op, cap, counter);
      const uint64_t m = __ballot(open);
      if (lane == 0 && i < n) mask[i >> 6] = m;
      wave_cnt += __popcll(m);
    }
    if (lane == 0 && wave_cnt) atomicAdd(&block_cnt, wave_cnt);
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[t] = block_cnt;
  }
  wave_flush(wo, ob, op, cap, counter);
}

// Copies the candidates marked by a raster pass to (ob, op): tile t's
// entries start at tile_off[t] (exclusive scan of the tile counts) and keep
// candidate order.
__global__ __launch_bounds__(kBlock) void k_open_scatter(const uint64_t* mask, const unsigned long long* tile_off,
                                                         uint64_t n, const uint32_t* cb, const uint32_t* cp,
                                                         uint32_t* ob, uint32_t* op) {
  constexpr int kWords = kRasterTile / 64;
  __shared__ unsigned int word_off[kWords];
  const uint64_t n_tiles = (n + kRasterTile - 1) / kRasterTile;
  const uint64_t n_words = (n + 63) / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    __syncthreads();
    if (threadIdx.x < kWords) {
      const uint64_t w = t * kWords + threadIdx.x;
      word_off[threadIdx.x] = w < n_words ? __popcll(mask[w]) : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned int acc = 0;
      for (int j = 0; j < kWords; j++) {
        const unsigned int v = word_off[j];
        word_off[j] = acc;
        acc += v;
      }
    }
    __syncthreads();
    const unsigned long long base = tile_off[t];
    for (int j = wave; j < kWords; j += kBlock / 64) {
      const uint64_t w = t * kWords + j;
      if (w >= n_words) break;
      const uint64_t m = mask[w];
      if ((m >> lane) & 1) {
        const uint64_t i = w * 64 + lane, o = base + word_off[j] + __popcll(m & ((1ull << lane) - 1));
        ob[o] = cb[i];
        op[o] = cp[i];
      }
    }
  }
}

// Line-raster test of one candidate (single-line build row b, point pt,
// distance dist >= 0 and finite): false when no segment of the row can be
// within dist of the point -- no raster cell that the square
// [pt - dist, pt + dist] meets (widened by one cell) may hold a segment,
// or the square misses the line's rect, or the point has no coordinates.
// Rows without a raster answer true.
__device__ __forceinline__ bool line_may_be_within(const Ctx& c, uint32_t b, double2 pt, double dist) {
  return line_raster_may_hold(c.R, b, pt, dist);
}

// Cell-raster pass of the distance refine over single-line build rows and
// point probes: candidates that line_may_be_within rules out are dropped,
// the others marked for the walk as in k_raster_pass.
__global__ __launch_bounds__(kBlock) void k_dist_raster_pass(Ctx c, double dist, const uint32_t* cb,
                                                             const uint32_t* cp, uint64_t n, uint64_t* mask,
                                                             unsigned long long* tile_cnt) {
  __shared__ unsigned int block_cnt;
  const uint64_t n_tiles = (n + kRasterTile - 1) / kRasterTile;
  const int lane = threadIdx.x & 63;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    if (threadIdx.x == 0) block_cnt = 0;
    __syncthreads();
    unsigned int wave_cnt = 0;
    for (int it = 0; it < kRasterTile / kBlock; it++) {
      const uint64_t i = t * kRasterTile + (uint64_t)it * kBlock + threadIdx.x;
      bool walk = false;
      if (i < n) walk = line_may_be_within(c, cb[i], c.P.xy[cp[i]], dist);
      // i - lane is a multiple of 64: one mask word per wavefront step
      const uint64_t m = __ballot(walk);
      if (lane == 0 && i < n) mask[i >> 6] = m;
      wave_cnt += __popcll(m);
    }
    if (lane == 0 && wave_cnt) atomicAdd(&block_cnt, wave_cnt);
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[t] = block_cnt;
  }
}

// One lane per candidate; long-ring candidates are deferred into a list.
__global__ __launch_bounds__(kBlock) void k_refine_lane(Ctx c, int pred, int precision, const uint32_t* cb,
                                                        const uint32_t* cp, uint64_t n, uint32_t* ob, uint32_t* op,
                                                        uint64_t cap, unsigned long long* counter,
                                                        uint32_t* long_list, unsigned long long* long_counter,
                                                        unsigned long long* stats, int* bad) {
  SJ_WAVE_OUT(wo);
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t tests = 0, fallbacks = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    uint64_t i = base + threadIdx.x;
    bool keep = false;
    uint32_t b = 0, p = 0;
    if (i < n) {
      b = cb[i];
      p = cp[i];
      if (long_list && is_long(c, b, p)) {
        unsigned long long w = atomicAdd(long_counter, 1ull);
        long_list[w] = (uint32_t)i;
      } else {
        int r = eval_relation(c, b, p, pred, precision, tests, fallbacks);
        if (r < 0) atomicOr(bad, 1);
        keep = r > 0;
      }
    }
    emit(wo, keep, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
  // per-wave statistics
  for (int o = 32; o > 0; o >>= 1) {
    tests += __shfl_xor(tests, o);
    fallbacks += __shfl_xor(fallbacks, o);
  }
  if ((threadIdx.x & 63) == 0 && stats && (tests | fallbacks)) {
    atomicAdd(&stats[0], (unsigned long long)tests);
    atomicAdd(&stats[1], (unsigned long long)fallbacks);
  }
}

// The general relation refine: one lane per candidate, every supported
// pair of geometry types (the polygon x polygon relations with boundaries
// and the line intersections included).
__global__ __launch_bounds__(kBlock) void k_refine_relate(Ctx c, int pred, int precision, const uint32_t* cb,
                                                          const uint32_t* cp, uint64_t n, uint32_t* ob, uint32_t* op,
                                                          uint64_t cap, unsigned long long* counter,
                                                          unsigned long long* stats, int* bad) {
  SJ_WAVE_OUT(wo);
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t tests = 0, fallbacks = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    uint64_t i = base + threadIdx.x;
    bool keep = false;
    uint32_t b = 0, p = 0;
    if (i < n) {
      b = cb[i];
      p = cp[i];
      int r = eval_relate_extended(c, b, p, pred);
      if (r == -2) r = eval_relation(c, b, p, pred, precision, tests, fallbacks);
      if (r < 0) atomicOr(bad, 1);
      keep = r > 0;
    }
    emit(wo, keep, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
  for (int o = 32; o > 0; o >>= 1) {
    tests += __shfl_xor(tests, o);
    fallbacks += __shfl_xor(fallbacks, o);
  }
  if ((threadIdx.x & 63) == 0 && stats && (tests | fallbacks)) {
    atomicAdd(&stats[0], (unsigned long long)tests);
    atomicAdd(&stats[1], (unsigned long long)fallbacks);
  }
}

// One wavefront per long-ring candidate: lanes split the segments.
__global__ __launch_bounds__(kBlock) void k_refine_wave(Ctx c, int pred, int precision, const uint32_t* cb,
                                                        const uint32_t* cp, const uint32_t* list,
                                                        const unsigned long long* list_n, uint32_t* ob,
                                                        uint32_t* op, uint64_t cap, unsigned long long* counter,
                                                        unsigned long long* stats) {
  const int lane = threadIdx.x & 63;
  const uint64_t nlist = *list_n;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint64_t tests = 0, fallbacks = 0;
  for (uint64_t k = wave; k < nlist; k += nwaves) {
    uint64_t i = list[k];
    uint32_t b = cb[i], p = cp[i];
    double px, py;
    bool valid = point_of(c.P, p, px, py);
    bool allow = (pred == kIntersects || pred == kCovers);
    int64_t ring = c.B.ring_off[c.B.part_off[b]];
    bool keep = false;
    if (valid && (pred == kIntersects || pred == kCovers || pred == kContains)) {
      double4 rb = c.R.ring_rect[ring];
      if (rect_covers(rb, px, py)) {
        RingRef r = ring_at(c.B, ring);
        int64_t ns = r.npts - 1;
        int on = 0, in = 0;
        if (precision == 0) {
          double2 o = c.R.origin[ring];
          float E = c.R.err[ring];
          float pxl = (float)(px - o.x), pyl = (float)(py - o.y);
          ring_segments_f32(r, c.R.xy32, E, pxl, pyl, px, py, lane, ns, 64, on, in, tests, fallbacks);
        } else {
          ring_segments_f64(r, lane, ns, 64, px, py, on, in, tests);
        }
        bool any_on = __ballot(on) != 0;
        bool parity = (__popcll(__ballot(in)) & 1) != 0;
        keep = any_on ? allow : parity;
      }
    }
    if (lane == 0 && keep) {
      unsigned long long w = atomicAdd(counter, 1ull);
      if (w < cap) { ob[w] = b; op[w] = p; }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    tests += __shfl_xor(tests, o);
    fallbacks += __shfl_xor(fallbacks, o);
  }
  if (lane == 0 && stats) {
    atomicAdd(&stats[0], (unsigned long long)tests);
    atomicAdd(&stats[1], (unsigned long long)fallbacks);
  }
}

// ---- distance ----------------------------------------------------------

// Segment j of a vertex run is a term <= d of the reference distance (on
// the segment: 0, for lines; else distance_coord_to_line).  A segment
// whose bounding box is more than dm away on either axis is rejected first:
// dm = d (1 + 1e-4) + 4E exceeds d by more than the rounding of the
// distance computation, which is relative to the ring's extent (E =
// 2^-22 H) and to d.
__device__ __forceinline__ bool segment_within(const double2* xy, int64_t j, double px, double py, double d,
                                               double dm, bool check_on) {
  double2 a = xy[j], b = xy[j + 1];
  double gx = fmax(fmin(a.x, b.x) - px, px - fmax(a.x, b.x));
  double gy = fmax(fmin(a.y, b.y) - py, py - fmax(a.y, b.y));
  if (fmax(gx, gy) > dm) return false;
  if (check_on && line_intersects_coord(a.x, a.y, b.x, b.y, px, py)) return true;
  return coord_to_line(px, py, a.x, a.y, b.x, b.y) <= d;
}

// Whether some segment of a build ring (or line) is within d of the point.
// Segments whose FP32 tree boxes lie farther than d from the point, plus
// the boxes' and the point's FP32 error bounds (4E) and the rounding of d,
// cannot be within d and are skipped.  With the ring-local FP32 vertices,
// each remaining segment is decided by segment_within_f32, and only the
// uncertain ones (counted in fb) by the FP64 term.
__device__ bool ring_any_within_at(const Ctx& c, int64_t v0, int64_t npts, int64_t toff, float E, double2 o,
                                   double px, double py, double d, bool check_on, uint64_t& tests, uint64_t& fb) {
  const double2* xy = c.B.xy + v0;
  const int64_t ns = npts - 1;
  const int ls = c.R.seg_leaf_shift;
  const double dm = E < INFINITY ? d * 1.0001 + 4.0 * (double)E : INFINITY;
  if (toff < 0 || ns < 1 || !(E < INFINITY)) {
    for (int64_t j = 0; j < ns; j++) {
      tests++;
      if (segment_within(xy, j, px, py, d, dm, check_on)) return true;
    }
    return false;
  }
  const float D = (float)(d * 1.0001) + 4.0f * E;
  const float pxl = (float)(px - o.x), pyl = (float)(py - o.y);
  const float qx0 = pxl - D, qx1 = pxl + D, qy0 = pyl - D, qy1 = pyl + D;
  const int nsi = (int)ns;
  // the point's own rounding to local FP32 (its distance from the origin
  // can exceed the ring's half-extent by d)
  const float Ep = fmaxf(fabsf(pxl), fabsf(pyl)) * 5.9604645e-08f * 1.0001f;
  const float2* q32 = c.R.xy32 + v0;
  auto pass = [&](const float4& b) { return b.x <= qx1 && b.z >= qx0 && b.y <= qy1 && b.w >= qy0; };
  auto leaf = [&](int lf) {
    const int s0 = lf << ls, s1 = min(s0 + (1 << ls), nsi);
    for (int j = s0; j < s1; j++) {
      tests++;
      const float2 qa = q32[j], qb = q32[j + 1];
      const int r = segment_within_f32(qa.x, qa.y, qb.x, qb.y, pxl, pyl, E, Ep, d);
      if (r == 1) return true;
      if (r == 0) continue;
      fb++;
      if (segment_within(xy, j, px, py, d, dm, check_on)) return true;
    }
    return false;
  };
  return walk_any(c.R.seg_tree + toff, seg_tree_shape(nsi, ls), pass, leaf);
}

__device__ __forceinline__ bool ring_any_within(const Ctx& c, int64_t ring, double px, double py, double d,
                                                bool check_on, uint64_t& tests, uint64_t& fb) {
  const int64_t v0 = c.B.pt_off[ring];
  return ring_any_within_at(c, v0, c.B.pt_off[ring + 1] - v0, c.R.seg_tree_off ? c.R.seg_tree_off[ring] : -1,
                            c.R.err[ring], c.R.origin[ring], px, py, d, check_on, tests, fb);
}

// Position of the point w.r.t. a closed build ring, exactly: 2 boundary,
// 1 inside, 0 outside (geo's coord_pos_relative_to_ring).  The certified
// FP32 crossing count over the segment tree decides it when no test is
// uncertain (then no vertex or edge lies within its margin of the point's
// ray, so every crossing-number rule counts the same crossings); otherwise
// the exact scan does.
__device__ int build_ring_position(const Ctx& c, int64_t ring, double px, double py, uint64_t& exact_scans) {
  const int64_t v0 = c.B.pt_off[ring], npts = c.B.pt_off[ring + 1] - v0;
  const double4 rb = c.R.ring_rect[ring];
  if (!rect_covers(rb, px, py)) return 0;
  if (npts >= 4) {
    uint64_t tests = 0, fb = 0;
    const int64_t toff = c.R.seg_tree_off ? c.R.seg_tree_off[ring] : -1;
    bool in = ring_contains_at<kPrecDeferred>(c, rb, v0, npts, c.R.origin[ring], c.R.err[ring], toff, px, py, true,
                                              tests, fb);
    if (fb == 0) return in ? 1 : 0;
  }
  exact_scans++;
  return exact_ring_position(ring_at(c.B, ring), px, py);
}

// The reference decision distance(build row, point) <= d
// (rust/sedona-spatial-join/src/refine/geo.rs:312-335) without the
// distance itself: the distance is a minimum over parts and segments (0
// inside a polygon and on a line; NaN terms ignored), and such a minimum is
// <= d exactly when one of its terms is.
//
// KIND narrows the build column's geometry types at compile time (the
// kernel instance is chosen from the column's type mask), so that a
// column of lines does not pay the registers of the polygon branch.
enum DistKind : int { kDistAny = 0, kDistLines = 1, kDistPolygons = 2 };

template <int KIND>
__device__ __forceinline__ bool within_distance(const Ctx& c, int64_t row, double px, double py, double d,
                                                uint64_t& tests, uint64_t& exact_scans, uint64_t& fb) {
  if (!(d >= 0)) return false;  // every distance is >= 0 or NaN
  const GeomView& g = c.B;
  const int t = g.type[row];
  if constexpr (KIND == kDistAny) {
    if (g.points_only || t == kPoint || t == kMultiPoint || c.R.ring_rect == nullptr)
      return distance_to_point(g, row, px, py) <= d;
  }
  if constexpr (KIND == kDistLines) {
    // a single line: its record holds the vertex run, tree and frame
    const PolyDesc r = c.R.desc[row];
    if (r.npts >= 0) {
      if (r.npts < 2) return true;  // point_to_run: 0
      return ring_any_within_at(c, r.v0, r.npts, r.tree_off, r.err, ring_origin(r.rect), px, py, d, true, tests, fb);
    }
  }
  for (int64_t q = g.part_off[row]; q < g.part_off[row + 1]; q++) {
    const int64_t r0 = g.ring_off[q], r1 = g.ring_off[q + 1];
    if (KIND == kDistLines || (KIND == kDistAny && (t == kLine || t == kMultiLine))) {
      if (g.pt_off[r0 + 1] - g.pt_off[r0] < 2) return true;  // point_to_run: 0
      if (ring_any_within(c, r0, px, py, d, true, tests, fb)) return true;
      continue;
    }
    if constexpr (KIND == kDistLines) continue;
    if (r0 == r1) return true;  // point_to_polygon: 0
    int st = build_ring_position(c, r0, px, py, exact_scans);
    if (st == 2) return true;
    if (st == 1) {
      bool in_hole = false;
      for (int64_t h = r0 + 1; h < r1; h++) {
        int hs = build_ring_position(c, h, px, py, exact_scans);
        if (hs == 2) return true;
        if (hs == 1) in_hole = true;
      }
      if (!in_hole) return true;
    }
    for (int64_t k = r0; k < r1; k++)
      if (ring_any_within(c, k, px, py, d, false, tests, fb)) return true;
  }
  return false;
}

// Kernel instance for a build column's type mask.
inline int dist_kind(uint32_t type_mask, bool points_only, const RingIndex& ri) {
  if (points_only || ri.ring_rect.p == nullptr || ri.desc.p == nullptr) return kDistAny;
  const uint32_t lines = (1u << kLine) | (1u << kMultiLine), polys = (1u << kPolygon) | (1u << kMultiPolygon);
  if (type_mask && (type_mask & ~lines) == 0) return kDistLines;
  if (type_mask && (type_mask & ~polys) == 0) return kDistPolygons;
  return kDistAny;
}

// plain != 0: the all-FP64 form (the distance itself over every segment,
// then <=), for comparison.  (Forcing 8 waves per SIMD here, as for the
// point-in-polygon walk, measured slower: the spills cost more.)
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_refine_distance(Ctx c, double dist, int plain, const uint32_t* cb,
                                                            const uint32_t* cp, uint64_t n, const uint32_t* list,
                                                            uint32_t* ob, uint32_t* op, uint64_t cap,
                                                            unsigned long long* counter, unsigned long long* stats,
                                                            const unsigned long long* n_dev,
                                                            const uint32_t* probe_ids) {
  const GeomView& B = c.B;
  const GeomView& P = c.P;
  SJ_WAVE_OUT(wo);
  if (n_dev) n = min(n, (uint64_t)*n_dev);  // candidate count produced on the device (n: its bound)
  uint64_t tests = 0, exact_scans = 0, fb = 0;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    uint64_t i = base + threadIdx.x;
    bool keep = false;
    uint32_t b = 0, p = 0;
    if (i < n) {
      if (list) i = list[i];
      b = cb[i];
      p = cp[i];
      double px, py;
      bool bok = B.type[b] != kNull && (B.points_only || B.part_off[b] != B.part_off[b + 1]);
      if (bok && point_of(P, p, px, py))
        keep = plain ? distance_to_point(B, b, px, py) <= dist
                     : within_distance<KIND>(c, b, px, py, dist, tests, exact_scans, fb);
      if (keep && probe_ids) p = probe_ids[p];
    }
    emit(wo, keep, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
  for (int o = 32; o > 0; o >>= 1) {
    tests += __shfl_xor(tests, o);
    exact_scans += __shfl_xor(exact_scans, o);
    fb += __shfl_xor(fb, o);
  }
  if ((threadIdx.x & 63) == 0 && stats) {
    atomicAdd(&stats[0], (unsigned long long)tests);
    atomicAdd(&stats[1], (unsigned long long)exact_scans);
    atomicAdd(&stats[2], (unsigned long long)fb);
  }
}

// The reference distance of each (build row, probe point) pair, FP64
// (NaN for a null or empty row).
__global__ __launch_bounds__(kBlock) void k_pair_distance(GeomView B, GeomView P, const uint32_t* cb,
                                                          const uint32_t* cp, uint64_t n, double* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = cb[i], p = cp[i];
  double px, py, d = NAN;
  const bool bok = B.type[b] != kNull && (B.points_only || B.part_off[b] != B.part_off[b + 1]);
  if (bok && point_of(P, p, px, py)) d = distance_to_point(B, b, px, py);
  out[i] = d;
}

// Fused filter + distance refine: the filter walk over the rect tree keeps
// a rect hit (build row b, probe point qid) only when within_distance
// holds, so candidates are never written out.
template <int KIND>
struct KeepWithin {
  Ctx c;
  double d;
  unsigned long long* stats;  // [0] rect hits, [1] segment tests, [2] exact ring scans, [3] FP64 segment terms
  uint64_t hits_seen, tests, scans, fb;
  __device__ __forceinline__ uint32_t operator()(uint32_t hits, const uint32_t* ids, uint32_t qid) {
    hits_seen += __popc(hits);
    const double2 pt = c.P.xy[qid];  // a point column: null and empty rows hold NaN
    if (!(pt.x == pt.x || pt.y == pt.y)) return 0;
    uint32_t kept = 0;
    for (uint32_t m = hits; m; m &= m - 1) {
      const int k = __ffs(m) - 1;
      const uint32_t b = ids[k];
      const bool bok = c.B.type[b] != kNull && (c.B.points_only || c.B.part_off[b] != c.B.part_off[b + 1]);
      if (bok && within_distance<KIND>(c, b, pt.x, pt.y, d, tests, scans, fb)) kept |= 1u << k;
    }
    return kept;
  }
  __device__ __forceinline__ void finish() {
    for (int o = 32; o > 0; o >>= 1) {
      hits_seen += __shfl_xor(hits_seen, o);
      tests += __shfl_xor(tests, o);
      scans += __shfl_xor(scans, o);
      fb += __shfl_xor(fb, o);
    }
    if ((threadIdx.x & 63) == 0 && stats) {
      atomicAdd(&stats[0], (unsigned long long)hits_seen);
      atomicAdd(&stats[1], (unsigned long long)tests);
      atomicAdd(&stats[2], (unsigned long long)scans);
      atomicAdd(&stats[3], (unsigned long long)fb);
    }
  }
};

// ---- ring index ------------------------------------------------------------

__global__ void k_ring_rects(GeomView g, int64_t n_rings, double4* rr, double2* origin, float* err) {
  int64_t ring = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int lane = threadIdx.x & 63;
  if (ring >= n_rings) return;
  int64_t v0 = g.pt_off[ring], v1 = g.pt_off[ring + 1];
  double xl = INFINITY, yl = INFINITY, xh = -INFINITY, yh = -INFINITY;
  for (int64_t v = v0 + lane; v < v1; v += 64) {
    double2 p = g.xy[v];
    xl = fmin0(xl, p.x); xh = fmax0(xh, p.x);
    yl = fmin0(yl, p.y); yh = fmax0(yh, p.y);
  }
  for (int o = 32; o > 0; o >>= 1) {
    xl = fmin0(xl, __shfl_xor(xl, o));
    yl = fmin0(yl, __shfl_xor(yl, o));
    xh = fmax0(xh, __shfl_xor(xh, o));
    yh = fmax0(yh, __shfl_xor(yh, o));
  }
  if (lane == 0) {
    double4 rb = make_double4(xl, yl, xh, yh);
    rr[ring] = rb;
    double ox = 0.5 * xl + 0.5 * xh, oy = 0.5 * yl + 0.5 * yh;
    double H = fmax(xh - xl, yh - yl) * 0.5;
    float E;
    // ring-local FP32 coordinates are within 2^-23 H of exact; use 2^-22 H
    if (v1 - v0 < 3 || !(H > 1e-15) || !(H < 1e15) || ox != ox || oy != oy) E = INFINITY;
    else E = (float)(H * 2.384185791015625e-07) * 1.0001f;
    origin[ring] = ring_origin(rb);
    err[ring] = E;
  }
}

__global__ void k_local32(GeomView g, int64_t n_rings, const double2* origin, float2* xy32) {
  int64_t ring = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int lane = threadIdx.x & 63;
  if (ring >= n_rings) return;
  int64_t v0 = g.pt_off[ring], v1 = g.pt_off[ring + 1];
  double2 o = origin[ring];
  for (int64_t v = v0 + lane; v < v1; v += 64) {
    double2 p = g.xy[v];
    xy32[v] = make_float2((float)(p.x - o.x), (float)(p.y - o.y));
  }
}

// Boxes per ring (0 for rings scanned whole).
__global__ void k_seg_counts(GeomView g, int64_t n_rings, int min_segs, int ls, int64_t* cnt) {
  int64_t ring = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ring >= n_rings) return;
  int64_t npts = g.pt_off[ring + 1] - g.pt_off[ring];
  int64_t ns = npts >= 3 ? npts - 1 : 0;
  int64_t total = 0;
  if (ns >= min_segs && ns < (1ll << 30)) total = seg_tree_shape((int)ns, ls).total();
  cnt[ring] = total;
}

__global__ void k_seg_offsets(const int64_t* cnt, const int64_t* scan, int64_t n_rings, int64_t* off) {
  int64_t ring = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ring >= n_rings) return;
  off[ring] = cnt[ring] ? scan[ring] : -1;
}

// Level 0: one wavefront per ring, lanes over leaves.
__global__ void k_seg_leaves(GeomView g, int64_t n_rings, const float2* xy32, const int64_t* off, int ls,
                             float4* tree) {
  int64_t ring = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int lane = threadIdx.x & 63;
  if (ring >= n_rings) return;
  int64_t o = off[ring];
  if (o < 0) return;
  int64_t v0 = g.pt_off[ring];
  int64_t ns = g.pt_off[ring + 1] - v0 - 1;
  int64_t nleaf = (ns + (1 << ls) - 1) >> ls;
  for (int64_t k = lane; k < nleaf; k += 64) {
    int64_t a = k << ls, e = min(a + (1 << ls), ns);  // vertices a..e
    float2 p = xy32[v0 + a];
    float4 b = make_float4(p.x, p.y, p.x, p.y);
    for (int64_t v = a + 1; v <= e; v++) {
      p = xy32[v0 + v];
      b.x = fminf(b.x, p.x); b.y = fminf(b.y, p.y);
      b.z = fmaxf(b.z, p.x); b.w = fmaxf(b.w, p.y);
    }
    tree[o + k] = b;
  }
}

// Upper levels: one thread per ring.
__global__ void k_seg_upper(GeomView g, int64_t n_rings, const int64_t* off, int ls, float4* tree) {
  int64_t ring = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ring >= n_rings) return;
  int64_t o = off[ring];
  if (o < 0) return;
  int64_t ns = g.pt_off[ring + 1] - g.pt_off[ring] - 1;
  TreeShape sh = seg_tree_shape((int)ns, ls);
  int64_t c[kSegMaxLevels] = {sh.c0, sh.c1, sh.c2, sh.c3, sh.c4};
  int64_t lo[kSegMaxLevels] = {sh.off<0>(), sh.off<1>(), sh.off<2>(), sh.off<3>(), sh.off<4>()};
  int L = sh.levels;
  for (int l = 1; l < L; l++) {
    const float4* ch = tree + o + lo[l - 1];
    float4* pa = tree + o + lo[l];
    for (int64_t i = 0; i < c[l]; i++) {
      int64_t c0 = i * kSegFanout, c1 = min(c0 + kSegFanout, c[l - 1]);
      float4 u = ch[c0];
      for (int64_t k = c0 + 1; k < c1; k++) {
        float4 b = ch[k];
        u.x = fminf(u.x, b.x); u.y = fminf(u.y, b.y);
        u.z = fmaxf(u.z, b.z); u.w = fmaxf(u.w, b.w);
      }
      pa[i] = u;
    }
  }
}

unsigned int blocks_for(uint64_t n) { return (unsigned int)((n + kBlock - 1) / kBlock); }

unsigned int grid_for(uint64_t n) {
  uint64_t b = blocks_for(n);
  return (unsigned int)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 16));
}

// Single-ring build rows (one polygon part with one ring, or one line) as
// one 64-byte record, so that the refine kernels reach a row's vertex run
// with one load.
__global__ void k_poly_desc(GeomView g, RingIndexView R, PolyDesc* out) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.n_rows) return;
  PolyDesc d;
  d.rect = make_double4(NAN, NAN, NAN, NAN);
  d.v0 = 0;
  d.tree_off = -1;
  d.npts = -1;
  d.err = INFINITY;
  d.raster_off = -1;
  const int t = g.type[r];
  if ((is_polyish(t) || t == kLine || t == kMultiLine) && g.part_off[r + 1] - g.part_off[r] == 1) {
    int64_t q = g.part_off[r];
    if (g.ring_off[q + 1] - g.ring_off[q] == 1) {
      int64_t ring = g.ring_off[q];
      int64_t v0 = g.pt_off[ring], n = g.pt_off[ring + 1] - v0;
      if (n <= 0x7fffffff) {
        d.rect = R.ring_rect[ring];
        d.v0 = v0;
        d.npts = (int32_t)n;
        d.err = R.err[ring];
        d.tree_off = R.seg_tree_off[ring];
      }
    }
  }
  out[r] = d;
}

// Rows that get a cell raster: one polygon ring (with or without a
// segment tree) with an FP32 frame and a rect of positive width and height.
__global__ void k_raster_flags(GeomView g, const PolyDesc* desc, int64_t* flag) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.n_rows) return;
  const PolyDesc d = desc[r];
  flag[r] = is_polyish(g.type[r]) && d.npts >= 4 && d.rect.z > d.rect.x && d.rect.w > d.rect.y &&
                    d.err < INFINITY
                ? 1
                : 0;
}

// Cell raster of one ring per wavefront.  A cell is left 0 when a segment's
// bounding box meets it or one of its 8 neighbours; so the point-in-ring
// result is the same at every point within one cell of a marked-clean
// cell: the reference's raycast decides each segment there by exact
// comparisons (the point lies outside every segment's box), and the
// crossing parity is constant off the boundary.  Clean cells get their
// centre's result (the certified FP32 walk, FP64 where uncertain).
__global__ __launch_bounds__(kBlock) void k_raster_build(Ctx c, PolyDesc* desc, const int64_t* flag,
                                                        const int64_t* roff, uint32_t* raster) {
  __shared__ uint32_t s_touch[kBlock / 64][kRasterCells * kRasterCells / 32];
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= c.B.n_rows || !flag[row]) return;  // wave-uniform
  uint32_t* touch = s_touch[threadIdx.x >> 6];
  for (int k = lane; k < kRasterCells * kRasterCells / 32; k += 64) touch[k] = 0;
  wave_lds_sync();
  PolyDesc d = desc[row];
  const double4 rb = d.rect;
  const double cw = (rb.z - rb.x) / kRasterCells, ch = (rb.w - rb.y) / kRasterCells;
  const double2* xy = c.B.xy + d.v0;
  const int64_t ns = d.npts - 1;
  for (int64_t j = lane; j < ns; j += 64) {
    const double2 a = xy[j], b = xy[j + 1];
    int i0 = (int)floor((fmin(a.x, b.x) - rb.x) / cw) - 1, i1 = (int)floor((fmax(a.x, b.x) - rb.x) / cw) + 1;
    int j0 = (int)floor((fmin(a.y, b.y) - rb.y) / ch) - 1, j1 = (int)floor((fmax(a.y, b.y) - rb.y) / ch) + 1;
    i0 = max(i0, 0); j0 = max(j0, 0);
    i1 = min(i1, kRasterCells - 1); j1 = min(j1, kRasterCells - 1);
    for (int cy = j0; cy <= j1; cy++)
      for (int cx = i0; cx <= i1; cx++) {
        const int cell = cy * kRasterCells + cx;
        atomicOr(&touch[cell >> 5], 1u << (cell & 31));
      }
  }
  wave_lds_sync();
  // lane k writes words k, k + 64, ...: 16 cells each
  const double2 o = ring_origin(rb);
  const float4* tree = d.tree_off >= 0 ? c.R.seg_tree + d.tree_off : nullptr;  // (short rings: scanned whole)
  const float2* q32 = c.R.xy32 + d.v0;
  for (int wi = lane; wi < kRasterWords; wi += 64) {
    uint32_t word = 0;
    uint64_t t = 0, f = 0;
    for (int k = 0; k < 16; k++) {
      const int cell = wi * 16 + k;
      if ((touch[cell >> 5] >> (cell & 31)) & 1u) continue;
      const double px = rb.x + ((cell % kRasterCells) + 0.5) * cw, py = rb.y + ((cell / kRasterCells) + 0.5) * ch;
      const bool in = ring_contains_run<0>(xy, q32, tree, c.R.seg_leaf_shift, d.npts, rb, o, d.err, px, py, true, t, f);
      word |= (in ? 1u : 2u) << (2 * k);
    }
    raster[roff[row] * kRasterWords + wi] = word;
  }
  if (lane == 0) desc[row].raster_off = roff[row] * kRasterWords;
}

__global__ void k_line_raster_flags(GeomView g, const PolyDesc* desc, int64_t* flag) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.n_rows) return;
  const PolyDesc d = desc[r];
  const double w = d.rect.z - d.rect.x, h = d.rect.w - d.rect.y;
  flag[r] = (g.type[r] == kLine || g.type[r] == kMultiLine) && d.npts >= 2 && w > 0 && h > 0 && w < INFINITY &&
                    h < INFINITY
                ? 1
                : 0;
}

// Cell index range [lo, hi] covering [v0, v1] (in cells from the grid's
// origin), widened by one cell and clamped to the grid; NaN gives cell 0.
__device__ __forceinline__ int line_cell_lo(double v) { return (int)fmin(fmax(floor(v) - 1.0, 0.0), kLineCells - 1.0); }
__device__ __forceinline__ int line_cell_hi(double v) { return (int)fmin(fmax(floor(v) + 1.0, 0.0), kLineCells - 1.0); }

// Cell raster of one line per wavefront.  A cell's bit is set when it lies
// in a segment's cell range (its bounding box, widened by one cell) and
// its centre is within 1.01 half diagonals of the segment's line (of its
// point, for a zero-length segment): every cell a segment passes through
// is set, with a margin far above the rounding of these tests.  Segments
// with a NaN coordinate set no cell; they are never within a distance.
__global__ __launch_bounds__(kBlock) void k_line_raster_build(GeomView g, PolyDesc* desc, const int64_t* flag,
                                                             const int64_t* roff, uint64_t* raster) {
  __shared__ unsigned long long s_words[kBlock / 64][kLineWords];
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= g.n_rows || !flag[row]) return;  // wave-uniform
  unsigned long long* words = s_words[threadIdx.x >> 6];
  for (int k = lane; k < kLineWords; k += 64) words[k] = 0;
  wave_lds_sync();
  const PolyDesc d = desc[row];
  const double4 rb = d.rect;
  const double cw = (rb.z - rb.x) / kLineCells, ch = (rb.w - rb.y) / kLineCells;
  const double r = 0.505 * sqrt(cw * cw + ch * ch);
  const double2* xy = g.xy + d.v0;
  for (int64_t j = lane; j < (int64_t)d.npts - 1; j += 64) {
    const double2 a = xy[j], b = xy[j + 1];
    const int i0 = line_cell_lo((fmin(a.x, b.x) - rb.x) / cw), i1 = line_cell_hi((fmax(a.x, b.x) - rb.x) / cw);
    const int j0 = line_cell_lo((fmin(a.y, b.y) - rb.y) / ch), j1 = line_cell_hi((fmax(a.y, b.y) - rb.y) / ch);
    const double ex = b.x - a.x, ey = b.y - a.y;
    const double len = sqrt(ex * ex + ey * ey);
    for (int cy = j0; cy <= j1; cy++) {
      const double yc = rb.y + (cy + 0.5) * ch;
      unsigned long long bits[kLineRowWords] = {};
      for (int cx = i0; cx <= i1; cx++) {
        const double xc = rb.x + (cx + 0.5) * cw;
        const double dl = len > 0 ? fabs(ex * (yc - a.y) - ey * (xc - a.x)) / len
                                  : sqrt((xc - a.x) * (xc - a.x) + (yc - a.y) * (yc - a.y));
        if (dl <= r) bits[cx >> 6] |= 1ull << (cx & 63);
      }
#pragma unroll
      for (int w = 0; w < kLineRowWords; w++)
        if (bits[w]) atomicOr(&words[cy * kLineRowWords + w], bits[w]);
    }
  }
  wave_lds_sync();
  for (int k = lane; k < kLineWords; k += 64) raster[roff[row] * kLineWords + k] = words[k];
  if (lane == 0) desc[row].raster_off = roff[row] * kLineWords;
}

}  // namespace

void build_ring_index(const GeomColumn& g, RingIndex& ri, const RingIndexParams& prm, hipStream_t s) {
  ri = RingIndex();
  if (g.points_only || g.n_rings == 0) return;
  ri.ring_rect.alloc(g.n_rings, s);
  ri.origin.alloc(g.n_rings, s);
  ri.err.alloc(g.n_rings, s);
  ri.xy32.alloc(g.n_pts > 0 ? g.n_pts : 1, s);
  uint64_t threads = (uint64_t)g.n_rings * 64;
  hipLaunchKernelGGL(k_ring_rects, blocks_for(threads), kBlock, 0, s, g.view(), g.n_rings, ri.ring_rect.p,
                     ri.origin.p, ri.err.p);
  hipLaunchKernelGGL(k_local32, blocks_for(threads), kBlock, 0, s, g.view(), g.n_rings, ri.origin.p, ri.xy32.p);
  // segment trees
  DevArray<int64_t> cnt(g.n_rings, s);
  if (prm.seg_leaf_shift != kSegLeafShift && prm.seg_leaf_shift != kSegLeafShift + 1)
    throw Error("segment tree leaves must hold 16 or 32 segments");
  ri.seg_leaf_shift = prm.seg_leaf_shift;
  hipLaunchKernelGGL(k_seg_counts, blocks_for(g.n_rings), kBlock, 0, s, g.view(), g.n_rings, prm.min_segs_for_tree,
                     ri.seg_leaf_shift, cnt.p);
  DevArray<int64_t> scan(g.n_rings + 1, s);
  size_t tmp = 0;
  SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt.p, scan.p, (int)g.n_rings, s));
  DevArray<uint8_t> t(tmp + 16, s);
  SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, cnt.p, scan.p, (int)g.n_rings, s));
  int64_t last[2];
  SJ_CHECK(hipMemcpyAsync(&last[0], scan.p + g.n_rings - 1, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipMemcpyAsync(&last[1], cnt.p + g.n_rings - 1, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  ri.n_boxes = last[0] + last[1];
  ri.seg_tree_off.alloc(g.n_rings, s);
  hipLaunchKernelGGL(k_seg_offsets, blocks_for(g.n_rings), kBlock, 0, s, cnt.p, scan.p, g.n_rings,
                     ri.seg_tree_off.p);
  ri.seg_tree.alloc(ri.n_boxes > 0 ? ri.n_boxes : 1, s);
  if (ri.n_boxes > 0) {
    hipLaunchKernelGGL(k_seg_leaves, blocks_for(threads), kBlock, 0, s, g.view(), g.n_rings, ri.xy32.p,
                       ri.seg_tree_off.p, ri.seg_leaf_shift, ri.seg_tree.p);
    hipLaunchKernelGGL(k_seg_upper, blocks_for(g.n_rings), kBlock, 0, s, g.view(), g.n_rings, ri.seg_tree_off.p,
                       ri.seg_leaf_shift, ri.seg_tree.p);
  }
  ri.desc.alloc(g.n_rows > 0 ? g.n_rows : 1, s);
  if (g.n_rows > 0) hipLaunchKernelGGL(k_poly_desc, blocks_for(g.n_rows), kBlock, 0, s, g.view(), ri.view(), ri.desc.p);
  SJ_CHECK(hipGetLastError());
  // cell rasters of the single-ring polygon rows with segment trees, and
  // of the single-line rows
  if (g.n_rows > 0 && prm.raster) {
    // rows flagged by `flags` get consecutive raster slots (roff); returns
    // their count
    DevArray<int64_t> flag(g.n_rows + 1, s), roff(g.n_rows + 1, s);
    auto slots = [&](auto flags) {
      SJ_CHECK(hipMemsetAsync(flag.p + g.n_rows, 0, 8, s));
      hipLaunchKernelGGL(flags, blocks_for(g.n_rows), kBlock, 0, s, g.view(), ri.desc.p, flag.p);
      // the scan runs over n_rows + 1 entries so that roff[n_rows] is the total
      size_t tb = 0;
      SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag.p, roff.p, (int)g.n_rows + 1, s));
      DevArray<uint8_t> tmp(tb + 16, s);
      SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flag.p, roff.p, (int)g.n_rows + 1, s));
      int64_t total = 0;
      SJ_CHECK(hipMemcpyAsync(&total, roff.p + g.n_rows, 8, hipMemcpyDeviceToHost, s));
      SJ_CHECK(hipStreamSynchronize(s));
      return total;
    };
    const uint32_t polys = (1u << kPolygon) | (1u << kMultiPolygon), lines = (1u << kLine) | (1u << kMultiLine);
    if (g.type_mask & polys) {
      ri.n_rasters = slots(k_raster_flags);
      if (ri.n_rasters > 0) {
        ri.raster.alloc((size_t)ri.n_rasters * kRasterWords, s);
        Ctx c;
        c.B = g.view();
        c.R = ri.view();
        hipLaunchKernelGGL(k_raster_build, blocks_for((uint64_t)g.n_rows * 64), kBlock, 0, s, c, ri.desc.p, flag.p,
                           roff.p, ri.raster.p);
        SJ_CHECK(hipGetLastError());
      }
    }
    if (g.type_mask & lines) {
      ri.n_line_rasters = slots(k_line_raster_flags);
      if (ri.n_line_rasters > 0) {
        ri.line_raster.alloc((size_t)ri.n_line_rasters * kLineWords, s);
        hipLaunchKernelGGL(k_line_raster_build, blocks_for((uint64_t)g.n_rows * 64), kBlock, 0, s, g.view(),
                           ri.desc.p, flag.p, roff.p, ri.line_raster.p);
        SJ_CHECK(hipGetLastError());
      }
    }
  }
  SJ_CHECK(hipStreamSynchronize(s));
}

void check_relation_types(uint32_t bm, uint32_t pm, int pred) {
  const uint32_t poly = (1u << kPolygon) | (1u << kMultiPolygon), pt = 1u << kPoint;
  const uint32_t line = (1u << kLine) | (1u << kMultiLine);
  bool ok = true;
  auto only = [](uint32_t m, uint32_t allowed) { return (m & ~allowed) == 0; };
  switch (pred) {
    case kIntersects:
      // points, (multi)linestrings and (multi)polygons in any combination
      ok = only(bm, poly | pt | line) && only(pm, poly | pt | line);
      break;
    case kContains: case kWithin: case kCovers: case kCoveredBy: case kTouches: case kEquals:
      // point x point, point x (multi)polygon in both roles, and (multi)polygon x (multi)polygon
      ok = only(bm, poly | pt) && only(pm, poly | pt);
      break;
    default:
      ok = false;
  }
  if (!ok)
    throw Error("unsupported predicate / geometry type combination for the GPU refine (predicate " +
                std::to_string(pred) + ")");
}

bool pip_refine_applies(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, int predicate) {
  const uint32_t polyish = (1u << kPolygon) | (1u << kMultiPolygon);
  return !build.points_only && probe.points_only && ri.ring_rect.p != nullptr && ri.desc.p != nullptr &&
         (build.type_mask & ~polyish) == 0 &&
         (predicate == kIntersects || predicate == kCovers || predicate == kContains);
}

// Runs a raster pass over the n candidates (cb, cp) (launch(grid, mask,
// tile_counts)) and copies the candidates it marked, in candidate order, to
// open (mask: the pass's marks); *count (device memory, in count_store)
// receives their number.
template <class Launch>
void list_marked(uint64_t n, Launch launch, const uint32_t* cb, const uint32_t* cp, PairBuf& open,
                 DevArray<uint64_t>& mask, DevArray<unsigned long long>& count_store,
                 const unsigned long long** count, hipStream_t s) {
  const uint64_t n_tiles = (n + kRasterTile - 1) / kRasterTile;
  fit(mask, (n + 63) / 64, s);
  DevArray<unsigned long long> cnt(n_tiles + 1, s);
  count_store.alloc(n_tiles + 1, s);
  fit(open.b, n, s);
  fit(open.p, n, s);
  SJ_CHECK(hipMemsetAsync(cnt.p + n_tiles, 0, 8, s));
  const unsigned int g = (unsigned int)std::min<uint64_t>(n_tiles, 4096);
  launch(g, mask.p, cnt.p);
  size_t tb = 0;
  SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.p, count_store.p, (int)(n_tiles + 1), s));
  DevArray<uint8_t> tmp(tb, s);
  SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt.p, count_store.p, (int)(n_tiles + 1), s));
  hipLaunchKernelGGL(k_open_scatter, g, kBlock, 0, s, mask.p, count_store.p, n, cb, cp, open.b.p, open.p.p);
  SJ_CHECK(hipGetLastError());
  // the scan's last entry is the number of marked candidates
  *count = count_store.p + n_tiles;
}

bool raster_split(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, int predicate,
                  int precision, const uint32_t* cb, const uint32_t* cp, uint64_t n, const uint32_t* probe_ids,
                  PairBuf& out, PairBuf& open, DevArray<uint64_t>& mask, uint64_t* n_open, hipStream_t s) {
  if (ri.n_rasters == 0 || precision != 0 || n > 0xffffffffull || build.points_only || ri.ring_rect.p == nullptr ||
      !pip_refine_applies(build, ri, probe, predicate))
    return false;
  out.counter.reserve(1, s);
  fit(out.b, n > 0 ? n : 1, s);
  fit(out.p, n > 0 ? n : 1, s);
  SJ_CHECK(hipMemsetAsync(out.counter.p, 0, 8, s));
  *n_open = 0;
  if (n == 0) return true;
  Ctx c;
  c.B = build.view();
  c.R = ri.view();
  c.P = probe.view();
  DevArray<unsigned long long> open_count;
  const unsigned long long* count = nullptr;
  list_marked(
      n,
      [&](unsigned int g, uint64_t* marks, unsigned long long* tile_cnt) {
        hipLaunchKernelGGL(k_raster_pass, g, kBlock, 0, s, c, cb, cp, n, probe_ids, out.b.p, out.p.p, out.cap(),
                           out.counter.p, marks, tile_cnt);
      },
      cb, cp, open, mask, open_count, &count, s);
  unsigned long long h = 0;
  SJ_CHECK(hipMemcpyAsync(&h, count, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  *n_open = h;
  return true;
}

bool line_rasters_apply(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                        int precision) {
  return precision != 1 && dist_kind(build.type_mask, build.points_only, ri) == kDistLines && ri.n_line_rasters > 0 &&
         probe.points_only && d >= 0 && d < INFINITY;
}

int dist_raster_kind(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d, int precision) {
  if (line_rasters_apply(build, ri, probe, d, precision)) return 1;
  if (precision == 0 && dist_kind(build.type_mask, build.points_only, ri) == kDistPolygons && ri.n_rasters > 0 &&
      probe.points_only && d >= 0 && d < INFINITY)
    return 2;
  return 0;
}

bool raster_split_distance(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                           int precision, const uint32_t* cb, const uint32_t* cp, uint64_t n, PairBuf& open,
                           DevArray<uint64_t>& mask, uint64_t* n_open, hipStream_t s) {
  if (precision == 1 || dist_kind(build.type_mask, build.points_only, ri) != kDistLines || ri.n_line_rasters == 0 ||
      !probe.points_only || !(d >= 0 && d < INFINITY) || n > 0xffffffffull)
    return false;
  *n_open = 0;
  if (n == 0) return true;
  Ctx c;
  c.B = build.view();
  c.R = ri.view();
  c.P = probe.view();
  DevArray<unsigned long long> open_count;
  const unsigned long long* count = nullptr;
  list_marked(
      n,
      [&](unsigned int g, uint64_t* marks, unsigned long long* tile_cnt) {
        hipLaunchKernelGGL(k_dist_raster_pass, g, kBlock, 0, s, c, d, cb, cp, n, marks, tile_cnt);
      },
      cb, cp, open, mask, open_count, &count, s);
  unsigned long long h = 0;
  SJ_CHECK(hipMemcpyAsync(&h, count, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  *n_open = h;
  return true;
}

uint64_t refine_relation(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, int predicate,
                         const uint32_t* cb, const uint32_t* cp, uint64_t n, int precision, PairBuf& out,
                         RefineStats* stats, hipStream_t s, KernelTimer* timer, const uint32_t* probe_ids,
                         int pip_waves, Workspace* ws, bool raster_done) {
  out.counter.reserve(1, s);
  fit(out.b, n > 0 ? n : 1, s);
  fit(out.p, n > 0 ? n : 1, s);
  if (n == 0) {
    if (!raster_done) return 0;
    // a raster split decided every candidate: out holds the kept ones
    unsigned long long cnt = 0;
    SJ_CHECK(hipMemcpyAsync(&cnt, out.counter.p, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    return cnt;
  }
  if (build.points_only || ri.ring_rect.p == nullptr) precision = 1;
  Ctx c;
  c.B = build.view();
  c.R = ri.view();
  c.P = probe.view();
  DevArray<unsigned long long> aux(6, s);  // counter of long list / deferred, stats[2], bad flag, rest count, queue
  DevArray<uint32_t> long_list;
  bool pip = pip_refine_applies(build, ri, probe, predicate);
  if (probe_ids && !pip) throw Error("probe id mapping needs the point-in-polygon refine");
  if (raster_done && !(pip && precision == 0)) throw Error("raster split without the certified point-in-polygon refine");
  // after raster_split, out already holds the decided pairs
  if (!raster_done) SJ_CHECK(hipMemsetAsync(out.counter.p, 0, 8, s));
  SJ_CHECK(hipMemsetAsync(aux.p, 0, 48, s));
  if (pip) {
    int allow = predicate == kContains ? 0 : 1;
    unsigned long long walked = n;  // candidates the walk decides (host copy)
    if (timer) timer->record_start(s);
    if (precision == 0 && n <= 0xffffffffull) {
      // certified FP32 pass; candidates with an uncertain test (aux[0] of
      // them) are refined again in FP64 by the second launch
      DevArray<uint32_t> redo_own, rest_own;
      PairBuf open_own;
      DevArray<uint64_t> mask_own;
      DevArray<uint32_t>& redo = ws ? ws->redo : redo_own;
      DevArray<uint32_t>& rest = ws ? ws->rest : rest_own;
      PairBuf& open = ws ? ws->open : open_own;
      DevArray<uint64_t>& mask = ws ? ws->mask : mask_own;
      fit(redo, n, s);
      fit(rest, n, s);
      // candidates in clean raster cells are decided first; the walk (and
      // the FP64 pass after it) take the rest, copied in order to `open`
      DevArray<unsigned long long> open_count;
      const unsigned long long* n_open = nullptr;
      if (ri.n_rasters > 0 && !raster_done)
        list_marked(
            n,
            [&](unsigned int g, uint64_t* marks, unsigned long long* tile_cnt) {
              hipLaunchKernelGGL(k_raster_pass, g, kBlock, 0, s, c, cb, cp, n, probe_ids, out.b.p, out.p.p,
                                 out.cap(), out.counter.p, marks, tile_cnt);
            },
            cb, cp, open, mask, open_count, &n_open, s);
      const uint32_t* wb = n_open ? open.b.p : cb;
      const uint32_t* wp = n_open ? open.p.p : cp;
      // single-ring rows: the lean walk (one ray direction per wavefront);
      // the other rows (listed in `rest`, count aux[4]): the general kernel.
      // Both list the candidates they defer in redo (count aux[0]).
      launch_ring_walk(ri, probe, wb, wp, n, n_open, probe_ids, raster_done || n_open != nullptr, out.b.p, out.p.p,
                       out.cap(), out.counter.p, aux.p + 1, redo.p, aux.p, rest.p, aux.p + 4, pip_waves, s,
                       aux.p + 5);
      // (a grid-stride launch of 2048 blocks: the list is often empty)
      auto fp32_pass = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, std::min(grid_for(n), 2048u), kBlock, 0, s, c, allow, wb, wp, n, rest.p, out.b.p, out.p.p,
                           out.cap(), out.counter.p, aux.p + 1, probe_ids, redo.p, aux.p, aux.p + 4);
      };
      // resident waves per SIMD (an autotuned choice)
      if (pip_waves == 4) fp32_pass(k_refine_pip<kPrecDeferred, 4>);
      else if (pip_waves == 6) fp32_pass(k_refine_pip<kPrecDeferred, 6>);
      else fp32_pass(k_refine_pip<kPrecDeferred, 8>);
      // the timer brackets the raster pass, the listing and the FP32 walk
      // (the launches the roofline figures describe)
      if (timer) timer->record_stop(s);
      timer = nullptr;
      hipLaunchKernelGGL(k_pip_redo_wave, 2048, kBlock, 0, s, c, allow, wb, wp, redo.p, aux.p, probe_ids, out.b.p,
                         out.p.p, out.cap(), out.counter.p);
      // (read before open_count is released on the stream)
      if (n_open) SJ_CHECK(hipMemcpyAsync(&walked, n_open, 8, hipMemcpyDeviceToHost, s));
    } else if (precision == 0) {
      hipLaunchKernelGGL(k_refine_pip<0>, grid_for(n), kBlock, 0, s, c, allow, cb, cp, n, nullptr, out.b.p,
                         out.p.p, out.cap(), out.counter.p, aux.p + 1, probe_ids, nullptr, nullptr, nullptr);
    } else {
      hipLaunchKernelGGL(k_refine_pip<1>, grid_for(n), kBlock, 0, s, c, allow, cb, cp, n, nullptr, out.b.p,
                         out.p.p, out.cap(), out.counter.p, aux.p + 1, probe_ids, nullptr, nullptr, nullptr);
    }
    SJ_CHECK(hipGetLastError());
    if (timer) timer->record_stop(s);
    unsigned long long h[4], cnt = 0;
    SJ_CHECK(hipMemcpyAsync(h, aux.p, 32, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipMemcpyAsync(&cnt, out.counter.p, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    if (stats) {
      stats->segment_tests += h[1];
      stats->fallback_tests += h[2];
      stats->walked += walked;
    }
    return cnt;
  }
  // polygon x polygon: the probe's rings get segment trees (and rasters) too
  const uint32_t polys = (1u << kPolygon) | (1u << kMultiPolygon);
  RingIndex probe_ri;
  if (predicate == kIntersects && (build.type_mask & polys) && (probe.type_mask & polys) && !probe.points_only &&
      c.R.ring_rect != nullptr) {
    RingIndexParams prm;
    build_ring_index(probe, probe_ri, prm, s);
    c.PR = probe_ri.view();
  }
  if (c.PR.ring_rect != nullptr && (build.type_mask & ~polys) == 0 && (probe.type_mask & ~polys) == 0) {
    if (timer) timer->record_start(s);
    // candidates decided by the rings' rects and cell rasters first; the
    // wavefront-per-candidate kernel takes the rest, copied in order
    PairBuf open_own;
    DevArray<uint64_t> mask_own;
    PairBuf& open = ws ? ws->open : open_own;
    DevArray<uint64_t>& mask = ws ? ws->mask : mask_own;
    DevArray<unsigned long long> open_count;
    const unsigned long long* n_open = nullptr;
    if ((ri.n_rasters > 0 || probe_ri.n_rasters > 0) && n <= 0xffffffffull)
      list_marked(
          n,
          [&](unsigned int g, uint64_t* marks, unsigned long long* tile_cnt) {
            hipLaunchKernelGGL(k_pp_raster_pass, g, kBlock, 0, s, c, cb, cp, n, out.b.p, out.p.p, out.cap(),
                               out.counter.p, marks, tile_cnt);
          },
          cb, cp, open, mask, open_count, &n_open, s);
    const uint32_t* wb = n_open ? open.b.p : cb;
    const uint32_t* wp = n_open ? open.p.p : cp;
    const uint64_t waves = std::min<uint64_t>(n, 1u << 18);
    const unsigned int grid = (unsigned int)((waves * 64 + kBlock - 1) / kBlock);
    if (precision == 0 && n <= 0xffffffffull) {
      // the certified phases (covered vertices, then two crossings, each
      // kernel lean), then the inline-precision kernel over the pairs they
      // leave (rows of several parts or rings, undecided)
      DevArray<uint32_t> lb_own, lp_own, db(n, s), dp(n, s);
      DevArray<uint32_t>& lb = ws ? ws->redo : lb_own;
      DevArray<uint32_t>& lp = ws ? ws->rest : lp_own;
      fit(lb, n, s);
      fit(lp, n, s);
      DevArray<unsigned long long> cnt(3, s);  // certain list, unsure list, deferred
      SJ_CHECK(hipMemsetAsync(cnt.p, 0, 24, s));
      hipLaunchKernelGGL(k_pp_vertices, grid, kBlock, 0, s, c, wb, wp, n, n_open, out.b.p, out.p.p, out.cap(),
                         out.counter.p, lb.p, lp.p, db.p, dp.p, cnt.p);
      hipLaunchKernelGGL(k_pp_crossings<false>, grid, kBlock, 0, s, c, lb.p, lp.p, n, cnt.p, out.b.p, out.p.p,
                         out.cap(), out.counter.p, db.p, dp.p, cnt.p + 2, n);
      hipLaunchKernelGGL(k_pp_crossings<true>, grid, kBlock, 0, s, c, lb.p, lp.p, n, cnt.p + 1, out.b.p, out.p.p,
                         out.cap(), out.counter.p, db.p, dp.p, cnt.p + 2, n);
      hipLaunchKernelGGL(k_refine_polypoly, std::min(grid, 4096u), kBlock, 0, s, c, db.p, dp.p, n, out.b.p, out.p.p,
                         out.cap(), out.counter.p, cnt.p + 2);
    } else {
      hipLaunchKernelGGL(k_refine_polypoly, grid, kBlock, 0, s, c, wb, wp, n, out.b.p, out.p.p, out.cap(),
                         out.counter.p, n_open);
    }
    SJ_CHECK(hipGetLastError());
    if (timer) timer->record_stop(s);
    unsigned long long cnt = 0, walked = n;
    SJ_CHECK(hipMemcpyAsync(&cnt, out.counter.p, 8, hipMemcpyDeviceToHost, s));
    if (n_open) SJ_CHECK(hipMemcpyAsync(&walked, n_open, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    if (stats) stats->walked += walked;
    return cnt;
  }
  const uint32_t lines = (1u << kLine) | (1u << kMultiLine);
  const bool general = ((build.type_mask | probe.type_mask) & lines) ||
                       (predicate != kIntersects && (build.type_mask & polys) && (probe.type_mask & polys));
  bool use_long = !general && !build.points_only && probe.points_only && c.R.ring_rect != nullptr &&
                  (predicate == kIntersects || predicate == kCovers || predicate == kContains);
  if (use_long) long_list.alloc(n, s);
  if (timer) timer->record_start(s);
  if (general)
    hipLaunchKernelGGL(k_refine_relate, grid_for(n), kBlock, 0, s, c, predicate, precision, cb, cp, n, out.b.p,
                       out.p.p, out.cap(), out.counter.p, aux.p + 1, (int*)(aux.p + 3));
  else
    hipLaunchKernelGGL(k_refine_lane, grid_for(n), kBlock, 0, s, c, predicate, precision, cb, cp, n, out.b.p,
                       out.p.p, out.cap(), out.counter.p, use_long ? long_list.p : nullptr, aux.p, aux.p + 1,
                       (int*)(aux.p + 3));
  SJ_CHECK(hipGetLastError());
  if (use_long) {
    // enough waves to fill the chip several times over
    hipLaunchKernelGGL(k_refine_wave, 256 * 16, kBlock, 0, s, c, predicate, precision, cb, cp, long_list.p, aux.p,
                       out.b.p, out.p.p, out.cap(), out.counter.p, aux.p + 1);
    SJ_CHECK(hipGetLastError());
  }
  if (timer) timer->record_stop(s);
  unsigned long long h[4], cnt = 0;
  SJ_CHECK(hipMemcpyAsync(h, aux.p, 32, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipMemcpyAsync(&cnt, out.counter.p, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  if (h[3]) throw Error("unsupported geometry type combination in refine");
  if (stats) {
    stats->segment_tests += h[1];
    stats->fallback_tests += h[2];
    stats->walked += n;
  }
  return cnt;
}

void pair_distances(const GeomColumn& build, const GeomColumn& probe, const uint32_t* cb, const uint32_t* cp,
                    uint64_t n, double* out, hipStream_t s) {
  if (n == 0) return;
  if (!probe.points_only && (probe.type_mask & ~(1u << kPoint)))
    throw Error("pair distances: probe rows must be points");
  hipLaunchKernelGGL(k_pair_distance, grid_for(n), kBlock, 0, s, build.view(), probe.view(), cb, cp, n, out);
  SJ_CHECK(hipGetLastError());
}

uint64_t refine_distance(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                         const uint32_t* cb, const uint32_t* cp, uint64_t n, PairBuf& out, hipStream_t s,
                         KernelTimer* timer, int precision, RefineStats* stats, Workspace* ws, bool raster_done,
                         const uint32_t* probe_ids, bool out_holds_kept) {
  out.counter.reserve(1, s);
  fit(out.b, n > 0 ? n : 1, s);
  fit(out.p, n > 0 ? n : 1, s);
  if (n == 0) {
    if (!out_holds_kept) return 0;
    unsigned long long cnt = 0;
    SJ_CHECK(hipMemcpyAsync(&cnt, out.counter.p, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    return cnt;
  }
  if (!probe.points_only && (probe.type_mask & ~(1u << kPoint)))
    throw Error("distance refine: probe rows must be points");
  if (!out_holds_kept) SJ_CHECK(hipMemsetAsync(out.counter.p, 0, 8, s));
  if (timer) timer->record_start(s);
  Ctx c;
  c.B = build.view();
  c.R = ri.view();
  c.P = probe.view();
  DevArray<unsigned long long> aux(3, s);  // segment tests, exact ring scans, FP64 segment terms
  SJ_CHECK(hipMemsetAsync(aux.p, 0, 24, s));
  const int plain = precision == 1 ? 1 : 0;
  const int kind = plain ? kDistAny : dist_kind(build.type_mask, build.points_only, ri);
  // lines: candidates whose point's square [p - d, p + d] meets no raster
  // cell of the line are dropped first; the walk takes the rest from a list
  PairBuf open_own;
  DevArray<uint64_t> mask_own;
  PairBuf& open = ws ? ws->open : open_own;
  DevArray<uint64_t>& mask = ws ? ws->mask : mask_own;
  DevArray<unsigned long long> open_count;
  const unsigned long long* n_open = nullptr;
  if (!raster_done && kind == kDistLines && ri.n_line_rasters > 0 && probe.points_only && d >= 0 && d < INFINITY &&
      n <= 0xffffffffull)
    list_marked(
        n,
        [&](unsigned int g, uint64_t* marks, unsigned long long* tile_cnt) {
          hipLaunchKernelGGL(k_dist_raster_pass, g, kBlock, 0, s, c, d, cb, cp, n, marks, tile_cnt);
        },
        cb, cp, open, mask, open_count, &n_open, s);
  const uint32_t* wb = n_open ? open.b.p : cb;
  const uint32_t* wp = n_open ? open.p.p : cp;
  // lines, certified form: the lean FP32 walk, then the FP64 kernel over
  // the candidates it lists (other rows, uncertain tests)
  const bool lean = (kind == kDistLines || kind == kDistPolygons) && precision == 0 && ri.desc.p != nullptr &&
                    probe.points_only && d >= 0 && d < INFINITY && n < 0xffffffffull - 64;
  DevArray<uint32_t> redo_own;
  DevArray<unsigned long long> walk_aux;  // redo count, queue, tests, uncertain tests
  switch (lean ? (kind == kDistLines ? -1 : -2) : kind) {
    case -2: {
      // single-ring polygons: the segments first (certainly within d:
      // kept); where every segment is certainly farther, the point-in-
      // polygon walk keeps the points certainly inside (with the cell
      // rasters) and lists the uncertain ones for its FP64 pass (inside or
      // on the boundary: distance 0); rows of several rings or parts and
      // uncertain segment terms go to the FP64 distance kernel
      DevArray<uint32_t>& redo = ws ? ws->redo : redo_own;
      fit(redo, n, s);
      DevArray<uint32_t> pb(n, s), pp(n, s), redo2(n, s);
      // [0] redo count, [1] queue, [2..3] walk stats, [4] pairs to the
      // inside test, [5] its redo count, [6] its queue, [7..8] its stats
      walk_aux.alloc(9, s);
      SJ_CHECK(hipMemsetAsync(walk_aux.p, 0, 72, s));
      launch_line_walk(ri, probe, wb, wp, n, n_open, probe_ids, d, out.b.p, out.p.p, out.cap(), out.counter.p,
                       walk_aux.p + 2, redo.p, walk_aux.p, s, walk_aux.p + 1, pb.p, pp.p, walk_aux.p + 4);
      launch_ring_walk(ri, probe, pb.p, pp.p, n, walk_aux.p + 4, probe_ids, false, out.b.p, out.p.p, out.cap(),
                       out.counter.p, walk_aux.p + 7, redo2.p, walk_aux.p + 5, nullptr, nullptr, 8, s, walk_aux.p + 6);
      hipLaunchKernelGGL(k_pip_redo_wave, 2048, kBlock, 0, s, c, 1, pb.p, pp.p, redo2.p, walk_aux.p + 5, probe_ids,
                         out.b.p, out.p.p, out.cap(), out.counter.p);
      hipLaunchKernelGGL(k_refine_distance<kDistPolygons>, std::min(grid_for(n), 2048u), kBlock, 0, s, c, d, 0, wb,
                         wp, n, redo.p, out.b.p, out.p.p, out.cap(), out.counter.p, aux.p, walk_aux.p, probe_ids);
      break;
    }
    case -1: {
      DevArray<uint32_t>& redo = ws ? ws->redo : redo_own;
      fit(redo, n, s);
      walk_aux.alloc(4, s);
      SJ_CHECK(hipMemsetAsync(walk_aux.p, 0, 32, s));
      launch_line_walk(ri, probe, wb, wp, n, n_open, probe_ids, d, out.b.p, out.p.p, out.cap(), out.counter.p,
                       walk_aux.p + 2, redo.p, walk_aux.p, s, walk_aux.p + 1);
      hipLaunchKernelGGL(k_refine_distance<kDistLines>, std::min(grid_for(n), 2048u), kBlock, 0, s, c, d, 0, wb, wp,
                         n, redo.p, out.b.p, out.p.p, out.cap(), out.counter.p, aux.p, walk_aux.p, probe_ids);
      break;
    }
    case kDistLines:
      hipLaunchKernelGGL(k_refine_distance<kDistLines>, grid_for(n), kBlock, 0, s, c, d, 0, wb, wp, n, nullptr,
                         out.b.p, out.p.p, out.cap(), out.counter.p, aux.p, n_open, probe_ids);
      break;
    case kDistPolygons:
      hipLaunchKernelGGL(k_refine_distance<kDistPolygons>, grid_for(n), kBlock, 0, s, c, d, 0, wb, wp, n, nullptr,
                         out.b.p, out.p.p, out.cap(), out.counter.p, aux.p, n_open, probe_ids);
      break;
    default:
      hipLaunchKernelGGL(k_refine_distance<kDistAny>, grid_for(n), kBlock, 0, s, c, d, plain, cb, cp, n,
                         nullptr, out.b.p, out.p.p, out.cap(), out.counter.p, aux.p, nullptr, probe_ids);
  }
  SJ_CHECK(hipGetLastError());
  if (timer) timer->record_stop(s);
  unsigned long long cnt = 0, h[3], walked = n, wh[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  SJ_CHECK(hipMemcpyAsync(&cnt, out.counter.p, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipMemcpyAsync(h, aux.p, 24, hipMemcpyDeviceToHost, s));
  if (walk_aux.p) SJ_CHECK(hipMemcpyAsync(wh, walk_aux.p, 8 * walk_aux.n, hipMemcpyDeviceToHost, s));
  if (n_open) SJ_CHECK(hipMemcpyAsync(&walked, n_open, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  if (stats) {
    stats->segment_tests += h[0] + wh[2] + wh[7];
    stats->fallback_tests += h[1] + h[2];  // FP64 re-evaluations: exact ring scans and segment terms
    stats->walked += walked;
  }
  return cnt;
}

void launch_pip_refine_dev(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, int predicate,
                           const uint32_t* cb, const uint32_t* cp, const unsigned long long* n_dev, uint64_t n_cap,
                           const uint32_t* probe_ids, uint32_t* redo, unsigned long long* redo_n, uint32_t* ob,
                           uint32_t* op, uint64_t out_cap, unsigned long long* out_n, unsigned long long* stats,
                           int pip_waves, hipStream_t s) {
  if (!pip_refine_applies(build, ri, probe, predicate)) throw Error("point-in-polygon refine does not apply");
  if (n_cap == 0) return;
  Ctx c;
  c.B = build.view();
  c.R = ri.view();
  c.P = probe.view();
  const int allow = predicate == kContains ? 0 : 1;
  // the lean walk; rows that are not single rings go to the FP64 pass
  launch_ring_walk(ri, probe, cb, cp, n_cap, n_dev, probe_ids, false, ob, op, out_cap, out_n, stats, redo, redo_n,
                   nullptr, nullptr, pip_waves, s);
  hipLaunchKernelGGL(k_refine_pip<1>, 1024, kBlock, 0, s, c, allow, cb, cp, n_cap, redo, ob, op, out_cap, out_n,
                     nullptr, probe_ids, nullptr, nullptr, redo_n);
  SJ_CHECK(hipGetLastError());
}

void launch_distance_refine_dev(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                                const uint32_t* cb, const uint32_t* cp, const unsigned long long* n_dev,
                                uint64_t n_cap, uint32_t* ob, uint32_t* op, uint64_t out_cap,
                                unsigned long long* out_n, unsigned long long* stats, hipStream_t s) {
  if (n_cap == 0) return;
  Ctx c;
  c.B = build.view();
  c.R = ri.view();
  c.P = probe.view();
  switch (dist_kind(build.type_mask, build.points_only, ri)) {
    case kDistLines:
      hipLaunchKernelGGL(k_refine_distance<kDistLines>, grid_for(n_cap), kBlock, 0, s, c, d, 0, cb, cp, n_cap, nullptr, ob,
                         op, out_cap, out_n, stats, n_dev, nullptr);
      break;
    case kDistPolygons:
      hipLaunchKernelGGL(k_refine_distance<kDistPolygons>, grid_for(n_cap), kBlock, 0, s, c, d, 0, cb, cp, n_cap, nullptr,
                         ob, op, out_cap, out_n, stats, n_dev, nullptr);
      break;
    default:
      hipLaunchKernelGGL(k_refine_distance<kDistAny>, grid_for(n_cap), kBlock, 0, s, c, d, 0, cb, cp, n_cap, nullptr, ob,
                         op, out_cap, out_n, stats, n_dev, nullptr);
  }
  SJ_CHECK(hipGetLastError());
}

uint64_t filter_refine_distance(const RectTree& t, const float4* q, int64_t nq, const uint32_t* qorder,
                                const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                                PairBuf& out, uint64_t* candidates, RefineStats* stats, hipStream_t s,
                                KernelTimer* timer) {
  if (!probe.points_only) throw Error("fused distance join: probe rows must be points");
  out.counter.reserve(1, s);
  // capacity: the tree's observed kept pairs per query (16ths), else 4
  uint64_t per16 = t.kept_per_query_x16.load(std::memory_order_relaxed);
  size_t want = std::max<int64_t>(1 << 16, per16 ? (int64_t)((per16 * (uint64_t)nq) / 16 * 9 / 8 + 1024) : nq * 4);
  out.b.reserve(want, s);
  out.p.reserve(want, s);
  if (candidates) *candidates = 0;
  if (nq == 0 || t.n_levels == 0) return 0;
  Ctx c;
  c.B = build.view();
  c.R = ri.view();
  c.P = probe.view();
  DevArray<unsigned long long> aux(4, s);
  TreeView v = tree_view(t);
  const int kind = dist_kind(build.type_mask, build.points_only, ri);
  auto launch = [&](auto kw) {
    kw.c = c;
    kw.d = d;
    kw.hits_seen = kw.tests = kw.scans = kw.fb = 0;
    kw.stats = aux.p;
    if (t.leaf == 4)
      hipLaunchKernelGGL((k_filter<RectTree::FANOUT, 4, decltype(kw)>), blocks_for(nq), kFilterBlock, 0, s, v, q, nq,
                         qorder, 0, kw, out.b.p, out.p.p, out.cap(), out.counter.p);
    else if (t.leaf == 16)
      hipLaunchKernelGGL((k_filter<RectTree::FANOUT, 16, decltype(kw)>), blocks_for(nq), kFilterBlock, 0, s, v, q, nq,
                         qorder, 0, kw, out.b.p, out.p.p, out.cap(), out.counter.p);
    else
      hipLaunchKernelGGL((k_filter<RectTree::FANOUT, 8, decltype(kw)>), blocks_for(nq), kFilterBlock, 0, s, v, q, nq,
                         qorder, 0, kw, out.b.p, out.p.p, out.cap(), out.counter.p);
  };
  for (int attempt = 0; attempt < 2; attempt++) {
    SJ_CHECK(hipMemsetAsync(out.counter.p, 0, 8, s));
    SJ_CHECK(hipMemsetAsync(aux.p, 0, 32, s));
    if (timer) timer->record_start(s);
    if (kind == kDistLines) launch(KeepWithin<kDistLines>{});
    else if (kind == kDistPolygons) launch(KeepWithin<kDistPolygons>{});
    else launch(KeepWithin<kDistAny>{});
    SJ_CHECK(hipGetLastError());
    if (timer) timer->record_stop(s);
    unsigned long long total = 0, h[4];
    SJ_CHECK(hipMemcpyAsync(&total, out.counter.p, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipMemcpyAsync(h, aux.p, 32, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    t.kept_per_query_x16.store((uint64_t)((total * 16) / (uint64_t)nq + 1), std::memory_order_relaxed);
    if (total <= out.cap()) {
      if (candidates) *candidates = h[0];
      if (stats) {
        stats->segment_tests += h[1];
        stats->fallback_tests += h[2] + h[3];
      }
      return total;
    }
    size_t cap = (size_t)(total + total / 8 + 1024);
    out.b.alloc(cap, s);
    out.p.alloc(cap, s);
  }
  throw Error("distance join output did not fit after resizing");
}

namespace {
__global__ void k_pack(const uint32_t* b, const uint32_t* p, uint64_t n, uint64_t* k) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) k[i] = ((uint64_t)p[i] << 32) | b[i];
}
__global__ void k_unpack(const uint64_t* k, uint64_t n, uint32_t* b, uint32_t* p) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    uint64_t v = k[i];
    b[i] = (uint32_t)v;
    p[i] = (uint32_t)(v >> 32);
  }
}
}  // namespace

namespace {
// bad[0] / bad[1]: the first position whose build / probe index is out of
// range (initially ~0)
__global__ void k_check_ranges(const uint32_t* cb, const uint32_t* cp, uint64_t n, uint64_t nb, uint64_t np,
                               unsigned long long* bad) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (cb[i] >= nb) atomicMin(&bad[0], (unsigned long long)i);
    if (cp[i] >= np) atomicMin(&bad[1], (unsigned long long)i);
  }
}
}  // namespace

void check_pair_ranges(const uint32_t* cb, const uint32_t* cp, uint64_t n, uint64_t nb, uint64_t np,
                       hipStream_t s) {
  if (n == 0) return;
  DevArray<unsigned long long> bad(2, s);
  SJ_CHECK(hipMemsetAsync(bad.p, 0xff, 16, s));
  hipLaunchKernelGGL(k_check_ranges, std::min(grid_for(n), 4096u), kBlock, 0, s, cb, cp, n, nb, np, bad.p);
  SJ_CHECK(hipGetLastError());
  unsigned long long h[2];
  SJ_CHECK(hipMemcpyAsync(h, bad.p, 16, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  if (h[0] == ~0ull && h[1] == ~0ull) return;
  // the first bad position, its build index checked before its probe index
  const bool build_bad = h[0] <= h[1];
  uint32_t v = 0;
  SJ_CHECK(hipMemcpyAsync(&v, (build_bad ? cb : cp) + (build_bad ? h[0] : h[1]), 4, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  throw Error(std::string(build_bad ? "build" : "probe") + " index out of range: " + std::to_string(v));
}

void sort_pairs_by_build(PairBuf& pb, uint64_t n, int64_t n_build, hipStream_t s) {
  if (n < 2) return;
  int bbits = 1;
  while (bbits < 32 && (1ll << bbits) < n_build) bbits++;
  DevArray<uint32_t> kb(n, s), vp(n, s);
  size_t tmp = 0;
  SJ_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, pb.b.p, kb.p, pb.p.p, vp.p, (int)n, 0, bbits, s));
  DevArray<uint8_t> t(tmp + 16, s);
  SJ_CHECK(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, pb.b.p, kb.p, pb.p.p, vp.p, (int)n, 0, bbits, s));
  pb.b = std::move(kb);
  pb.p = std::move(vp);
}

void sort_pairs(uint32_t* b, uint32_t* p, uint64_t n, hipStream_t s) {
  if (n < 2) return;
  DevArray<uint64_t> k1(n, s), k2(n, s);
  hipLaunchKernelGGL(k_pack, blocks_for(n), kBlock, 0, s, b, p, n, k1.p);
  size_t tmp = 0;
  SJ_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, k1.p, k2.p, (int)n, 0, 64, s));
  DevArray<uint8_t> t(tmp + 16, s);
  SJ_CHECK(hipcub::DeviceRadixSort::SortKeys(t.p, tmp, k1.p, k2.p, (int)n, 0, 64, s));
  hipLaunchKernelGGL(k_unpack, blocks_for(n), kBlock, 0, s, k2.p, n, b, p);
  SJ_CHECK(hipGetLastError());
}

}  // namespace sjhip
