// ring_walk.h -- the point-in-ring walk shared by the refine kernels:
// FP64 and certified-FP32 segment scans, and the depth-first walk over a
// ring's segment tree (fan-out kSegFanout, leaves of 1 << ls segments).
//
// The classification is tg_ring_contains_point (c/sedona-tg/src/tg/
// tg.c:2809-2822) over raycast (tg.c:1199-1275): any ON segment decides
// allow_on_edge, otherwise the parity of IN segments decides.
#pragma once

#include "geom_math.h"
#include "row_geom.h"
#include "sjhip.h"

namespace sjhip {

namespace {

__device__ __forceinline__ bool rect_covers(const double4& b, double px, double py) {
  return !(px < b.x || py < b.y || px > b.z || py > b.w);
}

// Accumulates the ray classification of segment (a, b) into (on, in).
__device__ __forceinline__ void acc_ray(int rc, int& on, int& in) {
  on |= rc == kRayOn;
  in ^= rc == kRayIn;
}

// FP64 re-evaluation of one segment, kept out of line so that the FP32 scan
// loops do not carry the FP64 path's registers.
__device__ __noinline__ int raycast_fallback(const double2* xy, int j, double px, double py) {
  double2 A = xy[j], B = xy[j + 1];
  return raycast(A.x, A.y, B.x, B.y, px, py);
}

// FP64 plain ring test over segments [s0, s1) with stride.
__device__ __forceinline__ void ring_segments_f64(const RingRef& r, int64_t s0, int64_t s1, int64_t stride,
                                                  double px, double py, int& on, int& in, uint64_t& tests) {
  for (int64_t j = s0; j < s1; j += stride) {
    double2 a = r.xy[j], b = r.xy[j + 1];
    if (py < fmin0(a.y, b.y) || py > fmax0(a.y, b.y)) continue;
    tests++;
    acc_ray(raycast(a.x, a.y, b.x, b.y, px, py), on, in);
  }
}

// FP32-certified ring test over segments [s0, s1) with stride.  With
// Defer, an uncertain segment is only counted (the caller re-refines the
// candidate in FP64); otherwise it is decided by the FP64 test at once.
template <bool Defer = false>
__device__ __forceinline__ void ring_segments_f32(const RingRef& r, const float2* xy32, float E, float pxl, float pyl,
                                                  double px, double py, int64_t s0, int64_t s1, int64_t stride,
                                                  int& on, int& in, uint64_t& tests, uint64_t& fallbacks) {
  const float2* q = xy32 + r.v0;
  for (int64_t j = s0; j < s1; j += stride) {
    float2 a = q[j], b = q[j + 1];
    int rc = raycast_f32_certified(a.x, a.y, b.x, b.y, pxl, pyl, E);
    if (rc == kRayOut) continue;
    tests++;
    if (rc < 0) {
      fallbacks++;
      if constexpr (Defer) rc = kRayOut;
      else rc = raycast_fallback(r.xy, (int)j, px, py);
    }
    acc_ray(rc, on, in);
  }
}

// ---- segment-tree walk ----------------------------------------------------
//
// A segment can classify the +x ray as IN or ON only when the point's
// height lies within the segment's y-range and the segment is not entirely
// left of the point (tg.c:1203-1213: a segment whose max x is below p.x is
// OUT, degenerate segments included).  The reference's own natural index
// prunes the same way (branch_maybe_in, tg.c:2762-2772).  Boxes are FP32
// ring-local and the point's local coordinates are within E of exact, so a
// margin of 2E keeps the pruning conservative.

// Level sizes of a ring's tree (levels <= kSegMaxLevels), as scalars so that
// nothing is indexed dynamically (dynamic indexing would place them in
// scratch memory).
struct TreeShape {
  int c0, c1, c2, c3, c4;  // node counts per level
  int levels;
  template <int L>
  __host__ __device__ __forceinline__ int cnt() const {
    return L == 0 ? c0 : L == 1 ? c1 : L == 2 ? c2 : L == 3 ? c3 : c4;
  }
  template <int L>
  __host__ __device__ __forceinline__ int off() const {
    return L == 0 ? 0 : L == 1 ? c0 : L == 2 ? c0 + c1 : L == 3 ? c0 + c1 + c2 : c0 + c1 + c2 + c3;
  }
  __host__ __device__ __forceinline__ int total() const { return c0 + c1 + c2 + c3 + c4; }
};

__host__ __device__ __forceinline__ int up_level(int c) { return (c + kSegFanout - 1) / kSegFanout; }

__host__ __device__ __forceinline__ TreeShape seg_tree_shape(int ns, int ls) {
  TreeShape t;
  t.c0 = (ns + (1 << ls) - 1) >> ls;
  t.c1 = t.c2 = t.c3 = t.c4 = 0;
  t.levels = 1;
  if (t.c0 > kSegFanout) {
    t.c1 = up_level(t.c0);
    t.levels = 2;
    if (t.c1 > kSegFanout) {
      t.c2 = up_level(t.c1);
      t.levels = 3;
      if (t.c2 > kSegFanout) {
        t.c3 = up_level(t.c2);
        t.levels = 4;
        if (t.c3 > kSegFanout) {
          t.c4 = up_level(t.c3);
          t.levels = 5;
        }
      }
    }
  }
  return t;
}

struct TreeWalk {
  const float4* box;  // ring's boxes, level 0 first
  int levels;         // tree levels (level sizes follow from ns; see level_geom)
  float pxl, pyl, m;
  const double2* xy;  // ring's FP64 points
  const float2* q32;  // ring's FP32 local points
  float E;
  double px, py;
  int ns;
  int ls;    // segments per leaf: 1 << ls
  // ray direction: along the x axis (swap = 0) or the y axis (swap = 1),
  // toward + (sx = 1) or - (sx = -1); the tests read (along, across)
  // coordinates (sx * x, y) or (sx * y, x)
  float sx;
  int swap;
  int on, in;
  uint32_t tests, fallbacks;
};

__device__ __forceinline__ bool box_pass(const TreeWalk& w, const float4& b) {
  const float c0 = w.swap ? b.x : b.y, c1 = w.swap ? b.z : b.w;  // across the ray
  const float a0 = w.swap ? b.y : b.x, a1 = w.swap ? b.w : b.z;  // along the ray
  const float pa = w.swap ? w.pyl : w.pxl, pc = w.swap ? w.pxl : w.pyl;
  return pc >= c0 - w.m && pc <= c1 + w.m && (w.sx > 0.f ? pa <= a1 + w.m : pa >= a0 - w.m);
}

// Ring-walk precision modes (template parameter PREC): 0 = certified FP32,
// uncertain segments decided by the FP64 test inline; 1 = FP64 throughout;
// kPrecDeferred = certified FP32, uncertain segments only counted, so that
// the caller re-refines the candidate with mode 1.
constexpr int kPrecDeferred = 2;

#ifndef SJ_LEAF_BATCH
#define SJ_LEAF_BATCH 8
#endif
constexpr int kLeafBatch = SJ_LEAF_BATCH;

// A leaf's (up to) LEAF segments from s0, walking the vertex run once.
// LEAF is a compile-time size so that the batched loads below unroll
// (the FP64 mode reads the leaf size from the walk instead).
template <int PREC, int LEAF>
__device__ __forceinline__ void scan_leaf_n(TreeWalk& w, const int s0) {
  const int s1 = min(s0 + LEAF, w.ns);
  if constexpr (PREC != 1) {
    // the vertices of kLeafBatch segments are requested before the first
    // test, so that their latencies overlap (the walk is latency-bound)
#pragma unroll
    for (int cb = 0; cb < LEAF; cb += kLeafBatch) {
      const int c0 = s0 + cb;
      if (c0 >= s1) break;
      float2 v[kLeafBatch + 1];
#pragma unroll
      for (int k = 0; k <= kLeafBatch; k++) v[k] = c0 + k <= s1 ? w.q32[c0 + k] : make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < kLeafBatch; k++) {
        if (c0 + k >= s1) break;
        const float2 a = v[k], b = v[k + 1];
        int rc = w.swap ? raycast_f32_certified(w.sx * a.y, a.x, w.sx * b.y, b.x, w.sx * w.pyl, w.pxl, w.E)
                        : raycast_f32_certified(w.sx * a.x, a.y, w.sx * b.x, b.y, w.sx * w.pxl, w.pyl, w.E);
        if (rc == kRayOut) continue;
        w.tests++;
        if (rc < 0) {
          w.fallbacks++;
          if constexpr (PREC == kPrecDeferred) rc = kRayOut;
          else rc = raycast_fallback(w.xy, c0 + k, w.px, w.py);
        }
        w.on |= rc == kRayOn;
        w.in ^= rc == kRayIn;
      }
    }
  } else {
    const int e = min(s0 + (1 << w.ls), w.ns);
    double2 a = w.xy[s0];
    for (int j = s0; j < e; j++) {
      double2 b = w.xy[j + 1];
      double2 aa = a;
      a = b;
      if (w.py < fmin0(aa.y, b.y) || w.py > fmax0(aa.y, b.y)) continue;
      w.tests++;
      int rc = raycast(aa.x, aa.y, b.x, b.y, w.px, w.py);
      w.on |= rc == kRayOn;
      w.in ^= rc == kRayIn;
    }
  }
}

// The index's leaf size is one of two (kSegLeafShift or kSegLeafShift + 1)
// and the same for every lane of a launch, so the branch is uniform.
// The FP64 mode (a plain loop) keeps the runtime size.
template <int PREC>
__device__ __forceinline__ void scan_leaf(TreeWalk& w, int leaf) {
  if constexpr (PREC == 1) scan_leaf_n<1, 0>(w, leaf << w.ls);
  else if (w.ls == kSegLeafShift + 1) scan_leaf_n<PREC, 2 << kSegLeafShift>(w, leaf << (kSegLeafShift + 1));
  else scan_leaf_n<PREC, 1 << kSegLeafShift>(w, leaf << kSegLeafShift);
}

#ifndef SJ_BOX_BATCH
#define SJ_BOX_BATCH 8
#endif
constexpr int kBoxBatch = SJ_BOX_BATCH;

// Tests up to kSegFanout boxes b[0..n), kBoxBatch loads in flight at once.
__device__ __forceinline__ uint32_t children_mask(const TreeWalk& w, const float4* b, int n) {
  uint32_t mask = 0;
#pragma unroll
  for (int c0 = 0; c0 < kSegFanout; c0 += kBoxBatch) {
    float4 bx[kBoxBatch];
#pragma unroll
    for (int k = 0; k < kBoxBatch; k++) bx[k] = c0 + k < n ? b[c0 + k] : make_float4(1.f, 1.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < kBoxBatch; k++)
      if (c0 + k < n && box_pass(w, bx[k])) mask |= 1u << (c0 + k);
  }
  return mask;
}

__device__ __forceinline__ int level_cnt(const TreeShape& t, int l) {
  return l == 0 ? t.c0 : l == 1 ? t.c1 : l == 2 ? t.c2 : l == 3 ? t.c3 : t.c4;
}
__device__ __forceinline__ int level_off(const TreeShape& t, int l) {
  int o = 0;
  if (l > 0) o += t.c0;
  if (l > 1) o += t.c1;
  if (l > 2) o += t.c2;
  if (l > 3) o += t.c3;
  return o;
}

// Offset (in boxes) and node count of level l of the tree over ns
// segments, recomputed instead of held in registers.
__device__ __forceinline__ void level_geom(int ns, int ls, int l, int& off, int& cnt) {
  int c = (ns + (1 << ls) - 1) >> ls, o = 0;
  for (int k = 0; k < l; k++) {
    o += c;
    c = (c + kSegFanout - 1) / kSegFanout;
  }
  off = o;
  cnt = c;
}

// Depth-first walk without recursion.  The pending children of every level
// are one byte each of `masks` (bit k of byte l: node base_l + k of level l
// still to visit).  Children of node v are 8v .. 8v+7, so the base of the
// level above is recovered as (base >> 3) & ~7.
template <int PREC>
__device__ __forceinline__ void walk_tree(TreeWalk& w) {
  const int top = w.levels - 1;
  int otop, ntop;
  level_geom(w.ns, w.ls, top, otop, ntop);
  const float4* btop = w.box + otop;
  for (int c0 = 0; c0 < ntop; c0 += kSegFanout) {
    uint64_t masks = (uint64_t)children_mask(w, btop + c0, min(kSegFanout, ntop - c0)) << (8 * top);
    int lvl = top, base = c0;
    while (masks) {
      uint32_t m = (uint32_t)(masks >> (8 * lvl)) & 0xffu;
      if (m == 0) {
        lvl++;
        base = (base >> 3) & ~(kSegFanout - 1);
        continue;
      }
      int k = __ffs(m) - 1;
      masks &= ~(1ull << (8 * lvl + k));
      int node = base + k;
      if (lvl == 0) {
        scan_leaf<PREC>(w, node);
        continue;
      }
      int cb = node * kSegFanout;
      int loff, lcnt;
      level_geom(w.ns, w.ls, lvl - 1, loff, lcnt);
      uint32_t cm = children_mask(w, w.box + loff + cb, min(kSegFanout, lcnt - cb));
      if (cm) {
        lvl--;
        base = cb;
        masks |= (uint64_t)cm << (8 * lvl);
      }
    }
  }
}

// Depth-first walk over a ring's segment tree (the layout walk_tree
// uses) visiting the leaves whose path of boxes passes `pass`; stops and
// returns true as soon as `leaf(leaf_index)` does.
template <class Pass, class Leaf>
__device__ __forceinline__ bool walk_any(const float4* box, const TreeShape& shape, Pass pass, Leaf leaf) {
  const int top = shape.levels - 1;
  const int ntop = level_cnt(shape, top);
  const float4* btop = box + level_off(shape, top);
  for (int c0 = 0; c0 < ntop; c0 += kSegFanout) {
    uint32_t m0 = 0;
    const int n0 = min(kSegFanout, ntop - c0);
    for (int k = 0; k < n0; k++)
      if (pass(btop[c0 + k])) m0 |= 1u << k;
    uint64_t masks = (uint64_t)m0 << (8 * top);
    int lvl = top, base = c0;
    while (masks) {
      uint32_t m = (uint32_t)(masks >> (8 * lvl)) & 0xffu;
      if (m == 0) {
        lvl++;
        base = (base >> 3) & ~(kSegFanout - 1);
        continue;
      }
      int k = __ffs(m) - 1;
      masks &= ~(1ull << (8 * lvl + k));
      int node = base + k;
      if (lvl == 0) {
        if (leaf(node)) return true;
        continue;
      }
      int cb = node * kSegFanout;
      int n = min(kSegFanout, level_cnt(shape, lvl - 1) - cb);
      const float4* bc = box + level_off(shape, lvl - 1) + cb;
      uint32_t cm = 0;
      for (int q = 0; q < n; q++)
        if (pass(bc[q])) cm |= 1u << q;
      if (cm) {
        lvl--;
        base = cb;
        masks |= (uint64_t)cm << (8 * lvl);
      }
    }
  }
  return false;
}

// tg_ring_contains_point over a ring given its rect, vertex run (FP64 xy
// and FP32 ring-local q32, both starting at the ring's first vertex), FP32
// frame (origin, bound E) and segment tree (nullptr: none).
template <int PREC>
__device__ __forceinline__ bool ring_contains_run(const double2* xy, const float2* q32, const float4* tree, int ls,
                                                  int64_t npts, const double4& rb, double2 o, float E, double px,
                                                  double py, bool allow_on_edge, uint64_t& tests,
                                                  uint64_t& fallbacks) {
  if (!rect_covers(rb, px, py)) return false;
  RingRef r;
  r.v0 = 0;
  r.npts = npts;
  r.xy = xy;
  int64_t ns = r.npts >= 3 ? r.npts - 1 : 0;
  int on = 0, in = 0;
  float pxl = (float)(px - o.x), pyl = (float)(py - o.y);
  if (tree) {
    TreeWalk w;
    w.box = tree;
    w.levels = seg_tree_shape((int)ns, ls).levels;
    w.pxl = pxl;
    w.pyl = pyl;
    w.m = 2.0f * E;
    w.xy = xy;
    w.q32 = q32;
    w.E = E;
    w.px = px;
    w.py = py;
    w.ns = (int)ns;
    w.ls = ls;
    // The certified pass of the deferred mode casts its ray toward the
    // nearest side of the ring's rect, so that fewer segments and tree
    // nodes lie on it.  Mirroring and transposing the coordinates are
    // exact, every ring is closed (the decoder rejects open rings) and the
    // crossing parity of a closed ring is the same for every ray from a
    // point off its boundary; a point on the boundary, or level with a
    // vertex across the ray, leaves a test uncertain, and the candidate is
    // then re-refined along +x in FP64 (the reference's raycast).  The
    // inline mode (PREC 0) decides uncertain tests segment by segment along
    // +x, which cannot be mixed with another ray: it walks along the
    // nearest side certified first and keeps that answer only when no test
    // was uncertain, else walks the ring again along +x.  The FP64 mode
    // keeps +x throughout.
    w.sx = 1.0f;
    w.swap = 0;
    w.on = 0;
    w.in = 0;
    w.tests = 0;
    w.fallbacks = 0;
    if constexpr (PREC != 1) {
      TreeWalk v = w;
      const double dl = px - rb.x, dr = rb.z - px, db = py - rb.y, dt = rb.w - py;
      const double dx = fmin(dl, dr), dy = fmin(db, dt);
      if (dy < dx) {
        v.swap = 1;
        v.sx = db < dt ? -1.0f : 1.0f;
      } else {
        v.sx = dl < dr ? -1.0f : 1.0f;
      }
      if constexpr (PREC == kPrecDeferred) {
        w = v;
      } else {
        // inline mode: a certified walk along the nearest side decides when
        // none of its tests is uncertain; else the ring is walked again
        // along +x with uncertain segments decided in FP64
        walk_tree<kPrecDeferred>(v);
        tests += v.tests;
        if (v.fallbacks == 0) return v.in != 0;  // (no test is ON when none is uncertain)
      }
    }
    walk_tree<PREC>(w);
    tests += w.tests;
    fallbacks += w.fallbacks;
    return w.on ? allow_on_edge : (w.in != 0);
  }
  if constexpr (PREC != 1) {
    ring_segments_f32<PREC == kPrecDeferred>(r, q32, E, pxl, pyl, px, py, 0, ns, 1, on, in, tests, fallbacks);
  } else {
    ring_segments_f64(r, 0, ns, 1, px, py, on, in, tests);
  }
  return on ? allow_on_edge : (in != 0);
}

// Cells per unit length along x and y of a raster of `cells` cells per
// side over rect r (positive width and height).
__device__ __forceinline__ double2 cells_per_unit(const double4& r, double cells) {
  return make_double2(cells / (r.z - r.x), cells / (r.w - r.y));
}

// Cell-raster class of a point for a single-ring polygon (record d, its
// ring index R, inv = cells_per_unit(d.rect, kRasterCells)): 1 inside, 2
// outside (a clean cell, or outside the ring's rect), 0 undecided (a cell
// near the boundary, or no raster).  A clean cell's 8 neighbours hold no
// boundary either, so a point whose computed cell is one off its own (the
// rounding of the product) still gets its own class.
__device__ __forceinline__ int raster_class_inv(const RingIndexView& R, const PolyDesc& d, double2 pt, double2 inv) {
  if (d.raster_off < 0) return 0;
  if (!rect_covers(d.rect, pt.x, pt.y)) return 2;
  const int ci = min(max((int)floor((pt.x - d.rect.x) * inv.x), 0), kRasterCells - 1);
  const int cj = min(max((int)floor((pt.y - d.rect.y) * inv.y), 0), kRasterCells - 1);
  const int cell = cj * kRasterCells + ci;
  return (R.raster[d.raster_off + (cell >> 4)] >> (2 * (cell & 15))) & 3;
}

__device__ __forceinline__ int raster_class(const RingIndexView& R, const PolyDesc& d, double2 pt) {
  if (d.raster_off < 0) return 0;
  return raster_class_inv(R, d, pt, cells_per_unit(d.rect, kRasterCells));
}

// Whether a line's cell raster leaves a segment possibly within dist of the
// point: some cell of the square [pt - dist, pt + dist] (widened by one
// cell, which also covers a computed cell one off the exact one) is set.
// False also when the square misses the line's rect or the point has no
// coordinates; rows without a raster answer true.  inv =
// cells_per_unit(d.rect, kLineCells).
__device__ __forceinline__ bool line_raster_may_hold_d(const RingIndexView& R, const PolyDesc& d, double2 inv,
                                                       double2 pt, double dist) {
  if (!(pt.x == pt.x || pt.y == pt.y)) return false;
  if (d.raster_off < 0) return true;
  const double fx0 = floor((pt.x - dist - d.rect.x) * inv.x) - 1.0, fx1 = floor((pt.x + dist - d.rect.x) * inv.x) + 1.0;
  const double fy0 = floor((pt.y - dist - d.rect.y) * inv.y) - 1.0, fy1 = floor((pt.y + dist - d.rect.y) * inv.y) + 1.0;
  if (fx1 < 0.0 || fy1 < 0.0 || fx0 > kLineCells - 1.0 || fy0 > kLineCells - 1.0) return false;
  if (!(fy1 - fy0 < 24.0)) return true;  // a window of many rows: left to the walk
  const int i0 = (int)fmax(fx0, 0.0), i1 = (int)fmin(fx1, kLineCells - 1.0);
  const int j0 = (int)fmax(fy0, 0.0), j1 = (int)fmin(fy1, kLineCells - 1.0);
  if (i0 > i1 || j0 > j1) return true;
  // bits i0..i1 of a row, split over its words
  uint64_t m[kLineRowWords];
#pragma unroll
  for (int w = 0; w < kLineRowWords; w++) {
    const int lo = max(i0 - 64 * w, 0), hi = min(i1 - 64 * w, 63);
    m[w] = lo <= hi ? (~0ull >> (63 - (hi - lo))) << lo : 0ull;
  }
  for (int r = j0; r <= j1; r++) {
    const uint64_t* rw = R.line_raster + d.raster_off + (int64_t)r * kLineRowWords;
    bool any = false;
#pragma unroll
    for (int w = 0; w < kLineRowWords; w++) any |= (rw[w] & m[w]) != 0;
    if (any) return true;
  }
  return false;
}

__device__ __forceinline__ bool line_raster_may_hold(const RingIndexView& R, uint32_t b, double2 pt, double dist) {
  if (!(pt.x == pt.x || pt.y == pt.y)) return false;
  const PolyDesc d = R.desc[b];
  if (d.raster_off < 0) return true;
  return line_raster_may_hold_d(R, d, cells_per_unit(d.rect, kLineCells), pt, dist);
}

// The FP32 frame origin of a ring: the centre of its rect (the record of a
// single-ring row stores the rect, so the origin is recomputed from it).
__device__ __forceinline__ double2 ring_origin(const double4& rb) {
  double ox = 0.5 * rb.x + 0.5 * rb.z, oy = 0.5 * rb.y + 0.5 * rb.w;
  if (!(ox == ox)) { ox = 0; oy = 0; }
  return make_double2(ox, oy);
}

}  // namespace

}  // namespace sjhip
