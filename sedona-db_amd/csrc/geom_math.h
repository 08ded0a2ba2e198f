// geom_math.h -- scalar geometry predicates shared by host and device code.
//
// The relation predicates reproduce the reference CPU refine's IEEE double
// expression sequence (TG, c/sedona-tg/src/tg/tg.c) so that every decision
// is bit-identical, including those inside its rounding band.  Every
// translation unit including this header is compiled with
// -ffp-contract=off: the reference is an x86-64 build without FMA.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define SJ_HD __host__ __device__ __forceinline__

namespace sjhip {

enum RayResult : int { kRayOut = 0, kRayIn = 1, kRayOn = 2 };

SJ_HD uint64_t dbits(double x) { return __builtin_bit_cast(uint64_t, x); }
SJ_HD double bitsd(uint64_t b) { return __builtin_bit_cast(double, b); }

// Smallest double greater than x (x not NaN, not +inf): the ulp step that
// nexttoward(x, +INFINITY) takes in tg.c:885-890, 1256.
SJ_HD double next_up(double x) {
  if (x != x || x == INFINITY) return x;
  if (x == 0.0) return bitsd(1ull);  // +denorm_min for +0 and -0
  uint64_t b = dbits(x);
  return bitsd(x > 0 ? b + 1 : b - 1);
}
SJ_HD double next_down(double x) { return -next_up(-x); }

// tg.c:853-855: equality that is false for NaN comparisons like the
// reference's !((x < y) | (x > y)).
SJ_HD bool feq(double x, double y) { return !((x < y) | (x > y)); }
// tg.c:816-823: x86 minsd/maxsd select semantics.
SJ_HD double fmin0(double x, double y) { return x < y ? x : y; }
SJ_HD double fmax0(double x, double y) { return x > y ? x : y; }

// tg.c:861-901: collinearity by comparing the two cross-product terms
// after a one-ulp correction of each term towards the side its division
// check says was lost.
SJ_HD bool collinear(double x1, double y1, double x2, double y2, double x3, double y3) {
  bool x1x2 = feq(x1, x2), x1x3 = feq(x1, x3), x2x3 = feq(x2, x3);
  bool y1y2 = feq(y1, y2), y1y3 = feq(y1, y3), y2y3 = feq(y2, y3);
  if (x1x2) return x1x3;
  if (y1y2) return y1y3;
  if ((x1x2 & y1y2) | (x1x3 & y1y3) | (x2x3 & y2y3)) return true;
  double cx1 = x3 - x1, cy1 = y3 - y1;
  double cx2 = x2 - x1, cy2 = y2 - y1;
  double s1 = cx1 * cy2;
  double s2 = cy1 * cx2;
  double s3 = (s1 / cy2) - cx1;
  double s4 = (s2 / cx2) - cy1;
  if (s3 < 0) s1 = next_down(s1);
  else if (s3 > 0) s1 = next_up(s1);
  if (s4 < 0) s2 = next_down(s2);
  else if (s4 > 0) s2 = next_up(s2);
  return feq(s1 - s2, 0.0);
}

// tg.c:1199-1275: the reference's classification of the +x ray from p
// against segment (a, b).
SJ_HD int raycast(double ax, double ay, double bx, double by, double px, double py) {
  double rminx = fmin0(ax, bx), rminy = fmin0(ay, by);
  double rmaxx = fmax0(ax, bx), rmaxy = fmax0(ay, by);
  if (py < rminy || py > rmaxy) return kRayOut;
  if (px < rminx) {
    if (py != rminy && py != rmaxy) return kRayIn;
  } else if (px > rmaxx) {
    if (rminy != rmaxy && rminx != rmaxx) return kRayOut;
  }
  if (by < ay) {
    double t = ax; ax = bx; bx = t;
    t = ay; ay = by; by = t;
  }
  if ((feq(px, ax) && feq(py, ay)) || (feq(px, bx) && feq(py, by))) return kRayOn;
  if (ay == by) {
    if (ax == bx) return kRayOut;
    if (py == by && !(px < rminx || px > rmaxx)) return kRayOn;
  }
  if (ax == bx && px == bx && py >= ay && py <= by) return kRayOn;
  if (collinear(ax, ay, bx, by, px, py)) {
    if (px < rminx) {
      if (rminy == rmaxy) return kRayOut;
    } else if (px > rmaxx) {
      return kRayOut;
    }
    return kRayOn;
  }
  if (py == ay || py == by) py = next_up(py);
  if (py < ay || py > by) return kRayOut;
  if (ax > bx) {
    if (px >= ax) return kRayOut;
    if (px <= bx) return kRayIn;
  } else {
    if (px >= bx) return kRayOut;
    if (px <= ax) return kRayIn;
  }
  if ((py - ay) / (px - ax) >= (by - ay) / (bx - ax)) return kRayIn;
  return kRayOut;
}

// Certified FP32 classification of the same ray test in ring-local
// coordinates.  Inputs are FP32 values within E of the exact local
// coordinates (E covers the FP32 rounding of the vertices and of the point
// and the FP64 rounding the reference's own arithmetic can commit).
// Returns kRayOut / kRayIn when the reference's FP64 decision is certain,
// -1 when it is not (the caller then runs raycast() in FP64).
// Host and device compile this one definition (the kernels call it on
// mirrored and transposed coordinates for rays toward -x, +y and -y).
SJ_HD int raycast_f32_certified(float ax, float ay, float bx, float by, float px, float py, float E) {
  float lo = ay < by ? ay : by, hi = ay > by ? ay : by;
  float e2 = 2.0f * E;
  if (py < lo - e2 || py > hi + e2) return kRayOut;
  if (!(py > lo + e2 && py < hi - e2)) return -1;  // near a vertex height
  float xl = ax < bx ? ax : bx, xh = ax > bx ? ax : bx;
  if (px < xl - e2) return kRayIn;
  if (px > xh + e2) return kRayOut;
  // a below b
  if (by < ay) {
    float t = ax; ax = bx; bx = t;
    t = ay; ay = by; by = t;
  }
  float d1 = py - ay, d2 = bx - ax, d3 = by - ay, d4 = px - ax;
  float det = d1 * d2 - d3 * d4;
  float a1 = fabsf(d1), a2 = fabsf(d2), a3 = fabsf(d3), a4 = fabsf(d4);
  const float u = 5.9604645e-08f;  // 2^-24
  float bound = 3.0f * E * (a1 + a2 + a3 + a4) + 8.0f * u * (a1 * a2 + a3 * a4) + 8.0f * E * E;
  if (det > bound) return kRayIn;
  if (det < -bound) return kRayOut;
  return -1;
}

// ---------------------------------------------------------------------------
// Reference segment x segment test: tg.c:1392-1443 (with point_on_segment,
// tg.c:1186-1197).

SJ_HD bool segment_covers_point(double ax, double ay, double bx, double by, double px, double py) {
  if (px < fmin0(ax, bx) || py < fmin0(ay, by) || px > fmax0(ax, bx) || py > fmax0(ay, by)) return false;
  return collinear(ax, ay, bx, by, px, py);
}

SJ_HD bool segment_intersects(double ax, double ay, double bx, double by, double cx, double cy, double dx,
                              double dy) {
  if (fmin0(cx, dx) > fmax0(ax, bx) || fmax0(cx, dx) < fmin0(ax, bx) || fmin0(cy, dy) > fmax0(ay, by) ||
      fmax0(cy, dy) < fmin0(ay, by))
    return false;
  if ((feq(ax, cx) && feq(ay, cy)) || (feq(ax, dx) && feq(ay, dy)) || (feq(bx, cx) && feq(by, cy)) ||
      (feq(bx, dx) && feq(by, dy)))
    return true;
  double cmpx = cx - ax, cmpy = cy - ay;
  double rx = bx - ax, ry = by - ay;
  double cmpxr = cmpx * ry - cmpy * rx;
  if (feq(cmpxr, 0.0)) {
    if (!(((cx - ax <= 0) != (cx - bx <= 0))) || ((cy - ay <= 0) != (cy - by <= 0))) {
      return segment_covers_point(ax, ay, bx, by, cx, cy) || segment_covers_point(ax, ay, bx, by, dx, dy) ||
             segment_covers_point(cx, cy, dx, dy, ax, ay) || segment_covers_point(cx, cy, dx, dy, bx, by);
    }
    return true;
  }
  double sx = dx - cx, sy = dy - cy;
  double rxs = rx * sy - ry * sx;
  if (feq(rxs, 0.0)) return false;
  double cmpxs = cmpx * sy - cmpy * sx;
  double rxsr = 1 / rxs;
  double t = cmpxs * rxsr;
  double u = cmpxr * rxsr;
  return (t >= 0) && (t <= 1) && (u >= 0) && (u <= 1);
}

// Certified FP32 orientation of c relative to the directed line a -> b:
// 1 left, -1 right, 0 uncertain.  Inputs are FP32 values within e (per
// coordinate) of exact ones; the bound has the form raycast_f32_certified
// uses for the same 2 x 2 determinant.
SJ_HD int orient_f32_certified(float ax, float ay, float bx, float by, float cx, float cy, float e) {
  const float d1 = bx - ax, d2 = cy - ay, d3 = by - ay, d4 = cx - ax;
  const float det = d1 * d2 - d3 * d4;
  const float a1 = fabsf(d1), a2 = fabsf(d2), a3 = fabsf(d3), a4 = fabsf(d4);
  const float u = 5.9604645e-08f;  // 2^-24
  const float bound = 3.0f * e * (a1 + a2 + a3 + a4) + 8.0f * u * (a1 * a2 + a3 * a4) + 8.0f * e * e;
  return det > bound ? 1 : det < -bound ? -1 : 0;
}

// Certified FP32 form of segment_intersects (below) for segments a-b and
// c-d given within e of exact: 1 when the segments certainly cross at a
// point interior to both (each one's endpoints strictly on opposite sides
// of the other's line), 0 when one segment certainly lies strictly on one
// side of the other's line, -1 otherwise.  Where it decides, the exact
// closed segments intersect (or not) with a margin far above the rounding
// of the reference's FP64 expressions, so segment_intersects agrees.
SJ_HD int segments_cross_f32_certified(float ax, float ay, float bx, float by, float cx, float cy, float dx, float dy,
                                       float e) {
  const float e2 = 2.0f * e;
  if (fminf(cx, dx) > fmaxf(ax, bx) + e2 || fmaxf(cx, dx) < fminf(ax, bx) - e2 ||
      fminf(cy, dy) > fmaxf(ay, by) + e2 || fmaxf(cy, dy) < fminf(ay, by) - e2)
    return 0;
  const int o1 = orient_f32_certified(cx, cy, dx, dy, ax, ay, e), o2 = orient_f32_certified(cx, cy, dx, dy, bx, by, e);
  if (o1 != 0 && o1 == o2) return 0;
  const int o3 = orient_f32_certified(ax, ay, bx, by, cx, cy, e), o4 = orient_f32_certified(ax, ay, bx, by, dx, dy, e);
  if (o3 != 0 && o3 == o4) return 0;
  if (o1 != 0 && o2 != 0 && o3 != 0 && o4 != 0) return 1;
  return -1;
}

// ---------------------------------------------------------------------------
// Exact orientation sign (Shewchuk's expansion arithmetic, "Adaptive
// Precision Floating-Point Arithmetic and Fast Robust Geometric Predicates",
// 1997): error-free differences and products, then summed into a
// non-overlapping expansion whose largest component carries the sign.
// The reference's distance refine uses robust orient2d (robust 1.2.0),
// whose sign is exact; any exact method gives the same sign.

SJ_HD void two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  double bv = x - a;
  double av = x - bv;
  y = (a - av) + (b - bv);
}

SJ_HD void two_prod(double a, double b, double& x, double& y) {
  x = a * b;
  y = fma(a, b, -x);
}

SJ_HD int expansion_grow(double* e, int n, double b) {
  double q = b;
  int k = 0;
  for (int i = 0; i < n; i++) {
    double s, t;
    two_sum(q, e[i], s, t);
    if (t != 0) e[k++] = t;
    q = s;
  }
  if (q != 0 || k == 0) e[k++] = q;
  return k;
}

// The exact expansion sum, out of line: rarely reached, and its 40-term
// buffer would otherwise occupy every caller's stack.
__host__ __device__ __noinline__ int orient2d_exact(double ax, double ay, double bx, double by, double cx,
                                                     double cy);

SJ_HD int orient2d_sign(double ax, double ay, double bx, double by, double cx, double cy) {
  // fast filter (Shewchuk's ccwerrboundA)
  double detl = (ax - cx) * (by - cy);
  double detr = (ay - cy) * (bx - cx);
  double det = detl - detr;
  double detsum = fabs(detl) + fabs(detr);
  const double errA = 3.3306690738754716e-16;  // (3 + 16 eps) eps, eps = 2^-53
  if (fabs(det) > errA * detsum) return det > 0 ? 1 : -1;
  return orient2d_exact(ax, ay, bx, by, cx, cy);
}

__host__ __device__ __noinline__ inline int orient2d_exact(double ax, double ay, double bx, double by, double cx,
                                                            double cy) {
  double u0, u1, v0, v1, w0, w1, z0, z1;
  two_sum(ax, -cx, u1, u0);
  two_sum(by, -cy, v1, v0);
  two_sum(ay, -cy, w1, w0);
  two_sum(bx, -cx, z1, z0);
  double u[2] = {u0, u1}, v[2] = {v0, v1}, w[2] = {w0, w1}, z[2] = {z0, z1};
  double e[40];
  int n = 0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      double p, pe;
      two_prod(u[i], v[j], p, pe);
      n = expansion_grow(e, n, p);
      n = expansion_grow(e, n, pe);
      two_prod(w[i], z[j], p, pe);
      n = expansion_grow(e, n, -p);
      n = expansion_grow(e, n, -pe);
    }
  for (int i = n - 1; i >= 0; i--) {
    if (e[i] > 0) return 1;
    if (e[i] < 0) return -1;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// hypot as computed by the reference's host (Rust f64::hypot -> glibc):
// sqrt(x^2 + y^2) followed by one correction step of Borges, "An Improved
// Algorithm for hypot(a,b)" (arXiv:1904.09481, 2019), in its FMA-free form,
// with power-of-two rescaling outside [2^-511, 2^511].

SJ_HD double hypot_corr_kernel(double ax, double ay) {
  double h = sqrt(ax * ax + ay * ay);
  double t1, t2;
  if (h <= 2.0 * ay) {
    double d = h - ay;
    t1 = ax * (2.0 * d - ax);
    t2 = (d - 2.0 * (ax - ay)) * d;
  } else {
    double d = h - ax;
    t1 = 2.0 * d * (ax - 2.0 * ay);
    t2 = (4.0 * d - ay) * ay + d * d;
  }
  h -= (t1 + t2) / (2.0 * h);
  return h;
}

SJ_HD double hypot_ref(double x, double y) {
  if (isinf(x) || isinf(y)) return INFINITY;
  if (x != x || y != y) return x + y;
  x = fabs(x);
  y = fabs(y);
  double ax = x < y ? y : x;
  double ay = x < y ? x : y;
  const double kEps = 5.551115123125783e-17;    // 2^-54
  const double kLarge = 6.703903964971299e+153;  // 2^511
  const double kTiny = 1.4916681462400413e-154;  // 2^-511
  const double kScale = 2.409919865102884e-181;  // 2^-600
  if (ax > kLarge) {
    if (ay <= ax * kEps) return ax + ay;
    return hypot_corr_kernel(ax * kScale, ay * kScale) / kScale;
  }
  if (ay < kTiny) {
    if (ax >= ay / kEps) return ax + ay;
    return hypot_corr_kernel(ax / kScale, ay / kScale) * kScale;
  }
  if (ay <= ax * kEps) return ax + ay;
  return hypot_corr_kernel(ax, ay);
}

// geo's distance_coord_to_line_generic (rust/sedona-geo-generic-alg/src/
// algorithm/line_measures/metric_spaces/euclidean/utils.rs:127-166).
SJ_HD double coord_to_line(double px, double py, double sx, double sy, double ex, double ey) {
  if (sx == ex && sy == ey) return hypot_ref(px - sx, py - sy);
  double dx = ex - sx, dy = ey - sy;
  double d2 = dx * dx + dy * dy;
  double r = ((px - sx) * dx + (py - sy) * dy) / d2;
  if (r <= 0) return hypot_ref(px - sx, py - sy);
  if (r >= 1) return hypot_ref(px - ex, py - ey);
  double s = ((sy - py) * dx - (sx - px) * dy) / d2;
  return fabs(s) * hypot_ref(dx, dy);
}

// intersects/mod.rs:118-138
SJ_HD bool in_between(double v, double b1, double b2) {
  if (b1 < b2) return v >= b1 && v <= b2;
  return v >= b2 && v <= b1;
}

// intersects/line.rs:27-43
SJ_HD bool line_intersects_coord(double sx, double sy, double ex, double ey, double px, double py) {
  return orient2d_sign(sx, sy, ex, ey, px, py) == 0 && in_between(px, sx, ex) && in_between(py, sy, ey);
}

// Rust f64::min (NaN-ignoring).
SJ_HD double rmin(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a < b ? a : b;
}

// Certified FP32 form of the DWithin term of one segment, over ring-local
// FP32 coordinates (a, b the segment, p the point):
// 1 when the reference's FP64 term is certainly <= d, 0 when certainly
// > d, -1 when it is not certain (the caller then evaluates the FP64 term).
//   Inputs: a, b within E of the exact local vertices per coordinate and the
//   point within Ep; the distance from a point to a segment moves by at most
//   the displacement of the point plus that of the farther endpoint, so the
//   rounded inputs change it by at most sqrt(2) (E + Ep).
//   Arithmetic: the projection branch (endpoint a, endpoint b, or the
//   perpendicular |w x v| / |v|) is evaluated in FP32; each form errs by at
//   most 12 u (|w| + |w2|) (u = 2^-24; w = p - a, w2 = p - b), and near a
//   branch boundary the two forms agree to the same order.
//   The reference's FP64 term (geo's distance_coord_to_line) errs by at most
//   16 u64 (|w| + |w2|) for the same reason; 1e-14 (|w| + |w2|) covers it.
// The FP32 distance must exceed d by more than the sum of these bounds (x
// 1.001) to decide 0, and fall short of it by as much to decide 1.
SJ_HD int segment_within_f32(float ax, float ay, float bx, float by, float pxl, float pyl, float Eab, float Ep,
                             double d) {
  const float vx = bx - ax, vy = by - ay;
  const float wx = pxl - ax, wy = pyl - ay;
  const float w2x = pxl - bx, w2y = pyl - by;
  const float nw = sqrtf(wx * wx + wy * wy), nw2 = sqrtf(w2x * w2x + w2y * w2y);
  const float c1 = wx * vx + wy * vy, c2 = vx * vx + vy * vy;
  float dist;
  if (c1 <= 0.f) dist = nw;
  else if (c1 >= c2) dist = nw2;
  else dist = fabsf(wx * vy - wy * vx) / sqrtf(c2);
  const float m = (1.4142136f * (Eab + Ep) + (12.f * 5.9604645e-08f + 1e-14f) * (nw + nw2)) * 1.001f;
  if ((double)dist + (double)m <= d) return 1;
  if ((double)dist - (double)m > d) return 0;
  return -1;
}

// ---------------------------------------------------------------------------
// Rects of geometries (a1/a2).

// f64 interval -> f32, outward (rust/sedona-spatial-join/src/utils/bounds.rs:147-162)
SJ_HD float2 outward(double lo, double hi) {
  float l = (float)lo, h = (float)hi;
  if ((double)l > lo) l = nextafterf(l, -INFINITY);
  if ((double)h < hi) h = nextafterf(h, INFINITY);
  return make_float2(l, h);
}

SJ_HD float4 finish_rect(double xl, double yl, double xh, double yh, bool nearest_point, double expand) {
  float4 r;
  if (nearest_point) {
    r = make_float4((float)xl, (float)yl, (float)xh, (float)yh);
  } else {
    float2 x = outward(xl, xh), y = outward(yl, yh);
    r = make_float4(x.x, y.x, x.y, y.y);
  }
  if (expand == expand && expand >= 0.0) {  // interval.rs:269-275
    float2 x = outward((double)r.x - expand, (double)r.z + expand);
    float2 y = outward((double)r.y - expand, (double)r.w + expand);
    r = make_float4(x.x, y.x, x.y, y.y);
  }
  return r;
}

}  // namespace sjhip
