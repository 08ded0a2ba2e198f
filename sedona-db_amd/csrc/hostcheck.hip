// hostcheck.hip -- host entry points over geom_math.h (see
// include/gpuspatial_hip_testing.h).  Compiled with the same flags as the
// device code, so the formulas the kernels use can be checked against the
// oracle on machines without a GPU.
#include "../../include/gpuspatial_hip_testing.h"
#include "geom_math.h"

using namespace sjhip;


extern "C" {
int SjhipHostRaycast(double ax, double ay, double bx, double by, double px, double py) {
  return raycast(ax, ay, bx, by, px, py);
}
int SjhipHostRaycastF32(float ax, float ay, float bx, float by, float px, float py, float e) {
  return raycast_f32_certified(ax, ay, bx, by, px, py, e);
}
int SjhipHostCollinear(double x1, double y1, double x2, double y2, double x3, double y3) {
  return collinear(x1, y1, x2, y2, x3, y3) ? 1 : 0;
}
int SjhipHostSegmentIntersects(double ax, double ay, double bx, double by, double cx, double cy, double dx,
                               double dy) {
  return segment_intersects(ax, ay, bx, by, cx, cy, dx, dy) ? 1 : 0;
}
int SjhipHostOrient2d(double ax, double ay, double bx, double by, double cx, double cy) {
  return orient2d_sign(ax, ay, bx, by, cx, cy);
}
double SjhipHostHypot(double x, double y) { return hypot_ref(x, y); }
double SjhipHostCoordToLine(double px, double py, double sx, double sy, double ex, double ey) {
  return coord_to_line(px, py, sx, sy, ex, ey);
}
void SjhipHostRaycastBatch(const double* s, int64_t n, int32_t* out) {
  for (int64_t i = 0; i < n; i++) out[i] = raycast(s[6 * i], s[6 * i + 1], s[6 * i + 2], s[6 * i + 3], s[6 * i + 4], s[6 * i + 5]);
}
int SjhipHostSegmentWithinF32(float ax, float ay, float bx, float by, float px, float py, float e_ab, float e_p,
                              double d) {
  return segment_within_f32(ax, ay, bx, by, px, py, e_ab, e_p, d);
}
int SjhipHostSegmentsCrossF32(float ax, float ay, float bx, float by, float cx, float cy, float dx, float dy, float e) {
  return segments_cross_f32_certified(ax, ay, bx, by, cx, cy, dx, dy, e);
}
void SjhipHostHypotBatch(const double* xy, int64_t n, double* out) {
  for (int64_t i = 0; i < n; i++) out[i] = hypot_ref(xy[2 * i], xy[2 * i + 1]);
}
}
