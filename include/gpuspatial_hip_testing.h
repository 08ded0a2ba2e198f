/*
 * gpuspatial_hip_testing.h -- host-side evaluation of the library's scalar
 * geometry predicates (the same source the device kernels compile), for
 * parity tests that run without a GPU.  No counterpart in the reference;
 * callers are test suites only.
 */
#ifndef GPUSPATIAL_HIP_TESTING_H
#define GPUSPATIAL_HIP_TESTING_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Ray classification of tg.c:1199-1275: 0 out, 1 in, 2 on. */
int SjhipHostRaycast(double ax, double ay, double bx, double by, double px, double py);
/* Certified FP32 form: 0 out, 1 in, -1 undecided (FP64 fallback needed). */
int SjhipHostRaycastF32(float ax, float ay, float bx, float by, float px, float py, float err_bound);
/* Collinearity test of tg.c:861-901. */
int SjhipHostCollinear(double x1, double y1, double x2, double y2, double x3, double y3);
/* Segment intersection of tg.c:1392-1443. */
int SjhipHostSegmentIntersects(double ax, double ay, double bx, double by, double cx, double cy, double dx, double dy);
/* Exact orientation sign. */
int SjhipHostOrient2d(double ax, double ay, double bx, double by, double cx, double cy);
/* hypot as the reference host computes it. */
double SjhipHostHypot(double x, double y);
/* Point to segment distance of the reference's geo refine. */
double SjhipHostCoordToLine(double px, double py, double sx, double sy, double ex, double ey);
/* Certified FP32 DWithin term of one segment in ring-local coordinates
 * (inputs within e_ab / e_p of exact per coordinate): 1 certainly within
 * d, 0 certainly not, -1 undecided (FP64 fallback needed). */
int SjhipHostSegmentWithinF32(float ax, float ay, float bx, float by, float px, float py, float e_ab, float e_p,
                              double d);
/* Certified FP32 form of the segment intersection test for segments a-b
 * and c-d given within e of exact per coordinate: 1 certainly
 * intersecting, 0 certainly not, -1 undecided (FP64 fallback needed). */
int SjhipHostSegmentsCrossF32(float ax, float ay, float bx, float by, float cx, float cy, float dx, float dy, float e);
/* Batched forms over arrays (n entries). */
void SjhipHostRaycastBatch(const double* seg_pts, int64_t n, int32_t* out);
void SjhipHostHypotBatch(const double* xy, int64_t n, double* out);

/* Device test hook: the kernels-only stable radix sort that captured
 * request sequences use, over device buffers of n (u32 key, u32 value)
 * pairs on device `device_id` (keys ordered by bits [begin_bit, end_bit)).
 * Returns 0, or EINVAL with the message in SjhipLastError(). */
int SjhipTestingGraphSortPairs(const uint32_t* keys, const uint32_t* values, uint32_t* keys_out,
                               uint32_t* values_out, int64_t n, int begin_bit, int end_bit, int device_id);

#ifdef __cplusplus
}
#endif
#endif
