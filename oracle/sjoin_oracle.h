/*
 * sjoin_oracle.h -- CPU restatement of the reference CPU spatial join.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and never by the product path (sedona-db_amd/), which must fail loudly
 * when its HIP library is missing.
 *
 * What it restates (reference paths relative to the apache/sedona-db tree):
 *   - WKB decode + per-row bounds: rust/sedona-geometry/src/bounds.rs:240-249,
 *     rust/sedona-spatial-join/src/utils/bounds.rs:49-63,147-162 (outward f32),
 *     rust/sedona-spatial-join-gpu/src/join_provider.rs:94-146 (points: nearest f32).
 *   - distance expansion of probe rects: rust/sedona-spatial-join/src/
 *     operand_evaluator.rs:67-107,605-607; rust/sedona-geometry/src/interval.rs:269-275.
 *   - relation refine (TG semantics): c/sedona-tg/src/tg/tg.c:816-823 (fmin0/fmax0),
 *     853-901 (feq, collinear), 1076-1078 (pteq), 1199-1275 (raycast),
 *     2126-2134 (num_segments), 2669-2822 (pip_eval_seg*, standard_pip,
 *     tg_ring_contains_point), 3113-3118 (tg_ring_empty), 3909-3955
 *     (poly_contains_point), 1392-1443 (segment x segment), 3089-3111,
 *     3036-3087, 3166-3190 (ring intersects segment / ring), 4100-4136
 *     (poly intersects poly), 6053-6100 and 6185-6290 (geometry dispatch).
 *   - distance refine (geo semantics): rust/sedona-geo-generic-alg/src/algorithm/
 *     line_measures/metric_spaces/euclidean/utils.rs:96-271, intersects/line.rs:27-43,
 *     kernels/robust.rs:35-65; compared with <= per
 *     rust/sedona-spatial-join/src/refine/geo.rs:312-335.
 *   - KNN ground truth: rust/sedona-spatial-join/tests/spatial_join_integration.rs:1373-1414
 *     (sort by distance, then by id; take k).
 */
#ifndef SJOIN_ORACLE_H
#define SJOIN_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry type codes (WKB 2-D base types). 0 = null row. */
enum { ORA_NULL = 0, ORA_POINT = 1, ORA_LINESTRING = 2, ORA_POLYGON = 3,
       ORA_MULTIPOINT = 4, ORA_MULTILINESTRING = 5, ORA_MULTIPOLYGON = 6 };

/* Predicates: same numbering as SedonaSpatialRelationPredicate, plus
 * ORA_PRED_DWITHIN for the distance join. */
enum { ORA_PRED_EQUALS = 0, ORA_PRED_DISJOINT = 1, ORA_PRED_TOUCHES = 2,
       ORA_PRED_CONTAINS = 3, ORA_PRED_COVERS = 4, ORA_PRED_INTERSECTS = 5,
       ORA_PRED_WITHIN = 6, ORA_PRED_COVEREDBY = 7, ORA_PRED_DWITHIN = 100 };

/* Flat geometry store: row -> parts -> rings -> points. */
typedef struct ora_store {
  int64_t n_geoms, n_parts, n_rings, n_pts;
  uint8_t* type;      /* n_geoms */
  int64_t* part_off;  /* n_geoms + 1 */
  int64_t* ring_off;  /* n_parts + 1 */
  int64_t* pt_off;    /* n_rings + 1 */
  double* xy;         /* 2 * n_pts */
} ora_store;

/* Decode an Arrow Binary/LargeBinary column given as int64 offsets (n+1),
 * data bytes and an optional validity byte-per-row array (1 = valid).
 * Returns 0 or a negative error code; *bad_row receives the failing row. */
int ora_store_from_wkb(const int64_t* offsets, const uint8_t* data,
                       const uint8_t* valid, int64_t n, ora_store** out,
                       int64_t* bad_row);
void ora_store_free(ora_store* s);
void ora_store_sizes(const ora_store* s, int64_t* out4);
/* Copy the SoA arrays out (caller sized them with ora_store_sizes). */
void ora_store_export(const ora_store* s, uint8_t* type, int64_t* part_off,
                      int64_t* ring_off, int64_t* pt_off, double* xy);

/* Per-row f32 rects [xmin,ymin,xmax,ymax]; empty rows -> NaN x4.
 * point_nearest != 0: points get the round-to-nearest degenerate rect
 * (GPU operator semantics); otherwise every row is rounded outward. */
void ora_rects(const ora_store* s, int point_nearest, float* out);
/* Expand f32 rects by a distance in f64 and round outward again. */
void ora_expand_rects(float* rects, int64_t n, double distance);

/* f64 -> f32 outward rounding of an interval (bounds.rs:147-162). */
void ora_f64_interval_to_f32(double lo, double hi, float* out_lo, float* out_hi);

/* Candidate filter: all (b,p) whose f32 rects intersect (inclusive);
 * NaN/inverted rects never match.  Output sorted by (probe, build).
 * Returns the pair count; when out arrays are NULL only counts. */
int64_t ora_filter(const float* build_rects, int64_t nb, const float* probe_rects,
                   int64_t np, uint32_t* out_build, uint32_t* out_probe,
                   int64_t cap, int nthreads);

/* Evaluate a predicate on candidate pairs.  mask[i] = 1 if true.
 * Returns 0, or -1 if a pair has an unsupported type combination. */
int ora_refine(const ora_store* build, const ora_store* probe, int pred,
               double distance, const uint32_t* bidx, const uint32_t* pidx,
               int64_t n, uint8_t* mask, int nthreads);

/* Full join = filter + refine; returns refined pair count (or <0 on error).
 * Pairs are sorted by (probe, build). */
int64_t ora_join(const ora_store* build, const ora_store* probe, int pred,
                 double distance, uint32_t* out_build, uint32_t* out_probe,
                 int64_t cap, int nthreads);

/* Distance from build geometry b to probe geometry p (geo semantics). */
double ora_distance(const ora_store* build, int64_t b, const ora_store* probe, int64_t p);

/* Exact KNN over point objects: for each probe point, the k nearest build
 * points by (hypot distance, build id).  Outputs n_probe*k entries; unused
 * slots get id 0xFFFFFFFF and distance +inf. */
int ora_knn(const ora_store* build, const ora_store* probe, int k,
            uint32_t* out_ids, double* out_dist, int nthreads);

/* Single-point helpers exposed for known-answer tests. */
int ora_raycast(double ax, double ay, double bx, double by, double px, double py);
int ora_collinear(double x1, double y1, double x2, double y2, double x3, double y3);
int ora_orient2d_sign(double ax, double ay, double bx, double by, double cx, double cy);
int ora_segment_intersects(double ax, double ay, double bx, double by,
                           double cx, double cy, double dx, double dy);

/* KNN by brute force over any geometry, one side points, rows of variable
 * length (tol >= 0: tie breakers within tol of the k-th distance). */
int64_t ora_knn_brute(const ora_store* B, const ora_store* P, int k, double tol, int64_t* out_off,
                      uint32_t* out_ids, double* out_d, int64_t cap, int nthreads);

/* ora_distance over n (build, probe) pairs. */
void ora_pair_distances(const ora_store* B, const ora_store* P, const uint32_t* bidx, const uint32_t* pidx,
                        int64_t n, double* out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
