/*
 * gpuspatial_hip.h -- C ABI of the MI355X spatial-join library (libsjhip.so).
 *
 * Part 1 declares exactly the structs and three create functions that the
 * reference's Rust FFI binds (bindgen over
 * c/sedona-libgpuspatial/libgpuspatial/include/gpuspatial/gpuspatial_c.h,
 * driven from c/sedona-libgpuspatial/build.rs:263-279).  The layouts,
 * member order, names (including the ABI's "precent" spelling) and
 * semantics are kept identical so that the Rust crates
 * c/sedona-libgpuspatial/src/libgpuspatial.rs and
 * rust/sedona-spatial-join-gpu/src/index/ (Rust) link against this library
 * unchanged.
 *
 * Part 2 declares additive extensions that the reference has no GPU
 * counterpart for: the distance refine (ST_DWithin), the KNN join, and a
 * device-resident fused join plus request loop.  Each cites the reference
 * interface it serves.
 *
 * Conventions for every entry point: 0 on success, EINVAL (or another
 * errno value) on failure with the message available from the object's
 * get_last_error; all input buffers are borrowed for the duration of the
 * call; the library never releases caller-owned Arrow arrays.
 */
#ifndef GPUSPATIAL_HIP_H
#define GPUSPATIAL_HIP_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arrow C Data Interface and C Device Data Interface: public Arrow
 * specification structs, guarded as the specification prescribes so that
 * they coexist with arrow/c/abi.h or nanoarrow. */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4
struct ArrowSchema {
  const char* format;
  const char* name;
  const char* metadata;
  int64_t flags;
  int64_t n_children;
  struct ArrowSchema** children;
  struct ArrowSchema* dictionary;
  void (*release)(struct ArrowSchema*);
  void* private_data;
};
struct ArrowArray {
  int64_t length;
  int64_t null_count;
  int64_t offset;
  int64_t n_buffers;
  int64_t n_children;
  const void** buffers;
  struct ArrowArray** children;
  struct ArrowArray* dictionary;
  void (*release)(struct ArrowArray*);
  void* private_data;
};
#endif

#ifndef ARROW_C_DEVICE_DATA_INTERFACE
#define ARROW_C_DEVICE_DATA_INTERFACE
typedef int32_t ArrowDeviceType;
#define ARROW_DEVICE_CPU 1
#define ARROW_DEVICE_ROCM 10
/* sync_event: NULL, or (for ROCm arrays) a pointer to the producer's
 * hipEvent_t.  Every entry point taking an ArrowDeviceArray waits on the
 * host for that event before it reads or queues the buffers. */
struct ArrowDeviceArray {
  struct ArrowArray array;
  int64_t device_id;
  ArrowDeviceType device_type;
  void* sync_event;
  int64_t reserved[3];
};
#endif

/* ======================================================================
 * Part 1: the reference ABI (gpuspatial_c.h)
 * ==================================================================== */

/* gpuspatial_c.h:28-37.  ptx_root: the reference loads OptiX PTX from this
 * directory; this library embeds its gfx950 code objects and only checks
 * that the directory exists, failing with "No such file or directory" like
 * the reference (test/c_wrapper_test.cc:45-65). */
struct GpuSpatialRuntimeConfig {
  const char* ptx_root;
  int device_id;
  bool use_cuda_memory_pool;
  int cuda_memory_pool_init_precent;
};

/* gpuspatial_c.h:39-50: one instance per process. */
struct GpuSpatialRuntime {
  int (*init)(struct GpuSpatialRuntime* self, struct GpuSpatialRuntimeConfig* config);
  void (*release)(struct GpuSpatialRuntime* self);
  const char* (*get_last_error)(struct GpuSpatialRuntime* self);
  void* private_data;
};

/* gpuspatial_c.h:53 */
void GpuSpatialRuntimeCreate(struct GpuSpatialRuntime* runtime);

/* gpuspatial_c.h:55-60 */
struct GpuSpatialIndexConfig {
  const struct GpuSpatialRuntime* runtime;
  uint32_t concurrency;
};

/* gpuspatial_c.h:62-65 */
struct SedonaSpatialIndexContext {
  void* private_data;
};

/* gpuspatial_c.h:67-122: FP32 2-D rectangle index.  push_build takes
 * n_rects x [min_x, min_y, max_x, max_y]; NaN or inverted rects are empty
 * and never match.  probe reports every (build, probe) pair whose rects
 * intersect (boundaries inclusive) through the callback, possibly in
 * several calls; a non-zero callback return stops the probe and is
 * returned by probe. */
struct SedonaFloatIndex2D {
  int (*clear)(struct SedonaFloatIndex2D* self);
  void (*create_context)(struct SedonaSpatialIndexContext* context);
  void (*destroy_context)(struct SedonaSpatialIndexContext* context);
  int (*push_build)(struct SedonaFloatIndex2D* self, const float* buf, uint32_t n_rects);
  int (*finish_building)(struct SedonaFloatIndex2D* self);
  int (*probe)(struct SedonaFloatIndex2D* self, struct SedonaSpatialIndexContext* context,
               const float* buf, uint32_t n_rects,
               int (*callback)(const uint32_t* build_indices, const uint32_t* probe_indices,
                               uint32_t length, void* user_data),
               void* user_data);
  const char* (*get_last_error)(struct SedonaFloatIndex2D* self);
  const char* (*context_get_last_error)(struct SedonaSpatialIndexContext* context);
  void (*release)(struct SedonaFloatIndex2D* self);
  void* private_data;
};

/* gpuspatial_c.h:127-128 */
int GpuSpatialIndexFloat2DCreate(struct SedonaFloatIndex2D* index,
                                 const struct GpuSpatialIndexConfig* config);

/* gpuspatial_c.h:130-140 */
struct GpuSpatialRefinerConfig {
  const struct GpuSpatialRuntime* runtime;
  uint32_t concurrency;
  bool compress_bvh;
  uint32_t pipeline_batches;
};

/* gpuspatial_c.h:142-151 (values 0..7). */
enum SedonaSpatialRelationPredicate {
  SedonaSpatialPredicateEquals = 0,
  SedonaSpatialPredicateDisjoint,
  SedonaSpatialPredicateTouches,
  SedonaSpatialPredicateContains,
  SedonaSpatialPredicateCovers,
  SedonaSpatialPredicateIntersects,
  SedonaSpatialPredicateWithin,
  SedonaSpatialPredicateCoveredBy
};

/* gpuspatial_c.h:154-211.  Geometry columns are Arrow Binary ("z"),
 * LargeBinary ("Z") or BinaryView ("vz") arrays of 2-D WKB.  refine
 * evaluates predicate(build, probe) for each candidate pair, rewrites the
 * caller's arrays in place with the pairs that hold, sorted by (probe,
 * build), and sets *new_indices_size.  Supported, with TG's semantics:
 * Intersects, Contains, Within, Covers, CoveredBy, Touches and Equals for
 * Point x Point, Point x (Multi)Polygon in both roles and (Multi)Polygon x
 * (Multi)Polygon; Intersects between (Multi)LineString and Point,
 * (Multi)LineString or (Multi)Polygon in either role.  Other combinations
 * fail with EINVAL. */
struct SedonaSpatialRefiner {
  int (*clear)(struct SedonaSpatialRefiner* self);
  int (*init_build_schema)(struct SedonaSpatialRefiner* self, const struct ArrowSchema* build_schema);
  int (*push_build)(struct SedonaSpatialRefiner* self, const struct ArrowArray* build_array);
  int (*finish_building)(struct SedonaSpatialRefiner* self);
  int (*refine)(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                const struct ArrowArray* probe_array, enum SedonaSpatialRelationPredicate predicate,
                uint32_t* build_indices, uint32_t* probe_indices, uint32_t indices_size,
                uint32_t* new_indices_size);
  const char* (*get_last_error)(struct SedonaSpatialRefiner* self);
  void (*release)(struct SedonaSpatialRefiner* self);
  void* private_data;
};

/* gpuspatial_c.h:216-217 */
int GpuSpatialRefinerCreate(struct SedonaSpatialRefiner* refiner,
                            const struct GpuSpatialRefinerConfig* config);

/* ======================================================================
 * Part 2: extensions
 * ==================================================================== */

#define SJHIP_EXT_VERSION 1
int SjhipExtVersion(void);

/* Distance refine for ST_DWithin(build, probe, distance).  Replaces the
 * reference's CPU Geo refiner (rust/sedona-spatial-join/src/refine/geo.rs:
 * 312-335: build.distance_ext(probe) <= distance) that the GPU operator
 * rejects today (rust/sedona-spatial-join-gpu/src/index/gpu_spatial_index.rs:
 * 119-121).  Probe rows must be points; build rows Point, LineString,
 * Polygon or their Multi* forms.  In/out contract as refine. */
int SjhipRefinerRefineDistance(struct SedonaSpatialRefiner* self,
                               const struct ArrowSchema* probe_schema,
                               const struct ArrowArray* probe_array, double distance,
                               uint32_t* build_indices, uint32_t* probe_indices,
                               uint32_t indices_size, uint32_t* new_indices_size);

/* KNN join: for every probe row the k nearest build rows by Euclidean
 * distance (geo's, FP64: hypot between points, point to (multi)linestring
 * or (multi)polygon), ties broken by build id, k in 1..32.  Replaces the
 * CPU SpatialIndex::query_knn (rust/sedona-spatial-join/src/index/
 * default_spatial_index.rs:306-502) that the GPU index leaves
 * unimplemented (rust/sedona-spatial-join-gpu/src/index/
 * gpu_spatial_index.rs:252-264).  Outputs are row-major n_probe x k; a
 * probe with fewer than k neighbours gets build id 0xFFFFFFFF and distance
 * +inf in the unused slots.  One side must be points: point probes against
 * build rows of any geometry, or probes of any geometry against point
 * build rows. */
/* The distance of each (build row, probe point) pair, FP64, as the
 * reference's distance refine computes it (geo's Euclidean distance,
 * rust/sedona-geo-generic-alg/src/algorithm/line_measures/metric_spaces/
 * euclidean/utils.rs:96-271); NaN for a null or empty row.  The ST_Distance
 * values behind an ST_DWithin decision; no counterpart in the reference's
 * C ABI.  Probe rows must be points. */
int SjhipRefinerPairDistances(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                              const struct ArrowArray* probe_array, const uint32_t* build_indices,
                              const uint32_t* probe_indices, uint64_t n, double* out_distances);

/* KNN with the reference's include_tie_breakers option
 * (default_spatial_index.rs:379-472): as SjhipRefinerKnn, and with
 * include_tie_breakers != 0 every further build row whose distance is
 * within DISTANCE_TOLERANCE = 1e-9 of the k-th distance is kept as well
 * (index/spatial_index.rs:32).  Rows have variable length: probe row q's
 * neighbours are entries [out_offsets[q], out_offsets[q+1]) of
 * out_build_indices / out_distances, sorted by (distance, build id); a
 * null or empty probe has none.  out_offsets holds n_probe + 1 entries,
 * the other two `capacity`; when the total exceeds capacity the call
 * fails with EINVAL and *out_total holds it (retry with larger arrays).
 * The order among ties at the k-th distance is by build id, as in the
 * reference's test ground truth. */
int SjhipRefinerKnnTies(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                        const struct ArrowArray* probe_array, uint32_t k, int32_t include_tie_breakers,
                        uint64_t* out_offsets, uint32_t* out_build_indices, double* out_distances, uint64_t capacity,
                        uint64_t* out_total);

int SjhipRefinerKnn(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                    const struct ArrowArray* probe_array, uint32_t k, uint32_t* out_build_indices,
                    double* out_distances);

/* SjhipRefinerKnn for a probe column in ROCm device (or CPU) memory, with
 * the outputs (n_probe x k, as for SjhipRefinerKnn) in device memory the
 * caller allocated.  The object tree over the build points is built by the
 * first KNN call and stays resident with the refiner.  *kernel_ms (if not
 * NULL) receives the device time of the KNN kernel. */
int SjhipRefinerKnnDevice(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                          const struct ArrowDeviceArray* probe_array, uint32_t k, uint32_t* out_build_indices,
                          double* out_distances, double* kernel_ms);

/* Result of a device-resident join: device pointers owned by the library,
 * valid until release is called. */
struct SjhipJoinResult {
  uint32_t* build_indices; /* device memory */
  uint32_t* probe_indices; /* device memory */
  uint64_t size;
  uint64_t candidates;      /* filter output count */
  uint64_t fallback_tests;  /* FP32 segment tests too close to call: decided
                               in FP64 (their candidate re-refined in FP64
                               by the point-in-polygon kernel) */
  uint64_t segment_tests;   /* segment tests performed */
  uint64_t refine_candidates; /* candidates the refine's walk decided: those
                                 the cell-raster passes left open */
  /* device time of the stages (ms, HIP events on the join's stream) when
   * options.profile is set: [0] probe decode + rects, [1] Morton order of
   * probes, [2] filter kernel, [3] refine kernels, [4] copy of the kept
   * pairs into the result (exact size) and output sort, [5] candidate
   * regrouping (refine_order 1; for the point-in-polygon refine with
   * cell rasters, the raster split before it) */
  double stage_ms[6];
  void (*release)(struct SjhipJoinResult* self);
  void* private_data;
};

/* Join options for the device-resident join. */
struct SjhipJoinOptions {
  int32_t predicate;   /* SedonaSpatialRelationPredicate, or 100 = DWithin */
  double distance;     /* DWithin threshold */
  int32_t precision;   /* 0 = FP32 certified + FP64 fallback, 1 = all-FP64 */
  int32_t sort_output; /* 1 = pairs sorted by (probe, build) */
  int32_t sort_probes; /* 1 = visit probes in Morton order */
  int32_t profile;     /* 1 = time the stages with HIP events (stage_ms) */
  int32_t refine_order; /* 0 = refine candidates in filter order, 1 = group
                           them by build row first (polygon-major), 2 =
                           (DWithin, point probes) decide the distance inside
                           the filter walk, writing no candidates; -1 = the
                           order SjhipAutotune chose (else 1; DWithin 0
                           unless the build side is single lines, whose
                           cell rasters drop candidates before the
                           grouping) */
  void* stream;        /* hipStream_t, or NULL for a stream of the call */
};

/* One probe batch of the GPU operator's query_batch
 * (rust/sedona-spatial-join-gpu/src/index/gpu_spatial_index.rs:163-229)
 * with filter and refine fused on the device: the probe WKB column is
 * already in device memory (ArrowDeviceArray with device_type ROCm) or in
 * host memory (device_type CPU, copied in).  The index and refiner must
 * have been built from the same build rows in the same order. */
int SjhipJoinDevice(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                    const struct ArrowSchema* probe_schema,
                    const struct ArrowDeviceArray* probe_array,
                    const struct SjhipJoinOptions* options, struct SjhipJoinResult* out);

/* Statistics of a partitioned join. */
struct SjhipPartitionStats {
  uint32_t partitions;          /* build partitions joined */
  uint64_t max_partition_rows;  /* rows of the largest partition */
  uint64_t max_partition_bytes; /* its estimated resident device bytes */
  uint64_t pairs;               /* pairs delivered */
  double build_ms;              /* building the partitions' refiners and indexes (host gather included) */
  double join_ms;               /* the partitions' joins, pairs copied back */
  double total_ms;              /* the whole call */
};

/* A join whose build side need not fit in device memory: the host build
 * column (Binary or LargeBinary WKB) is split into partitions whose
 * estimated resident size is at most device_budget_bytes, K-D-B style
 * (rust/sedona-spatial-join/src/partitioning/kdb.rs:36-40: along the
 * longer axis, at the median by bytes), every row in exactly one
 * partition; rows with an empty rect (null or empty) meet nothing and
 * are left out.  Partitions are built and joined one at a time against
 * the whole probe column (ROCm device memory or CPU memory) as
 * SjhipJoinDevice does, then released: the device holds one partition at
 * a time, the reference's memory-quota batching of the build side
 * (c/sedona-libgpuspatial/libgpuspatial/src/relate_engine.cu:468-479).
 * Each build row is in one partition, so every pair comes from exactly one
 * partition and no pair is repeated.  callback receives each partition's
 * pairs (build indices of the whole column, probe indices of the probe
 * column), from pinned memory valid during the call; its nonzero return
 * stops the join and is returned (one call with n = 0 when no build row
 * has a rect).  options.stream must be NULL; options.sort_output sorts
 * within a partition.  stats may be NULL. */
int SjhipPartitionedJoin(struct GpuSpatialRuntime* runtime, const struct ArrowSchema* build_schema,
                         const struct ArrowArray* build_array, const struct ArrowSchema* probe_schema,
                         const struct ArrowDeviceArray* probe_array, const struct SjhipJoinOptions* options,
                         uint64_t device_budget_bytes,
                         int (*callback)(const uint32_t*, const uint32_t*, uint32_t, void*), void* user_data,
                         struct SjhipPartitionStats* stats);

/* The join of a host-resident probe column (Binary or LargeBinary WKB in
 * CPU memory) in batches of batch_rows rows with `slots` (1..8) batches in
 * flight: each slot copies its batch in through pinned blocks, joins it on
 * its own stream as SjhipJoinDevice does, and copies the pairs back, so
 * that the copies of one batch overlap the join of another (the
 * reference's two-slot refine pipeline, c/sedona-libgpuspatial/
 * libgpuspatial/src/rt_spatial_refiner.cu:244-436, and pinned staging,
 * parallel_wkb_loader.hpp:505-534).  callback receives each batch's pairs
 * in batch order, probe indices global to the column, from pinned memory
 * valid during the call (one call with n = 0 for an empty column); its
 * nonzero return stops the join and is returned.  options.stream must be
 * NULL; with options.sort_output each batch (and so the whole output) is
 * sorted by (probe, build). */
int SjhipJoinHostStream(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                        const struct ArrowSchema* probe_schema, const struct ArrowArray* probe_array,
                        const struct SjhipJoinOptions* options, int64_t batch_rows, int32_t slots,
                        int (*callback)(const uint32_t*, const uint32_t*, uint32_t, void*), void* user_data);

/* The join of SjhipJoinDevice without allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  It allocates nothing: its buffers live in a library-owned
 * static device array of SJHIP_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends, and the pairs go to the caller's device arrays
 * out_build_indices / out_probe_indices of out_capacity entries, in refine
 * order.  Limits: a device-resident Binary or LargeBinary column of at most
 * SJHIP_NOALLOC_MAX_ROWS 21-byte WKB points without nulls; at most
 * SJHIP_NOALLOC_MAX_CANDIDATES candidate pairs; options.precision 0 and
 * options.sort_output 0; ST_Intersects, ST_Contains or ST_Covers over a
 * (multi)polygon build side, or ST_DWithin (predicate 100).  Beyond a limit
 * the call fails with EINVAL and a message; when only the output capacity
 * is short, *out_size holds the pair count.  options.stream (or the null
 * stream) runs the kernels; the call returns when they are done.  Every
 * call on one GPU uses that GPU's one array, whose contents the next call
 * overwrites: calls from threads of this process are serialised per GPU,
 * and calls from other processes have arrays of their own. */
#define SJHIP_NOALLOC_MAX_ROWS 65536
#define SJHIP_NOALLOC_MAX_CANDIDATES (1u << 20)
#define SJHIP_NOALLOC_SCRATCH_BYTES 17384192ull
int SjhipJoinDeviceNoAlloc(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                           const struct ArrowSchema* probe_schema, const struct ArrowDeviceArray* probe_array,
                           const struct SjhipJoinOptions* options, uint32_t* out_build_indices,
                           uint32_t* out_probe_indices, uint64_t out_capacity, uint64_t* out_size);

/* Parameters chosen by SjhipAutotune. */
struct SjhipTuneResult {
  int32_t leaf;         /* rect tree leaf size (4, 8 or 16 rects) */
  int32_t refine_order; /* candidate order before the refine (SjhipJoinOptions.refine_order) */
  int32_t pip_waves;    /* resident waves per SIMD of the point-in-polygon refine */
  int32_t evaluations;  /* configurations timed */
  double best_ms;       /* wall time of the chosen configuration on the sample */
  int32_t seg_leaf;     /* segments per leaf of the build rings' trees (16 or 32) */
};

/* Autotuner (no counterpart in the reference): times joins of a probe
 * sample against the built index and refiner under the options' predicate
 * and distance, choosing in turn the rect tree's leaf size (the tree is
 * rebuilt), the candidate order before the refine (per kind of refine:
 * point in polygon, distance, other), the point-in-polygon kernel's
 * resident waves and, for point-in-polygon samples, the segments per leaf
 * of the build rings' trees (the ring index is rebuilt).  The choices stay
 * with the index and refiner: joins whose options.refine_order is -1 use
 * the tuned order.  Must not run concurrently with joins on the same index
 * or refiner. */
int SjhipAutotune(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                  const struct ArrowSchema* probe_schema, const struct ArrowDeviceArray* probe_sample,
                  const struct SjhipJoinOptions* options, struct SjhipTuneResult* out);

/* Rebuilds the built refiner's ring index with `segments_per_leaf` (16 or
 * 32) segments per leaf of its rings' segment trees (the choice the
 * autotuner makes; no counterpart in the reference).  Must not run
 * concurrently with joins on the refiner. */
int SjhipRefinerSetSegmentLeaf(struct SedonaSpatialRefiner* refiner, int32_t segments_per_leaf);

/* Sets the join parameters the autotuner would choose for the index, so
 * that a run can reproduce a tuned configuration without tuning: the
 * candidate order for joins whose refine is `kind` (0 point in polygon,
 * 1 distance, 2 other relations; refine_order 0, 1 or 2, or -1 to clear)
 * and the point-in-polygon kernel's resident waves (4, 6 or 8; 0 keeps the
 * current value).  No counterpart in the reference.  Must not run
 * concurrently with joins on the index. */
int SjhipIndexSetTuning(struct SedonaFloatIndex2D* index, int32_t kind, int32_t refine_order, int32_t pip_waves);

/* Filter of the index's probes: 0 (the default) walks the cell grid for
 * point probes and the rect tree for others, and in the fused point-in-
 * polygon join over Morton-ordered points (polygon-major order) lets each
 * build rect test the points of the Morton cells it covers; 1 walks the
 * tree always; 2 the grid whenever the index has one (any probe rects); 3
 * is 0 without the build-rect pass.  Every mode gives the same candidates
 * and pairs; for tests and measurements.  No counterpart in the reference.
 * Must not run concurrently with probes of the index. */
int SjhipIndexSetFilterMode(struct SedonaFloatIndex2D* index, int32_t mode);

/* Precision of the refiner's refine and SjhipRefinerRefineDistance calls:
 * 0 (the default) runs the certified FP32 tests with FP64 re-evaluation
 * where a sign is uncertain, 1 the all-FP64 plain kernels (for polygon x
 * polygon Intersects: the kernel that decides uncertain segments in FP64
 * inline, without the certified phases).  Both give the same pair set; 1
 * exists for comparisons.  No counterpart in the reference. */
int SjhipRefinerSetPrecision(struct SedonaSpatialRefiner* refiner, int32_t precision);

/* Request loop over a resident index and refiner (the serving path of the
 * GPU operator: rust/sedona-spatial-join-gpu/src/index/gpu_spatial_index.rs:
 * 163-229 called for many probe batches at once).  Requests are joined
 * with SjhipJoinDevice semantics; up to `slots` of them run concurrently on
 * the device, each on its own stream, admitted in arrival order when a
 * slot is free and the request's device-memory estimate fits
 * `memory_budget` next to the running ones (a request larger than the
 * whole budget runs alone).  No counterpart in the reference. */
struct SjhipServerConfig {
  uint32_t slots;          /* concurrent requests (>= 1) */
  uint64_t memory_budget;  /* bytes; 0 = 90 % of the device memory free at creation */
  int32_t graphs;          /* 1 = replay captured HIP graphs for requests of
                              point probes in device memory (up to 65536
                              rows; point-in-polygon predicates and DWithin):
                              one graph per slot, predicate and power-of-two
                              size bucket, captured on first use; a graph
                              holds kernels and copies only (no memset or
                              allocation node).  Larger requests, and any
                              whose decode or candidates exceed the plan,
                              take the eager path */
};

/* Timing of one request, ms on a steady clock since the server's
 * creation: submitted, admitted to a slot, finished. */
struct SjhipRequestTiming {
  double submit_ms;
  double admit_ms;
  double done_ms;
  uint64_t estimate_bytes; /* the admission estimate */
};

int SjhipServerCreate(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                      const struct SjhipServerConfig* config, void** out_server);

/* Queues one request.  The probe arrays are borrowed until the request's
 * SjhipServerWait has returned; options->stream is ignored. */
int SjhipServerSubmit(void* server, const struct ArrowSchema* probe_schema,
                      const struct ArrowDeviceArray* probe_array, const struct SjhipJoinOptions* options,
                      uint64_t* out_request_id);

/* Waits up to timeout_ms (-1: no limit, 0: poll) for a request.  Returns 0
 * with the result moved into *out (release it as for SjhipJoinDevice),
 * EAGAIN while it is queued or running, or the request's own error code
 * (message from SjhipLastError).  A request is forgotten once 0 or its
 * error has been returned. */
int SjhipServerWait(void* server, uint64_t request_id, int64_t timeout_ms, struct SjhipJoinResult* out,
                    struct SjhipRequestTiming* timing);

/* Finishes the queued and running requests, then frees the server;
 * results never waited for are released. */
int SjhipServerDestroy(void* server);

/* Copies a device join result to host arrays of at least result->size
 * entries. */
int SjhipJoinResultCopyToHost(const struct SjhipJoinResult* result, uint32_t* build_indices,
                              uint32_t* probe_indices);

/* Builds the index from the refiner's decoded build geometry, with rects
 * computed on the device under the operator's semantics (points: nearest
 * FP32, others: outward FP32).  Device-resident counterpart of the host
 * rect marshalling in rust/sedona-spatial-join-gpu/src/index/
 * gpu_spatial_index_builder.rs:193-259 followed by push_build /
 * finish_building. */
int SjhipIndexBuildFromRefiner(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner);

/* push_build for an Arrow array whose buffers are in ROCm device memory
 * (or CPU memory), decoded on the device; same contract as push_build
 * (gpuspatial_c.h:168-174). */
int SjhipRefinerPushBuildDevice(struct SedonaSpatialRefiner* refiner, const struct ArrowSchema* schema,
                                const struct ArrowDeviceArray* array);

/* Per-row FP32 rects of a WKB column computed on the device: the GPU form
 * of the operator's EvaluatedGeometryArray bounds
 * (rust/sedona-spatial-join/src/operand_evaluator.rs:182-217 ->
 * rust/sedona-geometry/src/bounds.rs:240-249 -> Bounds2D::new,
 * rust/sedona-spatial-join/src/utils/bounds.rs:49-63, 147-162), with the
 * GPU operator's point special case when point_nearest != 0
 * (rust/sedona-spatial-join-gpu/src/join_provider.rs:108-140) and the
 * distance expansion of operand_evaluator.rs:605-607 when
 * expand_distance >= 0 (pass NaN for none).  Empty and null rows get NaN
 * x 4, the encoding push_build and probe expect.  out_rects holds
 * 4 * length floats. */
int SjhipComputeRects(const struct ArrowSchema* schema, const struct ArrowArray* array, int point_nearest,
                      double expand_distance, float* out_rects);

/* Device-memory plumbing for callers without a device allocator of their
 * own (no counterpart in the reference): allocate / free / copy on the
 * current device, synchronous. */
int SjhipDeviceAlloc(uint64_t bytes, void** out_ptr);
int SjhipDeviceFree(void* ptr);
int SjhipCopyToDevice(void* dst, const void* src, uint64_t bytes);
int SjhipCopyToHost(void* dst, const void* src, uint64_t bytes);
int SjhipSetDevice(int device_id);

/* Bandwidth calibration (no counterpart in the reference): copies `bytes`
 * (a multiple of 16) from src to dst with a plain vector-copy kernel,
 * `reps` times on one stream, and returns the mean kernel time in *ms.
 * Its HBM traffic is exactly 2 * bytes per launch, which scales the
 * profiler's FETCH_SIZE / WRITE_SIZE readings of the other kernels. */
int SjhipCopyKernel(void* dst, const void* src, uint64_t bytes, int reps, double* ms);

/* Error text of the last failed extension call on this thread. */
const char* SjhipLastError(void);

#ifdef __cplusplus
}
#endif
#endif /* GPUSPATIAL_HIP_H */
