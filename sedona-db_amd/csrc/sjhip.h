// sjhip.h -- internal declarations of the MI355X spatial-join library.
//
// Device data layout (all in HBM, structure of arrays):
//   geometry column  : type[row] (u8), part_off[row+1], ring_off[part+1],
//                      pt_off[ring+1] (int64 prefix offsets), xy[pt] (double2)
//                      A column holding only points keeps xy[row] directly
//                      (NaN for null / empty rows) and no offsets.
//   ring refine index: ring_rect[ring] (double4, the reference's ring->rect),
//                      FP32 ring-local vertices xy32[pt] (float2) with a
//                      per-ring origin and error bound for the certified
//                      FP32 fast path, and per ring of 32+ segments a
//                      segment tree in natural ring order (boxes over 16
//                      consecutive segments, fan-out 8).
//   rect tree        : Morton-sorted FP32 leaf rects (float4) in groups of
//                      LEAF, internal levels of fan-out FANOUT, all levels
//                      in one float4 array; sorted build ids (u32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace sjhip {

// ---------------------------------------------------------------------------
// Errors

struct Error : std::runtime_error {
  int code;
  explicit Error(const std::string& m, int c = 22) : std::runtime_error(m), code(c) {}
};

#define SJ_CHECK(expr)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::sjhip::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +  \
                           __FILE__ + ":" + std::to_string(__LINE__));                    \
  } while (0)

// ---------------------------------------------------------------------------
// Device memory (stream-ordered)

// A non-blocking stream per device on which buffers that outlive the
// stream they were made on are freed and copied.  (The legacy null stream
// synchronises with every blocking stream, which HIP refuses while another
// thread captures a graph.)
inline hipStream_t detached_stream() {
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev < 0 || dev >= 64) dev = 0;
  std::lock_guard<std::mutex> g(mu);
  if (!streams[dev]) (void)hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking);
  return streams[dev];
}

// Bytes requested from the stream-ordered pool for `bytes`: 256-byte
// granules, and above 1 MiB the next of eight size classes per power of
// two (at most 12.5 % more).  Buffers whose sizes vary a little from one
// join to the next (candidates, pairs) then fall into blocks the pool
// already holds; a block of a new size makes the driver map (and clear)
// fresh memory while the host waits, about 1 s per 10 GB.
inline size_t pool_bytes(size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (bytes <= (size_t(1) << 20)) return bytes;
  size_t top = size_t(1) << (63 - __builtin_clzll((unsigned long long)bytes));
  const size_t step = top >> 3;
  return (bytes + step - 1) / step * step;
}

template <typename T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  hipStream_t s = nullptr;
  DevArray() = default;
  DevArray(size_t count, hipStream_t stream) { alloc(count, stream); }
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  DevArray(DevArray&& o) noexcept : p(o.p), n(o.n), s(o.s) { o.p = nullptr; o.n = 0; }
  DevArray& operator=(DevArray&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; s = o.s; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevArray() { release(); }
  void alloc(size_t count, hipStream_t stream) {
    release();
    s = stream;
    n = count;
    if (count) SJ_CHECK(hipMallocAsync((void**)&p, pool_bytes(count * sizeof(T)), stream));
  }
  // grows (without preserving contents) when smaller than count
  void reserve(size_t count, hipStream_t stream) {
    if (count > n || p == nullptr) alloc(count < 1 ? 1 : count, stream);
  }
  void release() {
    if (p) (void)hipFreeAsync(p, s);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------------------
// Geometry type codes (2-D WKB base types)

enum GeomType : uint8_t { kNull = 0, kPoint = 1, kLine = 2, kPolygon = 3, kMultiPoint = 4, kMultiLine = 5, kMultiPolygon = 6 };

// Device view of a geometry column.
struct GeomView {
  int64_t n_rows = 0;
  int points_only = 0;
  const uint8_t* type = nullptr;
  const int64_t* part_off = nullptr;
  const int64_t* ring_off = nullptr;
  const int64_t* pt_off = nullptr;
  const double2* xy = nullptr;
};

// Segment tree of one ring: level 0 holds one FP32 ring-local box per
// 2^seg_leaf_shift consecutive segments (an index-wide choice, 16 unless
// the autotuner picks 32), level l+1 one box per kSegFanout nodes of level
// l, up to a top level of at most kSegFanout nodes.
constexpr int kSegLeafShift = 4;
constexpr int kSegFanout = 8;
constexpr int kSegMaxLevels = 5;

// Record of a build row that is one polygon with one ring, or one line
// (npts >= 0): its ring's rect, first vertex, vertex count, FP32 bound,
// segment-tree offset (-1: none) and cell-raster offset (-1: none; a
// polygon's offset indexes RingIndex::raster, a line's
// RingIndex::line_raster).  npts = -1 marks every other row.
struct alignas(32) PolyDesc {
  double4 rect;
  int64_t v0;
  int64_t tree_off;
  int32_t npts;
  float err;
  int64_t raster_off;
};

// Cell raster of a polygon ring: kRasterCells x kRasterCells cells over
// the ring's rect, 2 bits each (0: near the boundary, 1: inside, 2:
// outside), kRasterWords 32-bit words per ring.
#ifndef SJ_RASTER_CELLS
#define SJ_RASTER_CELLS 64
#endif
constexpr int kRasterCells = SJ_RASTER_CELLS;
static_assert(kRasterCells == 32 || kRasterCells == 64 || kRasterCells == 128, "raster cells per side: 32, 64 or 128");
constexpr int kRasterWords = kRasterCells * kRasterCells / 16;

// Cell raster of a line: kLineCells x kLineCells cells over the line's
// rect, one bit per cell (set: some segment may pass through the cell),
// kLineRowWords 64-bit words per row of cells (cell x of row y: bit x % 64
// of word y * kLineRowWords + x / 64).
constexpr int kLineCells = 128;
constexpr int kLineRowWords = kLineCells / 64;
constexpr int kLineWords = kLineCells * kLineRowWords;

// Device view of the ring refine index (build side).
struct RingIndexView {
  const double4* ring_rect = nullptr;   // xmin, ymin, xmax, ymax (double)
  const float2* xy32 = nullptr;         // per point: ring-local FP32 vertex
  const double2* origin = nullptr;      // per ring: local origin (double)
  const float* err = nullptr;           // per ring: FP32 certification bound E
  const int64_t* seg_tree_off = nullptr;  // per ring: first box in seg_tree, -1 = not indexed
  const float4* seg_tree = nullptr;       // boxes (xmin, ymin, xmax, ymax), ring-local FP32
  const PolyDesc* desc = nullptr;         // per build row
  const uint32_t* raster = nullptr;       // cell rasters (see PolyDesc::raster_off)
  const uint64_t* line_raster = nullptr;  // line cell rasters (see PolyDesc::raster_off)
  int seg_leaf_shift = kSegLeafShift;     // segments per tree leaf: 1 << seg_leaf_shift
};

// Owned geometry column on the device.
struct GeomColumn {
  int64_t n_rows = 0, n_parts = 0, n_rings = 0, n_pts = 0;
  bool points_only = false;
  uint32_t type_mask = 0;  // bit t set when a row of type t is present
  // point columns decoded from WKB: bounds of the valid points rounded to
  // FP32 (min x, min y, max x, max y; the centres of their point rects)
  // and their number, or pt_valid < 0 when not computed
  float pt_bounds[4] = {0.f, 0.f, 0.f, 0.f};
  int64_t pt_valid = -1;
  DevArray<uint8_t> type;
  DevArray<int64_t> part_off, ring_off, pt_off;
  DevArray<double2> xy;
  GeomView view() const {
    GeomView v;
    v.n_rows = n_rows;
    v.points_only = points_only ? 1 : 0;
    v.type = type.p;
    v.part_off = part_off.p;
    v.ring_off = ring_off.p;
    v.pt_off = pt_off.p;
    v.xy = xy.p;
    return v;
  }
};

struct RingIndex {
  DevArray<double4> ring_rect;
  DevArray<float2> xy32;
  DevArray<double2> origin;
  DevArray<float> err;
  DevArray<int64_t> seg_tree_off;
  DevArray<float4> seg_tree;
  DevArray<PolyDesc> desc;
  DevArray<uint32_t> raster;
  DevArray<uint64_t> line_raster;
  int64_t n_boxes = 0;
  int64_t n_rasters = 0;
  int64_t n_line_rasters = 0;
  int seg_leaf_shift = kSegLeafShift;
  // For an index built on a stream that may end before it (after that
  // stream has been synchronised): its buffers are then freed on the
  // detached stream.
  void detach_streams() {
    hipStream_t d = detached_stream();
    ring_rect.s = xy32.s = origin.s = err.s = seg_tree_off.s = seg_tree.s = desc.s = raster.s = line_raster.s = d;
  }
  RingIndexView view() const {
    RingIndexView v;
    v.ring_rect = ring_rect.p;
    v.xy32 = xy32.p;
    v.origin = origin.p;
    v.err = err.p;
    v.seg_tree_off = seg_tree_off.p;
    v.seg_tree = seg_tree.p;
    v.desc = desc.p;
    v.raster = raster.p;
    v.line_raster = line_raster.p;
    v.seg_leaf_shift = seg_leaf_shift;
    return v;
  }
};

// Host description of an Arrow WKB column (host or device buffers).
struct WkbInput {
  int64_t length = 0;
  int64_t offset = 0;         // Arrow array offset (rows)
  const uint8_t* validity = nullptr;
  const void* offsets = nullptr;  // int32 or int64 (see large)
  bool large = false;
  const uint8_t* data = nullptr;
  // BinaryView: views (16 bytes each) and variadic data buffers
  bool view = false;
  const uint8_t* views = nullptr;
  std::vector<const uint8_t*> view_buffers;
  std::vector<int64_t> view_buffer_sizes;
  bool on_device = false;
};

// Parses an ArrowSchema format ("z", "Z", "vz") and an ArrowArray.
WkbInput wkb_input_from_arrow(const struct ArrowSchema* schema, const struct ArrowArray* array, bool on_device);

// Decodes a WKB column into a device geometry column (copying host buffers in).
void decode_wkb(const WkbInput& in, GeomColumn& out, hipStream_t s);

// Per-row FP32 rects: rounding outward; points round to nearest when
// point_nearest is set (the GPU operator's point special case).  Empty rows
// get NaN rects.  Optionally expands by a distance (f64) and rounds outward.
void compute_rects(const GeomColumn& g, const RingIndex* ri, bool point_nearest, double expand,
                   DevArray<float4>& rects, hipStream_t s);

// Ring refine index for a build column.
struct RingIndexParams {
  int min_segs_for_tree = 32;  // rings with fewer segments are scanned whole
  bool raster = true;          // cell rasters for single-ring polygon rows
  int seg_leaf_shift = kSegLeafShift;  // segments per tree leaf: 1 << seg_leaf_shift (4 or 5)
};
void build_ring_index(const GeomColumn& g, RingIndex& ri, const RingIndexParams& p, hipStream_t s);

// Rect tree over FP32 build rects.
struct RectTree {
  static constexpr int FANOUT = 8;
  static constexpr int LEAF = 8;  // default rects per leaf node
  int leaf = LEAF;                // rects per leaf node: 4, 8 or 16
  int64_t n_valid = 0;
  int n_levels = 0;                 // internal + leaf-node levels
  std::vector<int64_t> level_off;   // start of level l in nodes (l = 0: leaf nodes)
  std::vector<int64_t> level_cnt;
  DevArray<float4> nodes;           // all node boxes
  DevArray<float4> leaf_rects;      // Morton-sorted valid rects
  DevArray<uint32_t> leaf_ids;      // their build ids
  DevArray<uint32_t> leaf_keys;     // their Morton keys (sorted)
  float bb[4] = {0, 0, 0, 0};       // bounds of rect centres (Morton frame)
  // query unions up to this extent (max of width, height) are walked by
  // a whole wavefront at once: twice the median extent of a leaf node
  float coop_extent = 0.f;
  // observed candidates (and fused-join kept pairs) per query x 16: sizes
  // the next probe's output
  mutable std::atomic<uint64_t> cand_per_query_x16{0};
  mutable std::atomic<uint64_t> kept_per_query_x16{0};
  // Uniform cell grid over the valid rects, built with the tree for the
  // filter of point probes.  The cell of a coordinate is
  // clamp(floor((v - grid_x0) * grid_inv_x), 0, grid_nx - 1) in FP32 (and
  // likewise in y), the same monotonic map for rects and probes, so a rect
  // is listed in every cell that can hold a point of it.  Cell c's rects
  // are grid_rects / grid_ids[grid_off[c] .. grid_off[c + 1]).  grid_nx = 0:
  // no grid.
  int grid_nx = 0, grid_ny = 0;
  float grid_x0 = 0.f, grid_y0 = 0.f, grid_inv_x = 0.f, grid_inv_y = 0.f;
  // 0: the grid filter for point queries, the tree walk otherwise; 1: the
  // tree walk always; 2: the grid filter whenever there is a grid
  int filter_mode = 0;
  DevArray<uint32_t> grid_off;
  DevArray<float4> grid_rects;
  DevArray<uint32_t> grid_ids;
  RectTree() = default;
  RectTree(RectTree&& o) noexcept { *this = std::move(o); }
  // For a tree that outlives the stream it was built on (after that
  // stream has been synchronised): its buffers are then freed on the
  // detached stream instead of the destroyed one.
  void detach_streams() {
    nodes.s = leaf_rects.s = leaf_ids.s = leaf_keys.s = detached_stream();
    grid_off.s = grid_rects.s = grid_ids.s = detached_stream();
  }
  RectTree& operator=(RectTree&& o) noexcept {
    leaf = o.leaf;
    n_valid = o.n_valid;
    n_levels = o.n_levels;
    level_off = std::move(o.level_off);
    level_cnt = std::move(o.level_cnt);
    nodes = std::move(o.nodes);
    leaf_rects = std::move(o.leaf_rects);
    leaf_ids = std::move(o.leaf_ids);
    leaf_keys = std::move(o.leaf_keys);
    for (int k = 0; k < 4; k++) bb[k] = o.bb[k];
    coop_extent = o.coop_extent;
    filter_mode = o.filter_mode;
    grid_nx = o.grid_nx;
    grid_ny = o.grid_ny;
    grid_x0 = o.grid_x0;
    grid_y0 = o.grid_y0;
    grid_inv_x = o.grid_inv_x;
    grid_inv_y = o.grid_inv_y;
    grid_off = std::move(o.grid_off);
    grid_rects = std::move(o.grid_rects);
    grid_ids = std::move(o.grid_ids);
    cand_per_query_x16.store(o.cand_per_query_x16.load());
    kept_per_query_x16.store(o.kept_per_query_x16.load());
    return *this;
  }
};
void build_rect_tree(const float4* rects_dev, int64_t n, RectTree& t, hipStream_t s, int leaf = RectTree::LEAF);

// Optional HIP-event bracket around a kernel launch (on its stream).
struct KernelTimer {
  hipEvent_t start = nullptr, stop = nullptr;
  void record_start(hipStream_t s) { if (start) (void)hipEventRecord(start, s); }
  void record_stop(hipStream_t s) { if (stop) (void)hipEventRecord(stop, s); }
};

// Candidate buffers
struct PairBuf {
  DevArray<uint32_t> b, p;
  DevArray<unsigned long long> counter;
  // pairs both arrays can hold (their sizes can differ once one of them has
  // been replaced, e.g. by a sort's output array)
  uint64_t cap() const { return (uint64_t)std::min(b.n, p.n); }
};

// Sizes a for at least count elements: kept when large enough, else
// reallocated with a quarter of headroom (contents not preserved).
template <class T>
void fit(DevArray<T>& a, size_t count, hipStream_t s) {
  if (a.p && a.n >= count) return;
  a.alloc(count + count / 4 + 256, s);
}

// Scratch of the polygon-major split, kept in a workspace between joins.
struct PmScratch {
  DevArray<uint32_t> start, ends, cnt, rec_i, rec_s;
  DevArray<uint64_t> off, flag, cnt2, off2, whits, hits;
  DevArray<uint8_t> rec_n, tmp;
  DevArray<ulonglong2> masks;
  DevArray<float4> qrect;  // ST_DWithin: the probe rect of each point
  void set_stream(hipStream_t s) {
    start.s = ends.s = cnt.s = rec_i.s = rec_s.s = s;
    off.s = flag.s = cnt2.s = off2.s = whits.s = hits.s = s;
    rec_n.s = tmp.s = masks.s = qrect.s = s;
  }
};

// Device scratch kept between joins: the filter's candidate pairs, the
// refine's lists and marks, and its output before the exact-size copy
// into the join's result.  Reused (and grown with headroom), so that joins
// of similar size take no fresh blocks from the stream-ordered pool: a
// block of a new size makes the driver map and clear memory while the host
// waits.  A workspace serves one join at a time; set_stream names the
// stream its buffers are freed on (the join's, then the detached stream).
struct Workspace {
  PairBuf cand, out;
  PairBuf open;  // the candidates a raster pass leaves to the walk
  DevArray<uint32_t> redo;
  DevArray<uint32_t> rest;  // candidates of rows the single-ring walk leaves to the general kernel
  DevArray<uint64_t> mask;
  PmScratch pm;
  void set_stream(hipStream_t s) {
    cand.b.s = cand.p.s = cand.counter.s = out.b.s = out.p.s = out.counter.s = s;
    open.b.s = open.p.s = open.counter.s = redo.s = rest.s = mask.s = s;
    pm.set_stream(s);
  }
};

// Runs the filter of queries against the tree into pb; returns the count.
// qorder (optional) gives the visiting order of the queries; with
// emit_rank the pairs carry the query's position i in qorder instead of
// its row qorder[i].
// point_probes: every query rect is a point (the grid filter applies);
// ordered_points (with emit_rank): the query points already gathered in
// qorder's order, which the grid filter then reads in sequence instead of
// the rects through qorder.
uint64_t filter_pairs(const RectTree& t, const float4* q, int64_t nq, const uint32_t* qorder, PairBuf& pb,
                      hipStream_t s, KernelTimer* timer = nullptr, bool emit_rank = false, bool point_probes = false,
                      const double2* ordered_points = nullptr);

// Gathers the points of a point column in the given row order
// (out.xy[i] = probe.xy[order[i]]), producing a point column of n rows
// without a type array (its rows are valid points or NaN).
void gather_points(const GeomColumn& probe, const uint32_t* order, int64_t n, GeomColumn& out, hipStream_t s);

// Morton frame of a set of query rects: the key of a rect with centre
// (cx, cy) interleaves qx = clamp((cx - x0) * sx, 0, 65535) (x in the even
// bits) and qy likewise, in FP32.
struct MortonFrame {
  float x0 = 0.f, y0 = 0.f, sx = 0.f, sy = 0.f;
};

// Morton order of query rects (a permutation of valid queries; invalid ones
// are dropped).  Returns the number of valid queries.  Optionally returns
// the frame and the sorted keys (the first n_valid of them are the valid
// queries').
int64_t morton_order(const float4* rects, int64_t n, DevArray<uint32_t>& order, hipStream_t s,
                     MortonFrame* frame = nullptr, DevArray<uint32_t>* sorted_keys = nullptr);

// Morton order of the valid rows of a point column by the keys its point
// rects (rounded to nearest) would get from morton_order: the same order,
// frame and sorted keys, without the rects.  Needs the bounds decode_wkb
// records for point columns.
int64_t morton_order_points(const GeomColumn& points, DevArray<uint32_t>& order, hipStream_t s, MortonFrame* frame,
                            DevArray<uint32_t>* sorted_keys);

struct PmSplitResult {
  uint64_t candidates = 0;  // (build, probe) rect hits
  uint64_t kept = 0;        // decided inside by a cell raster
  uint64_t open = 0;        // left to the walk
};

// Polygon-major filter of Morton-ordered point probes: xy[0, nq) are the
// valid probe points in Morton order (keys[0, nq) their sorted keys in
// frame f), probe_ids maps a position to its probe row.  Every build rect
// of t tests the points of the Morton cells it covers; the candidates come
// out grouped by build rect (rect-tree order), in Morton order within a
// cell.  With raster (certified point-in-polygon refine, cell rasters
// present) the candidates in clean raster cells are decided: kept pairs
// (build, probe row) go to out (out.counter = their number) and the rest
// to open as (build, position); without, every candidate goes to open.
// With dist >= 0 (ST_DWithin, raster false) each point's probe rect is its
// outward box widened by dist; dist_raster 1 (single-line build rows with
// line rasters) drops the hits the line rasters rule out, 2 (polygon build
// rows with cell rasters) keeps the hits in clean inside cells (distance
// 0) in out like the point-in-polygon split, 0 uses no raster; the rest go
// to open.  Returns false, having written nothing the caller reads, where it
// does not apply (no cell grid, a rect over too many cells); the caller
// then takes the probe-major filter.
bool pm_split(const RectTree& t, const RingIndex& ri, const double2* xy, const uint32_t* keys, int64_t nq,
              const MortonFrame& f, const uint32_t* probe_ids, bool raster, PairBuf& out, PairBuf& open,
              PmScratch& sc, PmSplitResult* res, hipStream_t s, double dist = -1.0, int dist_raster = 0);

// Refine statistics
struct RefineStats {
  uint64_t segment_tests = 0;
  uint64_t fallback_tests = 0;
  uint64_t walked = 0;  // candidates the refine's walk decided (after cell-raster passes)
};

// Whether refine_relation takes its point-in-polygon kernel (polygon
// build rows, point probes, Intersects / Covers / Contains).
bool pip_refine_applies(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, int predicate);

// Relation refine of candidate pairs; compacts kept pairs into out (count
// returned).  precision: 0 = FP32 certified + FP64 fallback, 1 = all-FP64.
// probe_ids (point-in-polygon kernel only): cp holds positions in probe,
// and the kept pairs carry probe_ids[position].
uint64_t refine_relation(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, int predicate,
                         const uint32_t* cb, const uint32_t* cp, uint64_t n, int precision, PairBuf& out,
                         RefineStats* stats, hipStream_t s, KernelTimer* timer = nullptr,
                         const uint32_t* probe_ids = nullptr, int pip_waves = 8, Workspace* ws = nullptr,
                         bool raster_done = false);

// Cell-raster split of point-in-polygon candidates, before they are
// regrouped: of the n candidates (cb, cp), those whose point lies in a
// clean raster cell are decided (kept pairs written to out, probe
// positions mapped through probe_ids; out.counter set to their number) and
// the others copied, in candidate order, to open (mask: scratch for the
// marks).  Returns false, doing nothing, where it does not apply (no
// rasters, not the certified point-in-polygon refine); else *n_open = the
// open candidates, which refine_relation(..., raster_done = true) then
// refines, appending to out.
bool raster_split(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, int predicate,
                  int precision, const uint32_t* cb, const uint32_t* cp, uint64_t n, const uint32_t* probe_ids,
                  PairBuf& out, PairBuf& open, DevArray<uint64_t>& mask, uint64_t* n_open, hipStream_t s);

// The same for ST_DWithin over single-line build rows and point probes
// (the line rasters only drop candidates): the candidates they cannot rule
// out are copied, in candidate order, to open, for
// refine_distance(..., raster_done = true).  Returns false, doing nothing,
// where the line rasters do not apply.
// Whether the line rasters of the distance refine apply: build rows single
// lines with rasters, point probes, a finite distance >= 0, not the
// all-FP64 form.
bool line_rasters_apply(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                        int precision);
// The rasters the ST_DWithin build-rect pass uses for this build side: 1
// line rasters (line_rasters_apply), 2 polygon cell rasters (build rows
// (multi)polygons, the certified refine), 0 none.
int dist_raster_kind(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d, int precision);
bool raster_split_distance(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                           int precision, const uint32_t* cb, const uint32_t* cp, uint64_t n, PairBuf& open,
                           DevArray<uint64_t>& mask, uint64_t* n_open, hipStream_t s);

// Distance refine (build.distance(probe) <= d); probe must be points.
// precision 1: the all-FP64 form over every segment.  stats: FP64 segment
// distance tests, and exact ring scans (uncertain certified crossing counts).
// FP64 reference distance (build row cb[i], probe point cp[i]) into out[i].
void pair_distances(const GeomColumn& build, const GeomColumn& probe, const uint32_t* cb, const uint32_t* cp,
                    uint64_t n, double* out, hipStream_t s);
// probe_ids (optional): cp holds positions in probe, and the kept pairs
// carry probe_ids[position].  out_holds_kept: out already holds kept pairs
// (out.counter of them), which the refine appends to.
uint64_t refine_distance(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                         const uint32_t* cb, const uint32_t* cp, uint64_t n, PairBuf& out, hipStream_t s,
                         KernelTimer* timer = nullptr, int precision = 0, RefineStats* stats = nullptr,
                         Workspace* ws = nullptr, bool raster_done = false, const uint32_t* probe_ids = nullptr,
                         bool out_holds_kept = false);

// ST_DWithin join of point probes in one kernel: the filter walk decides
// the distance predicate for each rect hit, so only kept pairs are
// written.  q: probe rects expanded by d; qorder: Morton visiting order.
// *candidates receives the number of rect hits.
uint64_t filter_refine_distance(const RectTree& t, const float4* q, int64_t nq, const uint32_t* qorder,
                                const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                                PairBuf& out, uint64_t* candidates, RefineStats* stats, hipStream_t s,
                                KernelTimer* timer = nullptr);

// Regroups candidate pairs by build row with a stable radix sort: within a
// row they keep their order (the probes' Morton order), so that a
// wavefront refines nearby points of one polygon.
void sort_pairs_by_build(PairBuf& pb, uint64_t n, int64_t n_build, hipStream_t s);

// Sorts pairs by (probe, build) in place.
void sort_pairs(uint32_t* b, uint32_t* p, uint64_t n, hipStream_t s);

// KNN over point columns.  The object tree over the build points is built
// on first use and kept in `index` (resident with the build side).
struct KnnIndex {
  RectTree tree;
  bool built = false;
};
// k nearest objects per query into fixed rows of k (id 0xffffffff and
// distance +inf where fewer exist).  One side must be points.
void knn_points(const GeomColumn& build, KnnIndex& index, const GeomColumn& probe, int k, uint32_t* out_ids_dev,
                double* out_d_dev, hipStream_t s, KernelTimer* timer = nullptr);
// The same with variable-length rows: query q's neighbours are
// [offsets[q], offsets[q+1]) of out_ids/out_d, sorted by (distance, id);
// with tol >= 0 every object within tol of the k-th distance is kept too
// (tie breakers), with tol < 0 at most k.  Returns the total.
uint64_t knn_with_ties(const GeomColumn& build, KnnIndex& index, const GeomColumn& probe, int k, double tol,
                       DevArray<unsigned long long>& offsets, DevArray<uint32_t>& out_ids, DevArray<double>& out_d,
                       hipStream_t s);

// ---------------------------------------------------------------------------
// Launch sequence without host synchronisation or allocation, for capture
// in a HIP graph and replay over new inputs (the request loop): inputs are
// described in device memory, counts are read on the device, buffers have
// fixed capacities.

// A column of 21-byte WKB points in device memory: data, row offsets
// (int32, or int64 when large), row count.
struct PointsDesc {
  const uint8_t* data;
  const void* offsets;
  int64_t n;
  int32_t large;
  int32_t pad;
};

// Writes a descriptor to device memory with a kernel (no copy from host
// memory, no allocation).
void store_points_desc(const PointsDesc& d, PointsDesc* dst, hipStream_t s);

// Decodes rows [0, d->n) of the described column into xy/type, rows up to
// cap as empty; rows that are not point records set *err.
void decode_points_desc(const PointsDesc* desc_dev, int64_t cap, double2* xy, uint8_t* type, unsigned long long* err,
                        hipStream_t s);
void rects_of_points_into(const double2* xy, const uint8_t* type, int64_t n, bool point_nearest, double expand,
                          float4* out, hipStream_t s);
void gather_points_into(const double2* xy, const uint32_t* order, int64_t n, double2* out, hipStream_t s);

// Stable LSD radix sort of (u32 key, u32 value) pairs over bits
// [begin_bit, end_bit), n <= 2^24, made of kernels only (no memset,
// allocation or host synchronisation): a captured graph may replay it.
// `temp` holds graph_sort_temp_bytes(n) bytes; inputs are not modified.
size_t graph_sort_temp_bytes(int64_t n);
void graph_sort_pairs(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout, int64_t n,
                      int begin_bit, int end_bit, void* temp, size_t temp_bytes, hipStream_t s);

// Morton order of n query rects in the tree's frame (invalid rects last):
// keys/ids and their sorted copies, with `temp` of
// morton_sort_temp_bytes(n).  Kernels only (graph_sort_pairs).
size_t morton_sort_temp_bytes(int64_t n);
void morton_order_fixed(const RectTree& t, const float4* rects, int64_t n, uint32_t* keys, uint32_t* ids,
                        uint32_t* keys_sorted, uint32_t* order, void* temp, size_t temp_bytes, hipStream_t s);

// One filter launch into fixed buffers (pairs beyond cap are counted, not
// written).
void launch_filter(const RectTree& t, const float4* q, int64_t nq, const uint32_t* qorder, bool emit_rank,
                   uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter, hipStream_t s,
                   bool point_probes = false);

// Fills n 64-bit words with v (a kernel: captured sequences record no
// memset nodes).
void fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);

// Certified FP32 point-in-polygon walk of candidates [0, n) (count
// min(n, *n_dev) when n_dev is given) over single-ring build rows, one ray
// direction per wavefront.  Kept pairs go to (ob, op) (probe ids mapped
// through probe_ids when given); candidates with an uncertain test are
// listed in redo (count *n_redo) for the FP64 pass; candidates of other
// rows in rest (count *n_rest), or in redo when rest is nullptr.
// raster_checked: a raster pass has already decided the candidates in
// clean cells.  stats[0..1]: tests and uncertain tests.  queue: a
// counter that is zero at launch, for wavefronts that take 64 candidates
// at a time (nullptr: a grid-stride loop).
void launch_ring_walk(const RingIndex& ri, const GeomColumn& probe, const uint32_t* cb, const uint32_t* cp,
                      uint64_t n, const unsigned long long* n_dev, const uint32_t* probe_ids, bool raster_checked,
                      uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter,
                      unsigned long long* stats, uint32_t* redo, unsigned long long* n_redo, uint32_t* rest,
                      unsigned long long* n_rest, int waves, hipStream_t s, unsigned long long* queue = nullptr);

// The certified FP32 ST_DWithin walk of candidates whose build row is one
// line: pairs certainly within d are appended to (ob, op); candidates of
// other rows, and those with an uncertain test and no certain one, are
// listed in redo (count *n_redo) for the FP64 kernel.  stats[0..1]: tests
// and uncertain tests.  queue: a counter that is zero at launch.  With a
// pair list (pb, pp, n_p), the build rows are single-ring polygons: the
// candidates whose ring segments are all certainly farther than d are
// listed there as (build row, probe position) for the point-in-polygon
// walk.
void launch_line_walk(const RingIndex& ri, const GeomColumn& probe, const uint32_t* cb, const uint32_t* cp,
                      uint64_t n, const unsigned long long* n_dev, const uint32_t* probe_ids, double d, uint32_t* ob,
                      uint32_t* op, uint64_t cap, unsigned long long* counter, unsigned long long* stats,
                      uint32_t* redo, unsigned long long* n_redo, hipStream_t s, unsigned long long* queue,
                      uint32_t* pb = nullptr, uint32_t* pp = nullptr, unsigned long long* n_p = nullptr);

// Throws "build index out of range: v" / "probe index out of range: v" for
// the first position (build index first) with cb[i] >= nb or cp[i] >= np.
void check_pair_ranges(const uint32_t* cb, const uint32_t* cp, uint64_t n, uint64_t nb, uint64_t np,
                       hipStream_t s);

// Refine launches over min(*n_dev, n_cap) candidates.  Point in polygon:
// certified FP32 pass (uncertain candidates to redo/*redo_n) and the FP64
// pass over them; stats[0..1] segment tests and uncertain tests.
void launch_pip_refine_dev(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, int predicate,
                           const uint32_t* cb, const uint32_t* cp, const unsigned long long* n_dev, uint64_t n_cap,
                           const uint32_t* probe_ids, uint32_t* redo, unsigned long long* redo_n, uint32_t* ob,
                           uint32_t* op, uint64_t out_cap, unsigned long long* out_n, unsigned long long* stats,
                           int pip_waves, hipStream_t s);
void launch_distance_refine_dev(const GeomColumn& build, const RingIndex& ri, const GeomColumn& probe, double d,
                                const uint32_t* cb, const uint32_t* cp, const unsigned long long* n_dev,
                                uint64_t n_cap, uint32_t* ob, uint32_t* op, uint64_t out_cap,
                                unsigned long long* out_n, unsigned long long* stats, hipStream_t s);

// Checks that every pair's geometry types are supported by the relation
// refine; throws otherwise.
void check_relation_types(uint32_t build_mask, uint32_t probe_mask, int predicate);

}  // namespace sjhip
