// knn.hip -- exact k-nearest-neighbour join.
//
// Semantics of the reference (rust/sedona-spatial-join/src/index/
// default_spatial_index.rs:306-502 with the ground truth of
// rust/sedona-spatial-join/tests/spatial_join_integration.rs:1373-1414):
// distance = geo's Euclidean distance in FP64 exactly as the reference's
// host computes it (hypot between points; point to (multi)linestring or
// (multi)polygon through distance_to_point), neighbours ordered by
// (distance, build id), first k kept.  With include_tie_breakers
// (default_spatial_index.rs:379-472) every object whose distance is within
// DISTANCE_TOLERANCE = 1e-9 of the k-th distance is kept as well
// (index/spatial_index.rs:32).  One side of each pair is a point: point
// queries against any objects, or any queries against point objects.
//
// Device plan: the objects sit in the Morton-ordered rect tree (FP32 rects
// rounded outward, so they contain the FP64 geometry).  Each lane owns one
// query and a bounded max-heap of K (distance, id) entries kept sorted in
// registers.  Phase 1 seeds the heap with the objects next to the query's
// Morton position (a binary search over the sorted leaf keys); phase 2
// walks the tree depth-first with its stack in LDS and prunes every node
// whose FP64 lower-bound distance exceeds the current k-th distance, so
// ties at the k-th distance are always examined.  Ties: a counting walk
// bounded by the k-th distance + 1e-9, a scan, and a filling walk whose
// per-query runs are then sorted by (distance, id).
#include <hipcub/hipcub.hpp>

#include "geom_math.h"
#include "row_geom.h"
#include "sjhip.h"

namespace sjhip {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxLevels = 14;

__device__ __forceinline__ uint32_t spread16(uint32_t x) {
  x &= 0xffff;
  x = (x | (x << 8)) & 0x00ff00ff;
  x = (x | (x << 4)) & 0x0f0f0f0f;
  x = (x | (x << 2)) & 0x33333333;
  x = (x | (x << 1)) & 0x55555555;
  return x;
}

__device__ __forceinline__ float down32(double v) {
  float f = (float)v;
  if ((double)f > v) f = nextafterf(f, -INFINITY);
  return f;
}
__device__ __forceinline__ float up32(double v) {
  float f = (float)v;
  if ((double)f < v) f = nextafterf(f, INFINITY);
  return f;
}

__global__ void k_point_rects(const double2* xy, const uint8_t* type, int64_t n, float4* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 p = xy[i];
  if (type[i] == 0 || p.x != p.x || p.y != p.y) {
    out[i] = make_float4(NAN, NAN, NAN, NAN);
    return;
  }
  out[i] = make_float4(down32(p.x), down32(p.y), up32(p.x), up32(p.y));
}

struct KnnTree {
  const float4* nodes;
  const float4* leaf_rects;
  const uint32_t* leaf_ids;
  const double2* obj;  // object points by build id (a points-only build column)
  GeomView objs;       // the object rows (any geometry)
  int64_t n_valid;
  int n_levels;
  int64_t off[kMaxLevels];
  int64_t cnt[kMaxLevels];
};

// Lower bound (FP64, rounded down) of the distance from q to any point of
// the FP32 box b.
__device__ __forceinline__ double box_lb(const float4& b, double qx, double qy) {
  double dx = fmax(fmax((double)b.x - qx, qx - (double)b.z), 0.0);
  double dy = fmax(fmax((double)b.y - qy, qy - (double)b.w), 0.0);
  double d = sqrt(dx * dx + dy * dy);
  return d * (1.0 - 1e-12);
}

template <int K>
struct Heap {
  double d[K];
  uint32_t id[K];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int i = 0; i < K; i++) {
      d[i] = INFINITY;
      id[i] = 0xffffffffu;
    }
  }
  __device__ __forceinline__ static bool less(double da, uint32_t ia, double db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
  }
  // sorted ascending; insert when better than the last entry
  __device__ __forceinline__ void push(double dd, uint32_t ii) {
    if (!less(dd, ii, d[K - 1], id[K - 1])) return;
    // skip duplicates (an id can be met twice: seeding and traversal)
#pragma unroll
    for (int i = 0; i < K; i++)
      if (id[i] == ii) return;
    double cd = dd;
    uint32_t ci = ii;
#pragma unroll
    for (int i = 0; i < K; i++) {
      if (less(cd, ci, d[i], id[i])) {
        double td = d[i];
        uint32_t ti = id[i];
        d[i] = cd;
        id[i] = ci;
        cd = td;
        ci = ti;
      }
    }
  }
  __device__ __forceinline__ double bound() const { return d[K - 1]; }
};

// Lower bound of the distance between the FP64 box q (xmin, ymin, xmax,
// ymax) and the FP32 box b, rounded down.
__device__ __forceinline__ double boxbox_lb(const float4& b, const double4& q) {
  double dx = fmax(fmax((double)b.x - q.z, q.x - (double)b.z), 0.0);
  double dy = fmax(fmax((double)b.y - q.w, q.y - (double)b.w), 0.0);
  return sqrt(dx * dx + dy * dy) * (1.0 - 1e-12);
}

// The query of one lane and its distances.  Mode 0: point queries, point
// objects (hypot); 1: point queries, objects of any geometry; 2: queries
// of any geometry, point objects.  Modes 1 and 2 evaluate geo's segment
// distance, whose rounding is absolute in the coordinates: their lower
// bounds also give up 2^-46 of the largest coordinate magnitude.
template <int MODE>
struct KnnQuery {
  double2 p;   // the query point (modes 0, 1); the query rect's centre (mode 2)
  double4 r;   // the query rect (mode 2)
  int64_t row; // the query row
  __device__ __forceinline__ double lb(const float4& b) const {
    if (MODE == 0) return box_lb(b, p.x, p.y);
    double scale = fmax(fmax(fabs((double)b.x), fabs((double)b.z)), fmax(fabs((double)b.y), fabs((double)b.w)));
    double l;
    if (MODE == 1) {
      scale = fmax(scale, fmax(fabs(p.x), fabs(p.y)));
      l = box_lb(b, p.x, p.y);
    } else {
      scale = fmax(scale, fmax(fmax(fabs(r.x), fabs(r.z)), fmax(fabs(r.y), fabs(r.w))));
      l = boxbox_lb(b, r);
    }
    return l - scale * 1.4210854715202004e-14;
  }
  __device__ __forceinline__ double dist(const KnnTree& t, const GeomView& q, uint32_t id) const {
    if (MODE == 0) {
      double2 o = t.obj[id];
      return hypot_ref(p.x - o.x, p.y - o.y);
    }
    if (MODE == 1) return distance_to_point(t.objs, id, p.x, p.y);
    double2 o = t.obj[id];
    return distance_to_point(q, row, o.x, o.y);
  }
};

// Sets up lane qi's query; false when the query row is null or empty.
template <int MODE>
__device__ __forceinline__ bool knn_query(const GeomView& q, const float4* qrects, int64_t qi, KnnQuery<MODE>& Q) {
  Q.row = qi;
  if (MODE == 2) {
    float4 b = qrects[qi];
    if (!(b.x <= b.z && b.y <= b.w)) return false;
    Q.r = make_double4(b.x, b.y, b.z, b.w);
    Q.p = make_double2(0.5 * ((double)b.x + (double)b.z), 0.5 * ((double)b.y + (double)b.w));
    return true;
  }
  double2 qp = q.xy[qi];
  Q.p = qp;
  return q.type[qi] != 0 && qp.x == qp.x && qp.y == qp.y;
}

// Depth-first walk over the tree from the top, with the per-level node and
// pending-children mask of lane tid in LDS: nodes whose lower bound exceeds
// bound() are pruned (re-checked on the way down as the bound shrinks),
// leaf objects go to leaf(id) when their rect passes.
template <int F, int LEAF, class LB, class Bound, class Leaf>
__device__ __forceinline__ void knn_walk(const KnnTree& t, uint32_t (*s_node)[kBlock], uint32_t (*s_mask)[kBlock],
                                         int tid, LB lb, Bound bound, Leaf leaf) {
  const int top = t.n_levels;
  int lvl = top;
  {
    uint32_t m = 0;
    const float4* nd = t.nodes + t.off[top - 1];
    int64_t nc = t.cnt[top - 1];
    for (int c = 0; c < F; c++)
      if (c < nc && !(lb(nd[c]) > bound())) m |= 1u << c;
    s_node[top][tid] = 0;
    s_mask[top][tid] = m;
  }
  while (true) {
    while (lvl <= top && s_mask[lvl][tid] == 0) lvl++;
    if (lvl > top) break;
    uint32_t m = s_mask[lvl][tid];
    int c = __ffs(m) - 1;
    s_mask[lvl][tid] = m & (m - 1);
    int64_t child = (int64_t)s_node[lvl][tid] * F + c;
    if (lvl - 1 == 0) {
      int64_t base = child * LEAF;
      int64_t e = min(base + LEAF, t.n_valid);
      for (int64_t j = base; j < e; j++) {
        float4 r = t.leaf_rects[j];
        if (lb(r) > bound()) continue;
        leaf(t.leaf_ids[j]);
      }
    } else {
      const float4* own = t.nodes + t.off[lvl - 1];
      if (lb(own[child]) > bound()) continue;
      int cl = lvl - 2;
      const float4* nd = t.nodes + t.off[cl];
      int64_t c0 = child * F;
      int64_t nc = t.cnt[cl] - c0;
      uint32_t mm = 0;
#pragma unroll
      for (int kk = 0; kk < F; kk++)
        if (kk < nc && !(lb(nd[c0 + kk]) > bound())) mm |= 1u << kk;
      lvl = lvl - 1;
      s_node[lvl][tid] = (uint32_t)child;
      s_mask[lvl][tid] = mm;
    }
  }
}

template <int K, int F, int LEAF, int MODE>
__global__ __launch_bounds__(kBlock) void k_knn(KnnTree t, const uint32_t* leaf_keys, float x0, float y0, float sx,
                                                float sy, GeomView q, const float4* qrects, int64_t nq,
                                                const uint32_t* qorder, int k, uint32_t* out_ids, double* out_d) {
  __shared__ uint32_t s_node[kMaxLevels + 1][kBlock];
  __shared__ uint32_t s_mask[kMaxLevels + 1][kBlock];
  const int tid = threadIdx.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  if (i >= nq) return;
  uint32_t qi = qorder ? qorder[i] : (uint32_t)i;
  Heap<K> h;
  h.init();
  KnnQuery<MODE> Q;
  bool valid = knn_query<MODE>(q, qrects, qi, Q) && t.n_levels > 0;
  if (valid) {
    // phase 1: neighbours in Morton order
    float fx = ((float)Q.p.x - x0) * sx, fy = ((float)Q.p.y - y0) * sy;
    uint32_t qx = (uint32_t)fminf(fmaxf(fx, 0.f), 65535.f);
    uint32_t qy = (uint32_t)fminf(fmaxf(fy, 0.f), 65535.f);
    uint32_t key = spread16(qx) | (spread16(qy) << 1);
    int64_t lo = 0, hi = t.n_valid;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (leaf_keys[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    int64_t a = max((int64_t)0, lo - K), b = min(t.n_valid, lo + K);
    for (int64_t j = a; j < b; j++) {
      uint32_t id = t.leaf_ids[j];
      h.push(Q.dist(t, q, id), id);
    }
    // phase 2: pruned depth-first walk
    knn_walk<F, LEAF>(
        t, s_node, s_mask, tid, [&](const float4& b) { return Q.lb(b); }, [&]() { return h.bound(); },
        [&](uint32_t id) { h.push(Q.dist(t, q, id), id); });
  }
  for (int j = 0; j < k; j++) {
    out_ids[(int64_t)qi * k + j] = h.id[j];
    out_d[(int64_t)qi * k + j] = h.d[j];
  }
}

// Tie breakers, pass 1: per query the number of objects whose distance is
// at most the k-th distance + tol (the k nearest included); a query with
// fewer than k objects keeps those it found, and so does every query when
// tol < 0 (no tie breakers).
template <int F, int LEAF, int MODE>
__global__ __launch_bounds__(kBlock) void k_knn_tie_count(KnnTree t, GeomView q, const float4* qrects, int64_t nq,
                                                          int k, const double* heap_d, double tol,
                                                          unsigned long long* counts) {
  __shared__ uint32_t s_node[kMaxLevels + 1][kBlock];
  __shared__ uint32_t s_mask[kMaxLevels + 1][kBlock];
  const int tid = threadIdx.x;
  int64_t qi = (int64_t)blockIdx.x * blockDim.x + tid;
  if (qi >= nq) return;
  const double dk = heap_d[qi * k + k - 1];
  unsigned long long cnt = 0;
  if (!(dk < INFINITY) || tol < 0) {
    for (int j = 0; j < k; j++) cnt += heap_d[qi * k + j] < INFINITY ? 1 : 0;
  } else {
    KnnQuery<MODE> Q;
    if (knn_query<MODE>(q, qrects, qi, Q) && t.n_levels > 0) {
      const double thr = dk + tol;
      knn_walk<F, LEAF>(
          t, s_node, s_mask, tid, [&](const float4& b) { return Q.lb(b); }, [&]() { return thr; },
          [&](uint32_t id) { cnt += Q.dist(t, q, id) <= thr ? 1 : 0; });
    }
  }
  counts[qi] = cnt;
}

// Tie breakers, pass 2: writes each query's objects within the bound of
// pass 1 at its offset, then sorts the run by (distance, id).
template <int F, int LEAF, int MODE>
__global__ __launch_bounds__(kBlock) void k_knn_tie_fill(KnnTree t, GeomView q, const float4* qrects, int64_t nq,
                                                         int k, const uint32_t* heap_ids, const double* heap_d,
                                                         double tol, const unsigned long long* offsets,
                                                         uint32_t* out_ids, double* out_d) {
  __shared__ uint32_t s_node[kMaxLevels + 1][kBlock];
  __shared__ uint32_t s_mask[kMaxLevels + 1][kBlock];
  const int tid = threadIdx.x;
  int64_t qi = (int64_t)blockIdx.x * blockDim.x + tid;
  if (qi >= nq) return;
  const uint64_t o0 = offsets[qi], o1 = offsets[qi + 1];
  const double dk = heap_d[qi * k + k - 1];
  if (!(dk < INFINITY) || tol < 0 || o1 - o0 == (uint64_t)k) {
    // the heap holds the answer (fewer than k objects, or no ties beyond k)
    for (uint64_t j = 0; j < o1 - o0; j++) {
      out_ids[o0 + j] = heap_ids[qi * k + j];
      out_d[o0 + j] = heap_d[qi * k + j];
    }
    return;
  }
  KnnQuery<MODE> Q;
  if (!(knn_query<MODE>(q, qrects, qi, Q) && t.n_levels > 0)) return;
  const double thr = dk + tol;
  uint64_t w = o0;
  knn_walk<F, LEAF>(
      t, s_node, s_mask, tid, [&](const float4& b) { return Q.lb(b); }, [&]() { return thr; },
      [&](uint32_t id) {
        double d = Q.dist(t, q, id);
        if (d <= thr && w < o1) {
          out_ids[w] = id;
          out_d[w] = d;
          w++;
        }
      });
  // insertion sort of the run by (distance, id)
  for (uint64_t a = o0 + 1; a < w; a++) {
    double dd = out_d[a];
    uint32_t ii = out_ids[a];
    uint64_t b = a;
    while (b > o0 && (out_d[b - 1] > dd || (out_d[b - 1] == dd && out_ids[b - 1] > ii))) {
      out_d[b] = out_d[b - 1];
      out_ids[b] = out_ids[b - 1];
      b--;
    }
    out_d[b] = dd;
    out_ids[b] = ii;
  }
}

unsigned int blocks_for(int64_t n) { return (unsigned int)((n + kBlock - 1) / kBlock); }

template <int K, int MODE>
void launch_knn(const KnnTree& kt, const uint32_t* keys, const float bb[4], float sx, float sy, const GeomColumn& probe,
                const float4* qrects, const uint32_t* order, int64_t nq, int k, uint32_t* ids, double* d,
                hipStream_t s) {
  hipLaunchKernelGGL((k_knn<K, RectTree::FANOUT, RectTree::LEAF, MODE>), blocks_for(nq), kBlock, 0, s, kt, keys,
                     bb[0], bb[1], sx, sy, probe.view(), qrects, nq, order, k, ids, d);
}

template <int MODE>
void launch_knn_k(const KnnTree& kt, const uint32_t* keys, const float bb[4], float sx, float sy,
                  const GeomColumn& probe, const float4* qrects, const uint32_t* ord, int64_t nq, int k, uint32_t* ids,
                  double* d, hipStream_t s) {
  if (k <= 1) launch_knn<1, MODE>(kt, keys, bb, sx, sy, probe, qrects, ord, nq, k, ids, d, s);
  else if (k <= 2) launch_knn<2, MODE>(kt, keys, bb, sx, sy, probe, qrects, ord, nq, k, ids, d, s);
  else if (k <= 4) launch_knn<4, MODE>(kt, keys, bb, sx, sy, probe, qrects, ord, nq, k, ids, d, s);
  else if (k <= 8) launch_knn<8, MODE>(kt, keys, bb, sx, sy, probe, qrects, ord, nq, k, ids, d, s);
  else if (k <= 16) launch_knn<16, MODE>(kt, keys, bb, sx, sy, probe, qrects, ord, nq, k, ids, d, s);
  else launch_knn<32, MODE>(kt, keys, bb, sx, sy, probe, qrects, ord, nq, k, ids, d, s);
}

// The distance mode of a (build, probe) pair of columns (see KnnQuery).
int knn_mode(const GeomColumn& build, const GeomColumn& probe) {
  const uint32_t pt = 1u << kPoint;
  const bool bpts = build.points_only || (build.type_mask & ~pt) == 0;
  const bool qpts = probe.points_only || (probe.type_mask & ~pt) == 0;
  if (bpts && qpts) return build.points_only && probe.points_only ? 0 : -1;
  if (qpts) return probe.points_only ? 1 : -1;
  if (bpts) return build.points_only ? 2 : -1;
  return -1;
}

struct KnnSetup {
  KnnTree kt{};
  float sx = 0, sy = 0;
  const float* bb = nullptr;
  DevArray<float4> qrects;
  DevArray<uint32_t> order;
  const uint32_t* ord = nullptr;
  int mode = 0;
};

void knn_setup(const GeomColumn& build, KnnIndex& index, const GeomColumn& probe, KnnSetup& S, hipStream_t s) {
  S.mode = knn_mode(build, probe);
  if (S.mode < 0)
    throw Error("knn: one side must be points (point queries against any objects, or any queries against point "
                "objects)");
  const int64_t nb = build.n_rows, nq = probe.n_rows;
  if (!index.built) {
    DevArray<float4> rects(nb > 0 ? nb : 1, s);
    if (build.points_only) {
      if (nb > 0)
        hipLaunchKernelGGL(k_point_rects, blocks_for(nb), kBlock, 0, s, build.xy.p, build.type.p, nb, rects.p);
    } else {
      compute_rects(build, nullptr, false, NAN, rects, s);
    }
    build_rect_tree(rects.p, nb, index.tree, s);
    SJ_CHECK(hipStreamSynchronize(s));
    index.tree.detach_streams();  // the index outlives the call's stream
    index.built = true;
  }
  const RectTree& tree = index.tree;
  S.bb = tree.bb;
  S.sx = S.bb[2] > S.bb[0] ? 65535.f / (S.bb[2] - S.bb[0]) : 0.f;
  S.sy = S.bb[3] > S.bb[1] ? 65535.f / (S.bb[3] - S.bb[1]) : 0.f;
  S.kt.nodes = tree.nodes.p;
  S.kt.leaf_rects = tree.leaf_rects.p;
  S.kt.leaf_ids = tree.leaf_ids.p;
  S.kt.obj = build.points_only ? build.xy.p : nullptr;
  S.kt.objs = build.view();
  S.kt.n_valid = tree.n_valid;
  S.kt.n_levels = tree.n_levels;
  for (int l = 0; l < tree.n_levels; l++) {
    S.kt.off[l] = tree.level_off[l];
    S.kt.cnt[l] = tree.level_cnt[l];
  }
  // visit queries in Morton order of their rects (outward FP32)
  S.qrects.alloc(nq > 0 ? nq : 1, s);
  if (probe.points_only) {
    if (nq > 0)
      hipLaunchKernelGGL(k_point_rects, blocks_for(nq), kBlock, 0, s, probe.xy.p, probe.type.p, nq, S.qrects.p);
  } else {
    compute_rects(probe, nullptr, false, NAN, S.qrects, s);
  }
  int64_t nvq = morton_order(S.qrects.p, nq, S.order, s);
  // invalid queries still need their output rows: run them unordered
  S.ord = nvq == nq ? S.order.p : nullptr;
}

}  // namespace

void knn_points(const GeomColumn& build, KnnIndex& index, const GeomColumn& probe, int k, uint32_t* out_ids,
                double* out_d, hipStream_t s, KernelTimer* timer) {
  if (k < 1 || k > 32) throw Error("knn: k must be in 1..32");
  const int64_t nq = probe.n_rows;
  if (nq == 0) return;
  KnnSetup S;
  knn_setup(build, index, probe, S, s);
  const uint32_t* keys = index.tree.leaf_keys.p;
  if (timer) timer->record_start(s);
  if (S.mode == 0) launch_knn_k<0>(S.kt, keys, S.bb, S.sx, S.sy, probe, S.qrects.p, S.ord, nq, k, out_ids, out_d, s);
  else if (S.mode == 1) launch_knn_k<1>(S.kt, keys, S.bb, S.sx, S.sy, probe, S.qrects.p, S.ord, nq, k, out_ids, out_d, s);
  else launch_knn_k<2>(S.kt, keys, S.bb, S.sx, S.sy, probe, S.qrects.p, S.ord, nq, k, out_ids, out_d, s);
  SJ_CHECK(hipGetLastError());
  if (timer) timer->record_stop(s);
  SJ_CHECK(hipStreamSynchronize(s));
}

uint64_t knn_with_ties(const GeomColumn& build, KnnIndex& index, const GeomColumn& probe, int k, double tol,
                       DevArray<unsigned long long>& offsets, DevArray<uint32_t>& out_ids, DevArray<double>& out_d,
                       hipStream_t s) {
  if (k < 1 || k > 32) throw Error("knn: k must be in 1..32");
  const int64_t nq = probe.n_rows;
  offsets.alloc(nq + 1, s);
  if (nq == 0) {
    SJ_CHECK(hipMemsetAsync(offsets.p, 0, 8, s));
    SJ_CHECK(hipStreamSynchronize(s));
    return 0;
  }
  DevArray<uint32_t> hid((size_t)nq * k, s);
  DevArray<double> hd((size_t)nq * k, s);
  knn_points(build, index, probe, k, hid.p, hd.p, s, nullptr);
  KnnSetup S;
  knn_setup(build, index, probe, S, s);
  DevArray<unsigned long long> counts(nq, s);
  auto count = [&](auto mode) {
    constexpr int M = decltype(mode)::value;
    hipLaunchKernelGGL((k_knn_tie_count<RectTree::FANOUT, RectTree::LEAF, M>), blocks_for(nq), kBlock, 0, s, S.kt,
                       probe.view(), S.qrects.p, nq, k, hd.p, tol, counts.p);
  };
  if (S.mode == 0) count(std::integral_constant<int, 0>());
  else if (S.mode == 1) count(std::integral_constant<int, 1>());
  else count(std::integral_constant<int, 2>());
  SJ_CHECK(hipGetLastError());
  // offsets = exclusive scan of the counts
  SJ_CHECK(hipMemsetAsync(offsets.p, 0, 8, s));
  size_t tb = 0;
  SJ_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tb, counts.p, offsets.p + 1, (int)nq, s));
  DevArray<uint8_t> tmp(tb + 16, s);
  SJ_CHECK(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, counts.p, offsets.p + 1, (int)nq, s));
  unsigned long long total = 0;
  SJ_CHECK(hipMemcpyAsync(&total, offsets.p + nq, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  out_ids.alloc(total > 0 ? total : 1, s);
  out_d.alloc(total > 0 ? total : 1, s);
  auto fill = [&](auto mode) {
    constexpr int M = decltype(mode)::value;
    hipLaunchKernelGGL((k_knn_tie_fill<RectTree::FANOUT, RectTree::LEAF, M>), blocks_for(nq), kBlock, 0, s, S.kt,
                       probe.view(), S.qrects.p, nq, k, hid.p, hd.p, tol, offsets.p, out_ids.p, out_d.p);
  };
  if (S.mode == 0) fill(std::integral_constant<int, 0>());
  else if (S.mode == 1) fill(std::integral_constant<int, 1>());
  else fill(std::integral_constant<int, 2>());
  SJ_CHECK(hipGetLastError());
  SJ_CHECK(hipStreamSynchronize(s));
  return total;
}

}  // namespace sjhip
