// knn.hip -- exact k-nearest-neighbour join over point columns.
//
// Semantics of the reference (rust/sedona-spatial-join/src/index/
// default_spatial_index.rs:306-502 with the ground truth of
// rust/sedona-spatial-join/tests/spatial_join_integration.rs:1373-1414):
// distance = hypot(dx, dy) in FP64 exactly as the reference's host
// computes it, neighbours ordered by (distance, build id), first k kept.
//
// Device plan: the object points sit in the Morton-ordered rect tree
// (FP32 rects rounded outward so they contain the FP64 points).  Each lane
// owns one query and a bounded max-heap of K (distance, id) entries kept
// sorted in registers.  Phase 1 seeds the heap with the objects next to
// the query's Morton position (a binary search over the sorted leaf
// keys); phase 2 walks the tree depth-first with its stack in LDS and
// prunes every node whose FP64 lower-bound distance exceeds the current
// k-th distance, so ties at the k-th distance are always examined.
#include <hipcub/hipcub.hpp>

#include "geom_math.h"
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
  const double2* obj;  // object points by build id
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

template <int K, int F, int LEAF>
__global__ __launch_bounds__(kBlock) void k_knn(KnnTree t, const uint32_t* leaf_keys, float x0, float y0, float sx,
                                                float sy, const double2* q, const uint8_t* qtype, int64_t nq,
                                                const uint32_t* qorder, int k, uint32_t* out_ids, double* out_d) {
  __shared__ uint32_t s_node[kMaxLevels + 1][kBlock];
  __shared__ uint32_t s_mask[kMaxLevels + 1][kBlock];
  const int tid = threadIdx.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  if (i >= nq) return;
  uint32_t qi = qorder ? qorder[i] : (uint32_t)i;
  Heap<K> h;
  h.init();
  double2 qp = q[qi];
  bool valid = qtype[qi] != 0 && qp.x == qp.x && qp.y == qp.y && t.n_levels > 0;
  if (valid) {
    // phase 1: neighbours in Morton order
    float fx = ((float)qp.x - x0) * sx, fy = ((float)qp.y - y0) * sy;
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
      double2 o = t.obj[id];
      h.push(hypot_ref(qp.x - o.x, qp.y - o.y), id);
    }
    // phase 2: pruned depth-first walk
    const int top = t.n_levels;
    int lvl = top;
    {
      uint32_t m = 0;
      const float4* nd = t.nodes + t.off[top - 1];
      int64_t nc = t.cnt[top - 1];
      for (int c = 0; c < F; c++)
        if (c < nc && !(box_lb(nd[c], qp.x, qp.y) > h.bound())) m |= 1u << c;
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
          if (box_lb(r, qp.x, qp.y) > h.bound()) continue;
          uint32_t id = t.leaf_ids[j];
          double2 o = t.obj[id];
          h.push(hypot_ref(qp.x - o.x, qp.y - o.y), id);
        }
      } else {
        // re-check this node against the (possibly smaller) bound
        const float4* own = t.nodes + t.off[lvl - 1];
        if (box_lb(own[child], qp.x, qp.y) > h.bound()) continue;
        int cl = lvl - 2;
        const float4* nd = t.nodes + t.off[cl];
        int64_t c0 = child * F;
        int64_t nc = t.cnt[cl] - c0;
        uint32_t mm = 0;
#pragma unroll
        for (int kk = 0; kk < F; kk++)
          if (kk < nc && !(box_lb(nd[c0 + kk], qp.x, qp.y) > h.bound())) mm |= 1u << kk;
        lvl = lvl - 1;
        s_node[lvl][tid] = (uint32_t)child;
        s_mask[lvl][tid] = mm;
      }
    }
  }
  for (int j = 0; j < k; j++) {
    out_ids[(int64_t)qi * k + j] = h.id[j];
    out_d[(int64_t)qi * k + j] = h.d[j];
  }
}

unsigned int blocks_for(int64_t n) { return (unsigned int)((n + kBlock - 1) / kBlock); }

template <int K>
void launch_knn(const KnnTree& kt, const uint32_t* keys, const float bb[4], float sx, float sy, const GeomColumn& probe,
                const uint32_t* order, int64_t nq, int k, uint32_t* ids, double* d, hipStream_t s) {
  hipLaunchKernelGGL((k_knn<K, RectTree::FANOUT, RectTree::LEAF>), blocks_for(nq), kBlock, 0, s, kt, keys, bb[0],
                     bb[1], sx, sy, probe.xy.p, probe.type.p, nq, order, k, ids, d);
}

}  // namespace

// bounds helper shared with bvh.hip semantics: centre bounds of valid rects
void knn_points(const GeomColumn& build, KnnIndex& index, const GeomColumn& probe, int k, uint32_t* out_ids,
                double* out_d, hipStream_t s, KernelTimer* timer) {
  if (k < 1 || k > 32) throw Error("knn: k must be in 1..32");
  if (!build.points_only) throw Error("knn: build rows must be points");
  if (!probe.points_only) throw Error("knn: probe rows must be points");
  const int64_t nb = build.n_rows, nq = probe.n_rows;
  if (nq == 0) return;
  if (!index.built) {
    DevArray<float4> rects(nb > 0 ? nb : 1, s);
    if (nb > 0)
      hipLaunchKernelGGL(k_point_rects, blocks_for(nb), kBlock, 0, s, build.xy.p, build.type.p, nb, rects.p);
    build_rect_tree(rects.p, nb, index.tree, s);
    SJ_CHECK(hipStreamSynchronize(s));
    index.tree.detach_streams();  // the index outlives the call's stream
    index.built = true;
  }
  const RectTree& tree = index.tree;
  const float* bb = tree.bb;
  float sx = bb[2] > bb[0] ? 65535.f / (bb[2] - bb[0]) : 0.f;
  float sy = bb[3] > bb[1] ? 65535.f / (bb[3] - bb[1]) : 0.f;
  const uint32_t* keys = tree.leaf_keys.p;
  KnnTree kt{};
  kt.nodes = tree.nodes.p;
  kt.leaf_rects = tree.leaf_rects.p;
  kt.leaf_ids = tree.leaf_ids.p;
  kt.obj = build.xy.p;
  kt.n_valid = tree.n_valid;
  kt.n_levels = tree.n_levels;
  for (int l = 0; l < tree.n_levels; l++) {
    kt.off[l] = tree.level_off[l];
    kt.cnt[l] = tree.level_cnt[l];
  }
  // visit queries in Morton order
  DevArray<float4> qrects(nq, s);
  hipLaunchKernelGGL(k_point_rects, blocks_for(nq), kBlock, 0, s, probe.xy.p, probe.type.p, nq, qrects.p);
  DevArray<uint32_t> order;
  int64_t nvq = morton_order(qrects.p, nq, order, s);
  // invalid queries still need their output rows: run them unordered
  const uint32_t* ord = nvq == nq ? order.p : nullptr;
  if (timer) timer->record_start(s);
  if (k <= 1) launch_knn<1>(kt, keys, bb, sx, sy, probe, ord, nq, k, out_ids, out_d, s);
  else if (k <= 2) launch_knn<2>(kt, keys, bb, sx, sy, probe, ord, nq, k, out_ids, out_d, s);
  else if (k <= 4) launch_knn<4>(kt, keys, bb, sx, sy, probe, ord, nq, k, out_ids, out_d, s);
  else if (k <= 8) launch_knn<8>(kt, keys, bb, sx, sy, probe, ord, nq, k, out_ids, out_d, s);
  else if (k <= 16) launch_knn<16>(kt, keys, bb, sx, sy, probe, ord, nq, k, out_ids, out_d, s);
  else launch_knn<32>(kt, keys, bb, sx, sy, probe, ord, nq, k, out_ids, out_d, s);
  SJ_CHECK(hipGetLastError());
  if (timer) timer->record_stop(s);
  SJ_CHECK(hipStreamSynchronize(s));
}

}  // namespace sjhip
