// bvh.hip -- linear bounding-volume hierarchy over FP32 rects and the
// filter kernel that walks it.
//
// Replaces the reference's OptiX acceleration structure and ray launches
// (c/sedona-libgpuspatial/libgpuspatial/src/rt_spatial_index.cu:309-476,
// src/rt/shaders/point_query.cu:30-73, box_query_forward.cu:30-89,
// box_query_backward.cu:30-83) and, for the oracle, the CPU R-tree
// (rust/sedona-spatial-join/src/index/default_spatial_index_builder.rs:
// 164-234).  Semantics kept: a (build, probe) pair is a candidate when
// their FP32 rects intersect with inclusive bounds; NaN / inverted rects
// never match.
//
// Build: Morton codes (16 bits per axis) of rect centres, a radix sort,
// then a bottom-up merge: LEAF consecutive sorted rects form a leaf node,
// FANOUT consecutive nodes form a parent, level by level.  The tree is
// implicit (children of node i are i*FANOUT .. i*FANOUT+FANOUT-1), so no
// child pointers are stored and each level is one contiguous float4 run.
//
// Filter: one lane per probe rect; the lane's traversal stack (node and
// remaining-children bitmask per level) lives in LDS; the loop is
// wavefront-uniform so that candidates found in one step are compacted
// with a wavefront prefix sum and a single atomic per wavefront.
#include <hipcub/hipcub.hpp>

#include "lbvh_filter.h"
#include <algorithm>
#include "sjhip.h"
#include "wave_out.h"

namespace sjhip {

namespace {

constexpr int kBlock = kFilterBlock;

__device__ __forceinline__ uint32_t spread16(uint32_t x) {
  x &= 0xffff;
  x = (x | (x << 8)) & 0x00ff00ff;
  x = (x | (x << 4)) & 0x0f0f0f0f;
  x = (x | (x << 2)) & 0x33333333;
  x = (x | (x << 1)) & 0x55555555;
  return x;
}

// float <-> order-preserving uint
__device__ __forceinline__ uint32_t f2o(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float o2f(uint32_t u) {
  uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  __builtin_memcpy(&f, &b, 4);
  return f;
}

// bounds of valid rect centres: [0] minx [1] miny [2] maxx [3] maxy as
// ordered uints, [4] valid count
__global__ void k_center_bounds(const float4* r, int64_t n, unsigned int* acc) {
  __shared__ unsigned int s[5][kBlock];
  int t = threadIdx.x;
  unsigned int mnx = 0xffffffffu, mny = 0xffffffffu, mxx = 0, mxy = 0, cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + t; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float4 b = r[i];
    if (!rect_ok(b)) continue;
    float cx = 0.5f * b.x + 0.5f * b.z, cy = 0.5f * b.y + 0.5f * b.w;
    unsigned int ox = f2o(cx), oy = f2o(cy);
    mnx = min(mnx, ox); mny = min(mny, oy);
    mxx = max(mxx, ox); mxy = max(mxy, oy);
    cnt++;
  }
  s[0][t] = mnx; s[1][t] = mny; s[2][t] = mxx; s[3][t] = mxy; s[4][t] = cnt;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (t < o) {
      s[0][t] = min(s[0][t], s[0][t + o]);
      s[1][t] = min(s[1][t], s[1][t + o]);
      s[2][t] = max(s[2][t], s[2][t + o]);
      s[3][t] = max(s[3][t], s[3][t + o]);
      s[4][t] += s[4][t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
    atomicMin(&acc[0], s[0][0]);
    atomicMin(&acc[1], s[1][0]);
    atomicMax(&acc[2], s[2][0]);
    atomicMax(&acc[3], s[3][0]);
    atomicAdd(&acc[4], s[4][0]);
  }
}

__global__ void k_morton(const float4* r, int64_t n, float x0, float y0, float sx, float sy, uint32_t* keys,
                         uint32_t* ids) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 b = r[i];
  uint32_t k = 0xffffffffu;
  if (rect_ok(b)) {
    float cx = 0.5f * b.x + 0.5f * b.z, cy = 0.5f * b.y + 0.5f * b.w;
    float fx = (cx - x0) * sx, fy = (cy - y0) * sy;
    uint32_t qx = (uint32_t)fminf(fmaxf(fx, 0.f), 65535.f);
    uint32_t qy = (uint32_t)fminf(fmaxf(fy, 0.f), 65535.f);
    k = spread16(qx) | (spread16(qy) << 1);
    if (k == 0xffffffffu) k = 0xfffffffeu;
  }
  keys[i] = k;
  ids[i] = (uint32_t)i;
}

__global__ void k_gather_leaves(const float4* r, const uint32_t* sorted_ids, int64_t n, float4* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = r[sorted_ids[i]];
}

// Extent (max of width, height) of up to n_out leaf nodes, every stride-th.
__global__ void k_leaf_extents(const float4* nodes, int64_t n_nodes, int64_t stride, int n_out, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  const int64_t k = (int64_t)i * stride;
  if (k >= n_nodes) {
    out[i] = 0.f;
    return;
  }
  const float4 b = nodes[k];
  out[i] = fmaxf(b.z - b.x, b.w - b.y);
}

__global__ void k_merge(const float4* child, int64_t n_child, int group, float4* parent, int64_t n_parent) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_parent) return;
  int64_t c0 = i * group, c1 = min(c0 + group, n_child);
  float4 u = child[c0];
  for (int64_t c = c0 + 1; c < c1; c++) {
    float4 b = child[c];
    u.x = fminf(u.x, b.x); u.y = fminf(u.y, b.y);
    u.z = fmaxf(u.z, b.z); u.w = fmaxf(u.w, b.w);
  }
  parent[i] = u;
}

unsigned int blocks_for(int64_t n) { return (unsigned int)((n + kBlock - 1) / kBlock); }

void center_bounds(const float4* r, int64_t n, float bb[4], int64_t* n_valid, hipStream_t s) {
  DevArray<unsigned int> acc(5, s);
  unsigned int init[5] = {0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  SJ_CHECK(hipMemcpyAsync(acc.p, init, sizeof(init), hipMemcpyHostToDevice, s));
  unsigned int grid = (unsigned int)std::min<int64_t>(4096, std::max<int64_t>(1, blocks_for(n)));
  if (n > 0) hipLaunchKernelGGL(k_center_bounds, grid, kBlock, 0, s, r, n, acc.p);
  unsigned int h[5];
  SJ_CHECK(hipMemcpyAsync(h, acc.p, sizeof(h), hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  *n_valid = h[4];
  for (int k = 0; k < 4; k++) bb[k] = h[4] ? o2f(h[k]) : 0.f;
}

void sort_keys(DevArray<uint32_t>& keys, DevArray<uint32_t>& ids, int64_t n, hipStream_t s) {
  DevArray<uint32_t> k2(n, s), i2(n, s);
  size_t tmp = 0;
  SJ_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys.p, k2.p, ids.p, i2.p, (int)n, 0, 32, s));
  DevArray<uint8_t> t(tmp + 16, s);
  SJ_CHECK(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, keys.p, k2.p, ids.p, i2.p, (int)n, 0, 32, s));
  keys = std::move(k2);
  ids = std::move(i2);
}

}  // namespace

void build_rect_tree(const float4* rects, int64_t n, RectTree& t, hipStream_t s, int leaf) {
  t.level_off.clear();
  t.level_cnt.clear();
  t.n_levels = 0;
  t.coop_extent = 0.f;
  float bb[4];
  center_bounds(rects, n, bb, &t.n_valid, s);
  if (t.n_valid == 0) {
    t.nodes.alloc(1, s);
    t.leaf_rects.alloc(1, s);
    t.leaf_ids.alloc(1, s);
    return;
  }
  DevArray<uint32_t> keys(n, s), ids(n, s);
  float sx = bb[2] > bb[0] ? 65535.f / (bb[2] - bb[0]) : 0.f;
  float sy = bb[3] > bb[1] ? 65535.f / (bb[3] - bb[1]) : 0.f;
  hipLaunchKernelGGL(k_morton, blocks_for(n), kBlock, 0, s, rects, n, bb[0], bb[1], sx, sy, keys.p, ids.p);
  sort_keys(keys, ids, n, s);
  const int64_t nv = t.n_valid;
  t.leaf_rects.alloc(nv, s);
  hipLaunchKernelGGL(k_gather_leaves, blocks_for(nv), kBlock, 0, s, rects, ids.p, nv, t.leaf_rects.p);
  t.leaf_ids.alloc(nv, s);
  SJ_CHECK(hipMemcpyAsync(t.leaf_ids.p, ids.p, nv * 4, hipMemcpyDeviceToDevice, s));
  t.leaf_keys.alloc(nv, s);
  SJ_CHECK(hipMemcpyAsync(t.leaf_keys.p, keys.p, nv * 4, hipMemcpyDeviceToDevice, s));
  for (int k = 0; k < 4; k++) t.bb[k] = bb[k];
  // level sizes
  std::vector<int64_t> cnt;
  if (leaf != 4 && leaf != 8 && leaf != 16) throw Error("rect tree leaf size must be 4, 8 or 16");
  t.leaf = leaf;
  int64_t c = (nv + leaf - 1) / leaf;
  cnt.push_back(c);
  while (c > RectTree::FANOUT) {
    c = (c + RectTree::FANOUT - 1) / RectTree::FANOUT;
    cnt.push_back(c);
  }
  if ((int)cnt.size() > kMaxLevels - 1) throw Error("rect tree too deep");
  int64_t tot = 0;
  for (int64_t x : cnt) {
    t.level_off.push_back(tot);
    tot += x;
  }
  t.level_cnt = cnt;
  t.n_levels = (int)cnt.size();
  t.nodes.alloc(tot, s);
  hipLaunchKernelGGL(k_merge, blocks_for(cnt[0]), kBlock, 0, s, t.leaf_rects.p, nv, leaf,
                     t.nodes.p + t.level_off[0], cnt[0]);
  for (size_t l = 1; l < cnt.size(); l++)
    hipLaunchKernelGGL(k_merge, blocks_for(cnt[l]), kBlock, 0, s, t.nodes.p + t.level_off[l - 1], cnt[l - 1],
                       RectTree::FANOUT, t.nodes.p + t.level_off[l], cnt[l]);
  SJ_CHECK(hipGetLastError());
  // the median extent of (a strided sample of) the leaf nodes
  const int n_sample = (int)std::min<int64_t>(cnt[0], 4096);
  const int64_t stride = std::max<int64_t>(1, cnt[0] / n_sample);
  DevArray<float> ext(n_sample, s);
  hipLaunchKernelGGL(k_leaf_extents, blocks_for(n_sample), kBlock, 0, s, t.nodes.p + t.level_off[0], cnt[0], stride,
                     n_sample, ext.p);
  SJ_CHECK(hipGetLastError());
  std::vector<float> h(n_sample);
  SJ_CHECK(hipMemcpyAsync(h.data(), ext.p, n_sample * sizeof(float), hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  std::nth_element(h.begin(), h.begin() + n_sample / 2, h.end());
  t.coop_extent = 2.0f * h[n_sample / 2];
}

#ifndef SJ_FILTER_WAVE
#define SJ_FILTER_WAVE 1
#endif

// The plain filter over the tree: wavefront-cooperative (k_filter_wave)
// when the queries come in a locality order, else one walk per lane.
void launch_plain_filter(const RectTree& t, const TreeView& v, const float4* q, int64_t nq, const uint32_t* qorder,
                         int er, uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter,
                         hipStream_t s) {
  if (SJ_FILTER_WAVE && qorder) {
    const unsigned grid = (unsigned)((nq + kFilterWaveBlock - 1) / kFilterWaveBlock);
    if (t.leaf == 4)
      hipLaunchKernelGGL((k_filter_wave<RectTree::FANOUT, 4>), grid, kFilterWaveBlock, 0, s, v, q, nq, qorder, er,
                         ob, op, cap, counter);
    else if (t.leaf == 16)
      hipLaunchKernelGGL((k_filter_wave<RectTree::FANOUT, 16>), grid, kFilterWaveBlock, 0, s, v, q, nq, qorder, er,
                         ob, op, cap, counter);
    else
      hipLaunchKernelGGL((k_filter_wave<RectTree::FANOUT, 8>), grid, kFilterWaveBlock, 0, s, v, q, nq, qorder, er,
                         ob, op, cap, counter);
  } else if (t.leaf == 4) {
    hipLaunchKernelGGL((k_filter<RectTree::FANOUT, 4, KeepAll>), blocks_for(nq), kBlock, 0, s, v, q, nq, qorder, er,
                       KeepAll{}, ob, op, cap, counter);
  } else if (t.leaf == 16) {
    hipLaunchKernelGGL((k_filter<RectTree::FANOUT, 16, KeepAll>), blocks_for(nq), kBlock, 0, s, v, q, nq, qorder, er,
                       KeepAll{}, ob, op, cap, counter);
  } else {
    hipLaunchKernelGGL((k_filter<RectTree::FANOUT, 8, KeepAll>), blocks_for(nq), kBlock, 0, s, v, q, nq, qorder, er,
                       KeepAll{}, ob, op, cap, counter);
  }
}

uint64_t filter_pairs(const RectTree& t, const float4* q, int64_t nq, const uint32_t* qorder, PairBuf& pb,
                      hipStream_t s, KernelTimer* timer, bool emit_rank) {
  if (pb.counter.n < 1) pb.counter.alloc(1, s);
  // first capacity: the tree's observed candidates per query (16ths) from
  // earlier probes, else 4 per query; one resize-and-rerun on overflow
  uint64_t per16 = t.cand_per_query_x16.load(std::memory_order_relaxed);
  size_t want = std::max<int64_t>(1 << 16, per16 ? (int64_t)((per16 * (uint64_t)nq) / 16 * 9 / 8 + 1024) : nq * 4);
  if (pb.cap() < want) {
    pb.b.alloc(want, s);
    pb.p.alloc(want, s);
  }
  if (nq == 0 || t.n_levels == 0) return 0;
  TreeView v = tree_view(t);
  for (int attempt = 0; attempt < 2; attempt++) {
    SJ_CHECK(hipMemsetAsync(pb.counter.p, 0, 8, s));
    if (timer) timer->record_start(s);
    launch_plain_filter(t, v, q, nq, qorder, emit_rank ? 1 : 0, pb.b.p, pb.p.p, pb.cap(), pb.counter.p, s);
    SJ_CHECK(hipGetLastError());
    if (timer) timer->record_stop(s);
    unsigned long long total = 0;
    SJ_CHECK(hipMemcpyAsync(&total, pb.counter.p, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    if (nq > 0) t.cand_per_query_x16.store((uint64_t)((total * 16) / (uint64_t)nq + 1), std::memory_order_relaxed);
    if (total <= pb.cap()) return total;
    size_t cap = (size_t)(total + total / 8 + 1024);
    pb.b.alloc(cap, s);
    pb.p.alloc(cap, s);
  }
  throw Error("filter output did not fit after resizing");
}

namespace {
__global__ void k_query_keys(const float4* r, int64_t n, float x0, float y0, float sx, float sy, uint32_t* keys,
                             uint32_t* ids) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 b = r[i];
  uint32_t k = 0xffffffffu;
  if (rect_ok(b)) {
    float cx = 0.5f * b.x + 0.5f * b.z, cy = 0.5f * b.y + 0.5f * b.w;
    uint32_t qx = (uint32_t)fminf(fmaxf((cx - x0) * sx, 0.f), 65535.f);
    uint32_t qy = (uint32_t)fminf(fmaxf((cy - y0) * sy, 0.f), 65535.f);
    k = spread16(qx) | (spread16(qy) << 1);
    if (k == 0xffffffffu) k = 0xfffffffeu;
  }
  keys[i] = k;
  ids[i] = (uint32_t)i;
}
}  // namespace

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_u64, blocks_for(n), kBlock, 0, s, p, n, v);
  SJ_CHECK(hipGetLastError());
}

size_t morton_sort_temp_bytes(int64_t n) { return graph_sort_temp_bytes(n); }

void morton_order_fixed(const RectTree& t, const float4* rects, int64_t n, uint32_t* keys, uint32_t* ids,
                        uint32_t* keys_sorted, uint32_t* order, void* temp, size_t temp_bytes, hipStream_t s) {
  if (n <= 0) return;
  const float* bb = t.bb;
  float sx = bb[2] > bb[0] ? 65535.f / (bb[2] - bb[0]) : 0.f;
  float sy = bb[3] > bb[1] ? 65535.f / (bb[3] - bb[1]) : 0.f;
  hipLaunchKernelGGL(k_query_keys, blocks_for(n), kBlock, 0, s, rects, n, bb[0], bb[1], sx, sy, keys, ids);
  SJ_CHECK(hipGetLastError());
  // not rocPRIM's onesweep sort: it clears its digit counters and lookback
  // states with hipMemsetAsync, and a memset node in a replayed graph does
  // not reliably clear them again
  graph_sort_pairs(keys, ids, keys_sorted, order, n, 0, 32, temp, temp_bytes, s);
}

void launch_filter(const RectTree& t, const float4* q, int64_t nq, const uint32_t* qorder, bool emit_rank,
                   uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter, hipStream_t s) {
  if (nq <= 0 || t.n_levels == 0) return;
  TreeView v = tree_view(t);
  launch_plain_filter(t, v, q, nq, qorder, emit_rank ? 1 : 0, ob, op, cap, counter, s);
  SJ_CHECK(hipGetLastError());
}

int64_t morton_order(const float4* rects, int64_t n, DevArray<uint32_t>& order, hipStream_t s) {
  float bb[4];
  int64_t nv;
  center_bounds(rects, n, bb, &nv, s);
  if (nv == 0) {
    order.alloc(1, s);
    return 0;
  }
  DevArray<uint32_t> keys(n, s), ids(n, s);
  float sx = bb[2] > bb[0] ? 65535.f / (bb[2] - bb[0]) : 0.f;
  float sy = bb[3] > bb[1] ? 65535.f / (bb[3] - bb[1]) : 0.f;
  hipLaunchKernelGGL(k_query_keys, blocks_for(n), kBlock, 0, s, rects, n, bb[0], bb[1], sx, sy, keys.p, ids.p);
  sort_keys(keys, ids, n, s);
  order = std::move(ids);
  return nv;
}

}  // namespace sjhip
