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
#include <cmath>
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

// ---- cell grid (RectTree::grid_*) ----------------------------------------

namespace {

// bounds of the valid rects as ordered uints: min x0, min y0, max x1, max y1
__global__ void k_rect_bounds(const float4* r, int64_t n, unsigned int* acc) {
  __shared__ unsigned int s[4][kBlock];
  const int t = threadIdx.x;
  unsigned int a = 0xffffffffu, b = 0xffffffffu, c = 0, d = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + t; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 q = r[i];
    if (!rect_ok(q)) continue;
    a = min(a, f2o(q.x)); b = min(b, f2o(q.y));
    c = max(c, f2o(q.z)); d = max(d, f2o(q.w));
  }
  s[0][t] = a; s[1][t] = b; s[2][t] = c; s[3][t] = d;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (t < o) {
      s[0][t] = min(s[0][t], s[0][t + o]);
      s[1][t] = min(s[1][t], s[1][t + o]);
      s[2][t] = max(s[2][t], s[2][t + o]);
      s[3][t] = max(s[3][t], s[3][t + o]);
    }
    __syncthreads();
  }
  if (t == 0) {
    atomicMin(&acc[0], s[0][0]);
    atomicMin(&acc[1], s[1][0]);
    atomicMax(&acc[2], s[2][0]);
    atomicMax(&acc[3], s[3][0]);
  }
}

// max(width, height) of every stride-th rect (-1 for an invalid one)
__global__ void k_rect_extents(const float4* r, int64_t n, int64_t stride, int n_out, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  const int64_t k = (int64_t)i * stride;
  if (k >= n) {
    out[i] = -1.f;
    return;
  }
  const float4 b = r[k];
  out[i] = rect_ok(b) ? fmaxf(b.z - b.x, b.w - b.y) : -1.f;
}

__global__ void k_grid_counts(const float4* r, int64_t n, GridView g, uint32_t* counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 b = r[i];
  uint32_t c = 0;
  if (rect_ok(b)) {
    const int cx0 = grid_cell(b.x, g.x0, g.inv_x, g.nx), cx1 = grid_cell(b.z, g.x0, g.inv_x, g.nx);
    const int cy0 = grid_cell(b.y, g.y0, g.inv_y, g.ny), cy1 = grid_cell(b.w, g.y0, g.inv_y, g.ny);
    const uint64_t cc = (uint64_t)(cx1 - cx0 + 1) * (uint64_t)(cy1 - cy0 + 1);
    c = cc > 0xffffffffull ? 0xffffffffu : (uint32_t)cc;
  }
  counts[i] = c;
}

__global__ void k_grid_fill(const float4* r, int64_t n, GridView g, const uint64_t* start, uint32_t* keys,
                            uint32_t* vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 b = r[i];
  if (!rect_ok(b)) return;
  const int cx0 = grid_cell(b.x, g.x0, g.inv_x, g.nx), cx1 = grid_cell(b.z, g.x0, g.inv_x, g.nx);
  const int cy0 = grid_cell(b.y, g.y0, g.inv_y, g.ny), cy1 = grid_cell(b.w, g.y0, g.inv_y, g.ny);
  uint64_t w = start[i];
  for (int cy = cy0; cy <= cy1; cy++)
    for (int cx = cx0; cx <= cx1; cx++) {
      keys[w] = (uint32_t)cy * (uint32_t)g.nx + (uint32_t)cx;
      vals[w] = (uint32_t)i;
      w++;
    }
}

// off[c] = first entry of cell c (entries sorted by cell), off[ncells] = T
__global__ void k_grid_starts(const uint32_t* keys, int64_t T, int64_t ncells, uint32_t* off) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e > T) return;
  const int64_t lo = e == 0 ? 0 : (int64_t)keys[e - 1] + 1;
  const int64_t hi = e == T ? ncells : (int64_t)keys[e];
  for (int64_t c = lo; c <= hi; c++) off[c] = (uint32_t)e;
}

__global__ void k_grid_gather(const float4* r, const uint32_t* vals, int64_t T, float4* out_r, uint32_t* out_id) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T) return;
  const uint32_t v = vals[e];
  out_r[e] = r[v];
  out_id[e] = v;
}

}  // namespace

// Cell width as a multiple of the median rect extent: narrower cells list
// each rect in more cells but leave fewer rects to test per query.
#ifndef SJ_GRID_CELL_SCALE
#define SJ_GRID_CELL_SCALE 1.0
#endif
constexpr double kGridCellScale = SJ_GRID_CELL_SCALE;

// Builds t's cell grid over the n rects.  Cells are kGridCellScale times
// the median rect extent wide (coarsened while the rects would be listed
// more than 16 times each on average); no grid when the rects' bounds are
// degenerate or too many entries remain.
void build_rect_grid(const float4* rects, int64_t n, RectTree& t, hipStream_t s) {
  t.grid_nx = t.grid_ny = 0;
  t.grid_off.release();
  t.grid_rects.release();
  t.grid_ids.release();
  if (n <= 0 || t.n_valid <= 0) return;
  DevArray<unsigned int> acc(4, s);
  const unsigned int init[4] = {0xffffffffu, 0xffffffffu, 0u, 0u};
  SJ_CHECK(hipMemcpyAsync(acc.p, init, sizeof(init), hipMemcpyHostToDevice, s));
  const unsigned grid = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, blocks_for(n)));
  hipLaunchKernelGGL(k_rect_bounds, grid, kBlock, 0, s, rects, n, acc.p);
  const int n_sample = (int)std::min<int64_t>(n, 4096);
  const int64_t stride = std::max<int64_t>(1, n / n_sample);
  DevArray<float> ext(n_sample, s);
  hipLaunchKernelGGL(k_rect_extents, blocks_for(n_sample), kBlock, 0, s, rects, n, stride, n_sample, ext.p);
  SJ_CHECK(hipGetLastError());
  unsigned int hb[4];
  std::vector<float> he(n_sample);
  SJ_CHECK(hipMemcpyAsync(hb, acc.p, sizeof(hb), hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipMemcpyAsync(he.data(), ext.p, n_sample * sizeof(float), hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  const float X0 = o2f(hb[0]), Y0 = o2f(hb[1]), X1 = o2f(hb[2]), Y1 = o2f(hb[3]);
  if (!(X1 > X0) || !(Y1 > Y0) || !std::isfinite(X1 - X0) || !std::isfinite(Y1 - Y0)) return;
  std::vector<float> valid;
  for (float v : he)
    if (v >= 0.f) valid.push_back(v);
  if (valid.empty()) return;
  std::nth_element(valid.begin(), valid.begin() + valid.size() / 2, valid.end());
  double cs = std::max<double>(valid[valid.size() / 2] * kGridCellScale, std::max(X1 - X0, Y1 - Y0) / 4096.0);
  DevArray<uint32_t> counts(n, s);
  DevArray<uint64_t> start(n + 1, s);
  size_t tb = 0;
  SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts.p, start.p, (int)(n + 1), s));
  DevArray<uint8_t> tmp(tb + 16, s);
  GridView g{};
  uint64_t total = 0;
  for (int attempt = 0;; attempt++) {
    const double nxd = std::ceil((X1 - X0) / cs), nyd = std::ceil((Y1 - Y0) / cs);
    if (nxd * nyd > double(1 << 24)) {
      cs *= 2;
      continue;
    }
    g.nx = std::max(1, (int)nxd);
    g.ny = std::max(1, (int)nyd);
    g.x0 = X0;
    g.y0 = Y0;
    g.inv_x = (float)(g.nx / (double(X1) - X0));
    g.inv_y = (float)(g.ny / (double(Y1) - Y0));
    hipLaunchKernelGGL(k_grid_counts, blocks_for(n), kBlock, 0, s, rects, n, g, counts.p);
    // start[i] = entries before rect i (start[n] = the total), in 64 bits
    SJ_CHECK(hipMemsetAsync(start.p + n, 0, 8, s));
    SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, counts.p, start.p, (int)n, s));
    uint64_t last_start = 0;
    uint32_t last_cnt = 0;
    SJ_CHECK(hipMemcpyAsync(&last_start, start.p + n - 1, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipMemcpyAsync(&last_cnt, counts.p + n - 1, 4, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    total = last_start + last_cnt;
    if (total <= (uint64_t)16 * (uint64_t)t.n_valid + 1024 && total < 0xffffffffull) break;
    if (attempt >= 8) return;  // no grid: the tree serves these rects
    cs *= 2;
  }
  const int64_t T = (int64_t)total;
  const int64_t ncells = (int64_t)g.nx * g.ny;
  DevArray<uint32_t> keys(T, s), vals(T, s), k2(T, s), v2(T, s);
  hipLaunchKernelGGL(k_grid_fill, blocks_for(n), kBlock, 0, s, rects, n, g, start.p, keys.p, vals.p);
  int bits = 1;
  while ((int64_t(1) << bits) < ncells) bits++;
  size_t sb = 0;
  SJ_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sb, keys.p, k2.p, vals.p, v2.p, (int)T, 0, bits, s));
  DevArray<uint8_t> st(sb + 16, s);
  SJ_CHECK(hipcub::DeviceRadixSort::SortPairs(st.p, sb, keys.p, k2.p, vals.p, v2.p, (int)T, 0, bits, s));
  t.grid_off.alloc(ncells + 1, s);
  hipLaunchKernelGGL(k_grid_starts, blocks_for(T + 1), kBlock, 0, s, k2.p, T, ncells, t.grid_off.p);
  t.grid_rects.alloc(T > 0 ? T : 1, s);
  t.grid_ids.alloc(T > 0 ? T : 1, s);
  if (T > 0) hipLaunchKernelGGL(k_grid_gather, blocks_for(T), kBlock, 0, s, rects, v2.p, T, t.grid_rects.p, t.grid_ids.p);
  SJ_CHECK(hipGetLastError());
  SJ_CHECK(hipStreamSynchronize(s));
  t.grid_nx = g.nx;
  t.grid_ny = g.ny;
  t.grid_x0 = g.x0;
  t.grid_y0 = g.y0;
  t.grid_inv_x = g.inv_x;
  t.grid_inv_y = g.inv_y;
}

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
  build_rect_grid(rects, n, t, s);
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

// The grid filter for point queries when the tree has a grid (the tree
// walk otherwise).
bool use_grid(const RectTree& t, bool point_probes) {
#ifdef SJ_NO_GRID_FILTER
  return false;
#else
  if (t.grid_nx <= 0 || t.filter_mode == 1) return false;
  return point_probes || t.filter_mode == 2;
#endif
}

void launch_grid_filter(const RectTree& t, const float4* q, const double2* xy, int64_t nq, const uint32_t* qorder,
                        int er, uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter, hipStream_t s) {
  const unsigned grid = (unsigned)((nq + kFilterGridBlock - 1) / kFilterGridBlock);
  hipLaunchKernelGGL(k_filter_grid, grid, kFilterGridBlock, 0, s, grid_view(t), q, xy, nq, qorder, er, ob, op, cap,
                     counter);
}

uint64_t filter_pairs(const RectTree& t, const float4* q, int64_t nq, const uint32_t* qorder, PairBuf& pb,
                      hipStream_t s, KernelTimer* timer, bool emit_rank, bool point_probes,
                      const double2* ordered_points) {
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
    if (use_grid(t, point_probes))
      launch_grid_filter(t, q, emit_rank ? ordered_points : nullptr, nq, qorder, emit_rank ? 1 : 0, pb.b.p, pb.p.p,
                         pb.cap(), pb.counter.p, s);
    else
      launch_plain_filter(t, v, q, nq, qorder, emit_rank ? 1 : 0, pb.b.p, pb.p.p, pb.cap(), pb.counter.p, s);
    SJ_CHECK(hipGetLastError());
    if (timer) timer->record_stop(s);
    unsigned long long total = 0;
    SJ_CHECK(hipMemcpyAsync(&total, pb.counter.p, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    if (nq > 0) t.cand_per_query_x16.store((uint64_t)((total * 16) / (uint64_t)nq + 1), std::memory_order_relaxed);
    // the refine kernels index candidates with 32 bits
    if (total >= 0xffffffffull - 64) throw Error("filter: 2^32 - 64 candidates or more: probe in smaller batches");
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

namespace {
// k_query_keys of the point rects (fx, fy, fx, fy), fx = x rounded to
// FP32, read from the points themselves; rows that are empty, null or NaN
// get the invalid key
__global__ void k_point_keys(const double2* xy, const uint8_t* type, int64_t n, float x0, float y0, float sx,
                             float sy, uint32_t* keys, uint32_t* ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 p = xy[i];
  const float fx = (float)p.x, fy = (float)p.y;
  uint32_t k = 0xffffffffu;
  if (type[i] != 0 && fx == fx && fy == fy) {
    const float cx = 0.5f * fx + 0.5f * fx, cy = 0.5f * fy + 0.5f * fy;
    const uint32_t qx = (uint32_t)fminf(fmaxf((cx - x0) * sx, 0.f), 65535.f);
    const uint32_t qy = (uint32_t)fminf(fmaxf((cy - y0) * sy, 0.f), 65535.f);
    k = spread16(qx) | (spread16(qy) << 1);
    if (k == 0xffffffffu) k = 0xfffffffeu;
  }
  keys[i] = k;
  ids[i] = (uint32_t)i;
}
}  // namespace

int64_t morton_order_points(const GeomColumn& g, DevArray<uint32_t>& order, hipStream_t s, MortonFrame* frame,
                            DevArray<uint32_t>* sorted_keys) {
  if (!g.points_only || g.pt_valid < 0 || (g.n_rows > 0 && g.type.p == nullptr))
    throw Error("morton_order_points needs a decoded point column");
  const int64_t n = g.n_rows, nv = g.pt_valid;
  if (nv == 0) {
    order.alloc(1, s);
    return 0;
  }
  const float* bb = g.pt_bounds;
  const float sx = bb[2] > bb[0] ? 65535.f / (bb[2] - bb[0]) : 0.f;
  const float sy = bb[3] > bb[1] ? 65535.f / (bb[3] - bb[1]) : 0.f;
  DevArray<uint32_t> keys(n, s), ids(n, s);
  hipLaunchKernelGGL(k_point_keys, blocks_for(n), kBlock, 0, s, g.xy.p, g.type.p, n, bb[0], bb[1], sx, sy, keys.p,
                     ids.p);
  SJ_CHECK(hipGetLastError());
  // (all 32 bits: rocPRIM's merge-sort path, which it takes below about 1M
  // keys, builds its comparator's mask as (1 << (begin_bit + bits)) - 1,
  // a 32-bit shift when a range [begin_bit > 0, 32) is asked for, which
  // yields the low bits instead; its merge phase then works on runs sorted
  // by another key than its own and indexes out of bounds)
  sort_keys(keys, ids, n, s);
  order = std::move(ids);
  if (frame) {
    frame->x0 = bb[0];
    frame->y0 = bb[1];
    frame->sx = sx;
    frame->sy = sy;
  }
  if (sorted_keys) *sorted_keys = std::move(keys);
  return nv;
}

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
                   uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter, hipStream_t s,
                   bool point_probes) {
  if (nq <= 0 || t.n_levels == 0) return;
  if (use_grid(t, point_probes)) {
    launch_grid_filter(t, q, nullptr, nq, qorder, emit_rank ? 1 : 0, ob, op, cap, counter, s);
  } else {
    TreeView v = tree_view(t);
    launch_plain_filter(t, v, q, nq, qorder, emit_rank ? 1 : 0, ob, op, cap, counter, s);
  }
  SJ_CHECK(hipGetLastError());
}

int64_t morton_order(const float4* rects, int64_t n, DevArray<uint32_t>& order, hipStream_t s, MortonFrame* frame,
                     DevArray<uint32_t>* sorted_keys) {
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
  if (frame) {
    frame->x0 = bb[0];
    frame->y0 = bb[1];
    frame->sx = sx;
    frame->sy = sy;
  }
  if (sorted_keys) *sorted_keys = std::move(keys);
  return nv;
}

}  // namespace sjhip
