// lbvh_filter.h -- the filter walk over the rect tree (see bvh.hip for
// the tree's layout), shared by the plain filter and the fused filter +
// refine kernels.
//
// One lane per probe rect; the lane's remaining-children bitmask per level
// lives in LDS (the node path itself is derived: children of node v are
// v*F .. v*F+F-1); the loop is wavefront-uniform so that the pairs found in
// one step are compacted with a wavefront prefix sum into a per-wavefront
// LDS buffer flushed with one atomic per kWaveOutCap pairs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sjhip.h"
#include "wave_out.h"

namespace sjhip {

constexpr int kFilterBlock = 256;
constexpr int kMaxLevels = 14;

__device__ __forceinline__ bool rect_ok(const float4& r) {
  return r.x <= r.z && r.y <= r.w;  // false for NaN
}

__device__ __forceinline__ bool hit(const float4& a, const float4& b) {
  return !(b.x > a.z || b.z < a.x || b.y > a.w || b.w < a.y);
}

struct TreeView {
  const float4* nodes;
  const float4* leaf_rects;
  const uint32_t* leaf_ids;
  int64_t n_valid;
  int n_levels;
  float coop_extent;  // see RectTree::coop_extent
  int64_t off[kMaxLevels];
  int64_t cnt[kMaxLevels];
};

// Keep policy of the plain filter: every rect hit is a candidate.
struct KeepAll {
  __device__ __forceinline__ uint32_t operator()(uint32_t hits, const uint32_t* ids, uint32_t qid) {
    (void)ids;
    (void)qid;
    return hits;
  }
  __device__ __forceinline__ void finish() {}
};

// Keep(hits, ids, qid) returns the subset of `hits` (bit k: build id
// ids[k] of the current leaf) that the pair with query row qid keeps: the
// plain filter keeps all, a fused filter+refine decides the predicate.
// Each lane works on its own copy; finish() is called by every lane at the
// end (statistics).
template <int F, int LEAF, class Keep>
__global__ __launch_bounds__(kFilterBlock) void k_filter(TreeView t, const float4* q, int64_t nq,
                                                         const uint32_t* qorder, int emit_rank, Keep keep,
                                                         uint32_t* ob, uint32_t* op, uint64_t cap,
                                                         unsigned long long* counter) {
  // per lane and level: the children of the level's current node still to
  // visit (the node itself is `cur`, its parent cur / F)
  __shared__ uint8_t s_mask[kMaxLevels + 1][kFilterBlock];
  __shared__ uint32_t s_ob[kFilterBlock / 64][kWaveOutCap], s_op[kFilterBlock / 64][kWaveOutCap];
  const int tid = threadIdx.x;
  WaveOut wo{s_ob[tid >> 6], s_op[tid >> 6], 0};
  Keep kp = keep;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  bool active = i < nq && t.n_levels > 0;
  uint32_t qid = 0, pid = 0;  // query row; probe id emitted (row, or position i in qorder)
  float4 box = make_float4(0, 0, 0, 0);
  const int top = t.n_levels;
  int lvl = top;
  uint32_t cur = 0;  // node at level lvl (level top: the virtual root)
  if (active) {
    qid = qorder ? qorder[i] : (uint32_t)i;
    pid = emit_rank ? (uint32_t)i : qid;
    box = q[qid];
    if (!rect_ok(box)) active = false;
  }
  if (active) {
    // virtual root: children are all nodes of level top-1 (<= F of them)
    uint32_t m = 0;
    const float4* nd = t.nodes + t.off[top - 1];
    int64_t nc = t.cnt[top - 1];
    for (int c = 0; c < F; c++)
      if (c < nc && hit(nd[c], box)) m |= 1u << c;
    s_mask[top][tid] = m;
  }
  while (__any(active)) {
    uint32_t hits = 0;
    int64_t hbase = 0;
    if (active) {
      while (lvl <= top && s_mask[lvl][tid] == 0) {
        lvl++;
        cur /= F;
      }
      if (lvl > top) {
        active = false;
      } else {
        uint32_t m = s_mask[lvl][tid];
        int c = __ffs(m) - 1;
        s_mask[lvl][tid] = m & (m - 1);
        int64_t child = (int64_t)cur * F + c;  // index at level lvl-1
        if (lvl - 1 == 0) {
          hbase = child * LEAF;
          int64_t e = min(hbase + LEAF, t.n_valid);
          const float4* lr = t.leaf_rects + hbase;
#pragma unroll
          for (int k = 0; k < LEAF; k++)
            if (hbase + k < e && hit(lr[k], box)) hits |= 1u << k;
          if (hits) hits = kp(hits, t.leaf_ids + hbase, qid);
        } else {
          int cl = lvl - 2;  // level of the child's children
          const float4* nd = t.nodes + t.off[cl];
          int64_t c0 = child * F;
          int64_t nc = t.cnt[cl] - c0;
          uint32_t mm = 0;
#pragma unroll
          for (int k = 0; k < F; k++)
            if (k < nc && hit(nd[c0 + k], box)) mm |= 1u << k;
          lvl = lvl - 1;
          cur = (uint32_t)child;
          s_mask[lvl][tid] = mm;
        }
      }
    }
    // wavefront compaction of this step's candidates into the LDS buffer
    if (__ballot(hits != 0)) wave_append_mask(wo, hits, t.leaf_ids, hbase, pid, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
  kp.finish();
}

// The walk of k_filter_wave for one wavefront with at least one active
// lane; the pairs go to wo (flushed by the caller at the end).
template <int F, int LEAF>
__device__ __forceinline__ void filter_wave_walk(const TreeView& t, bool active, const float4& box, const float4& ub,
                                                 uint32_t pid, WaveOut& wo, uint32_t* ob, uint32_t* op, uint64_t cap,
                                                 unsigned long long* counter) {
  const int lane = threadIdx.x & 63;
  // children still to visit, 8 bits per level: levels 0..7 in mlo, 8..15 in mhi
  uint64_t mlo = 0, mhi = 0;
  auto get_mask = [&](int l) -> uint32_t {
    return (uint32_t)((l < 8 ? mlo >> (8 * l) : mhi >> (8 * (l - 8))) & 0xffu);
  };
  auto set_mask = [&](int l, uint32_t m) {
    if (l < 8) mlo = (mlo & ~(0xffull << (8 * l))) | ((uint64_t)m << (8 * l));
    else mhi = (mhi & ~(0xffull << (8 * (l - 8)))) | ((uint64_t)m << (8 * (l - 8)));
  };
  const int top = t.n_levels;  // level top: the virtual root over the <= F nodes of level top-1
  {
    const int64_t nc = t.cnt[top - 1];
    const bool h = lane < F && lane < nc && hit(t.nodes[t.off[top - 1] + lane], ub);
    set_mask(top, (uint32_t)__ballot(h) & ((1u << F) - 1));
  }
  int lvl = top;
  int64_t cur = 0;  // node at level lvl
  for (;;) {
    while (lvl <= top && get_mask(lvl) == 0) {
      lvl++;
      cur /= F;
    }
    if (lvl > top) break;
    const uint32_t m = __builtin_amdgcn_readfirstlane(get_mask(lvl));
    const int c = __ffs(m) - 1;
    set_mask(lvl, m & (m - 1));
    const int64_t child = cur * F + c;  // index at level lvl-1
    if (lvl == 1) {
      const int64_t hbase = child * LEAF;
      const int64_t e = min(hbase + LEAF, t.n_valid);
      const float4* lr = t.leaf_rects + hbase;
      uint32_t hits = 0;
#pragma unroll
      for (int k = 0; k < LEAF; k++)
        if (hbase + k < e) {
          const float4 r = lr[k];
          if (active && hit(r, box)) hits |= 1u << k;
        }
      if (__ballot(hits != 0)) wave_append_mask(wo, hits, t.leaf_ids, hbase, pid, ob, op, cap, counter);
    } else {
      const int cl = lvl - 2;  // level of the child's children
      const int64_t c0 = child * F;
      const int64_t nc = t.cnt[cl] - c0;
      const bool h = lane < F && lane < nc && hit(t.nodes[t.off[cl] + c0 + lane], ub);
      lvl--;
      cur = child;
      set_mask(lvl, (uint32_t)__ballot(h) & ((1u << F) - 1));
    }
  }
}

// One walk per lane (the loop of k_filter with every hit kept), for a
// wavefront whose queries lie too far apart to share a walk.  mask: the
// block's per-level child masks, mask[l * blockDim.x + threadIdx.x].
template <int F, int LEAF>
__device__ __forceinline__ void filter_lane_walk(const TreeView& t, bool active, const float4& box, uint32_t pid,
                                                 uint8_t* mask, WaveOut& wo, uint32_t* ob, uint32_t* op,
                                                 uint64_t cap, unsigned long long* counter) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int top = t.n_levels;
  int lvl = top;
  int64_t cur = 0;
  if (active) {
    uint32_t m = 0;
    const float4* nd = t.nodes + t.off[top - 1];
    const int64_t nc = t.cnt[top - 1];
    for (int c = 0; c < F; c++)
      if (c < nc && hit(nd[c], box)) m |= 1u << c;
    mask[top * bs + tid] = (uint8_t)m;
  }
  while (__any(active)) {
    uint32_t hits = 0;
    int64_t hbase = 0;
    if (active) {
      while (lvl <= top && mask[lvl * bs + tid] == 0) {
        lvl++;
        cur /= F;
      }
      if (lvl > top) {
        active = false;
      } else {
        const uint32_t m = mask[lvl * bs + tid];
        const int c = __ffs(m) - 1;
        mask[lvl * bs + tid] = (uint8_t)(m & (m - 1));
        const int64_t child = cur * F + c;
        if (lvl == 1) {
          hbase = child * LEAF;
          const int64_t e = min(hbase + LEAF, t.n_valid);
          const float4* lr = t.leaf_rects + hbase;
#pragma unroll
          for (int k = 0; k < LEAF; k++)
            if (hbase + k < e && hit(lr[k], box)) hits |= 1u << k;
        } else {
          const int cl = lvl - 2;
          const float4* nd = t.nodes + t.off[cl];
          const int64_t c0 = child * F;
          const int64_t nc = t.cnt[cl] - c0;
          uint32_t mm = 0;
#pragma unroll
          for (int k = 0; k < F; k++)
            if (k < nc && hit(nd[c0 + k], box)) mm |= 1u << k;
          lvl--;
          cur = child;
          mask[lvl * bs + tid] = (uint8_t)mm;
        }
      }
    }
    if (__ballot(hits != 0)) wave_append_mask(wo, hits, t.leaf_ids, hbase, pid, ob, op, cap, counter);
  }
}

// Wavefront-cooperative filter for queries visited in a locality order
// (qorder: Morton order of the query rects).  The wavefront walks the
// tree once with the union of its (up to) 64 query rects: the children of
// a node are tested by F lanes at once and the walk state is one ballot
// mask per level, uniform across the wavefront.  At each leaf reached,
// every lane tests its own rect against the leaf's rects (uniform loads).
// A rect that meets a lane's query also meets the union, so the pairs are
// exactly those of k_filter (one walk per lane); they come out leaf by
// leaf, lanes in order.  A wavefront whose union is wider than
// t.coop_extent (its queries lie far apart, as in a sparse probe) walks
// one lane per query instead.
//
// The pairs a wavefront still holds at the end are written with one
// atomic per block (kFilterWaveBlock threads): a single output counter
// takes one same-address atomic per flush, and one per wavefront would be
// about one per 64 queries.
constexpr int kFilterWaveBlock = 512;

template <int F, int LEAF>
__global__ __launch_bounds__(kFilterWaveBlock) void k_filter_wave(TreeView t, const float4* q, int64_t nq,
                                                                  const uint32_t* qorder, int emit_rank,
                                                                  uint32_t* ob, uint32_t* op, uint64_t cap,
                                                                  unsigned long long* counter) {
  static_assert(F <= 8 && LEAF <= 32, "fan-out of at most 8 (8-bit masks), leaves of at most 32 rects");
  constexpr int kWaves = kFilterWaveBlock / 64;
  __shared__ uint32_t s_ob[kWaves][kWaveOutCap], s_op[kWaves][kWaveOutCap];
  __shared__ uint32_t s_n[kWaves];
  __shared__ unsigned long long s_base;
  __shared__ uint8_t s_mask[(kMaxLevels + 1) * kFilterWaveBlock];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  WaveOut wo{s_ob[wave], s_op[wave], 0};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  bool active = i < nq && t.n_levels > 0;
  uint32_t pid = 0;
  float4 box = make_float4(0, 0, 0, 0);
  if (active) {
    const uint32_t qid = qorder ? qorder[i] : (uint32_t)i;
    pid = emit_rank ? (uint32_t)i : qid;
    box = q[qid];
    if (!rect_ok(box)) active = false;
  }
  if (__any(active)) {
    // union of the active lanes' rects
    float ux0 = active ? box.x : INFINITY, uy0 = active ? box.y : INFINITY;
    float ux1 = active ? box.z : -INFINITY, uy1 = active ? box.w : -INFINITY;
    for (int o = 32; o > 0; o >>= 1) {
      ux0 = fminf(ux0, __shfl_xor(ux0, o));
      uy0 = fminf(uy0, __shfl_xor(uy0, o));
      ux1 = fmaxf(ux1, __shfl_xor(ux1, o));
      uy1 = fmaxf(uy1, __shfl_xor(uy1, o));
    }
    const float4 ub = make_float4(ux0, uy0, ux1, uy1);
    if (fmaxf(ux1 - ux0, uy1 - uy0) <= t.coop_extent)  // wavefront-uniform
      filter_wave_walk<F, LEAF>(t, active, box, ub, pid, wo, ob, op, cap, counter);
    else
      filter_lane_walk<F, LEAF>(t, active, box, pid, s_mask, wo, ob, op, cap, counter);
  }
  // block flush: every thread of the block reaches the barriers
  if (lane == 0) s_n[wave] = wo.n;
  __syncthreads();
  if (tid == 0) {
    unsigned long long tot = 0;
    for (int k = 0; k < kWaves; k++) tot += s_n[k];
    s_base = tot ? atomicAdd(counter, tot) : 0ull;
  }
  __syncthreads();
  unsigned long long base = s_base;
  for (int k = 0; k < wave; k++) base += s_n[k];
  for (uint32_t k = lane; k < wo.n; k += 64) {
    const unsigned long long d = base + k;
    if (d < cap) {
      ob[d] = wo.sb[k];
      op[d] = wo.sp[k];
    }
  }
}

// ---------------------------------------------------------------------------
// Grid filter (RectTree's cell grid): one lane per query in the given
// order; a query visits the cells its rect covers and tests their rects.
// A pair found in several cells is kept in one: the cell holding the
// corner (max of the two min x, max of the two min y) of the two rects'
// overlap, which lies in both rects and so in a cell listing the build
// rect.  (A point query covers one cell.)  Each lane counts its pairs,
// the block scans the counts and takes one output range with one atomic,
// then each lane writes its pairs contiguously, so they come out in query
// order.  The same inclusive FP32 rect test as the tree walk decides a
// hit.

struct GridView {
  const uint32_t* off;
  const float4* rects;
  const uint32_t* ids;
  int nx, ny;
  float x0, y0, inv_x, inv_y;
};

__device__ __forceinline__ int grid_cell(float v, float v0, float inv, int n) {
  float f = (v - v0) * inv;
  f = fminf(fmaxf(f, 0.f), (float)(n - 1));
  return (int)f;
}

// Counts (WRITE = false) or writes the pairs of one query.  For a query
// in one cell the counting pass records which of the cell's first 64
// rects it hits in *mask, and the writing pass then reads only their ids.
template <bool WRITE>
__device__ __forceinline__ uint32_t grid_visit(const GridView& g, const float4& box, uint32_t pid, uint32_t* ob,
                                               uint32_t* op, uint64_t base, uint64_t cap, uint64_t* mask) {
  const int cx0 = grid_cell(box.x, g.x0, g.inv_x, g.nx), cx1 = grid_cell(box.z, g.x0, g.inv_x, g.nx);
  const int cy0 = grid_cell(box.y, g.y0, g.inv_y, g.ny), cy1 = grid_cell(box.w, g.y0, g.inv_y, g.ny);
  const bool one = cx0 == cx1 && cy0 == cy1;
  uint32_t n = 0;
  if (one) {
    const int64_t cell = (int64_t)cy0 * g.nx + cx0;
    const uint32_t k0 = g.off[cell], k1 = g.off[cell + 1];
    if (k1 - k0 <= 64) {
      if (!WRITE) {
        uint64_t m = 0;
        for (uint32_t k = k0; k < k1; k++)
          if (hit(g.rects[k], box)) m |= 1ull << (k - k0);
        *mask = m;
        return (uint32_t)__popcll(m);
      }
      for (uint64_t m = *mask; m; m &= m - 1) {
        const uint32_t k = k0 + (uint32_t)(__ffsll((unsigned long long)m) - 1);
        const uint64_t d = base + n;
        if (d < cap) {
          ob[d] = g.ids[k];
          op[d] = pid;
        }
        n++;
      }
      return n;
    }
  }
  for (int cy = cy0; cy <= cy1; cy++)
    for (int cx = cx0; cx <= cx1; cx++) {
      const int64_t cell = (int64_t)cy * g.nx + cx;
      const uint32_t k0 = g.off[cell], k1 = g.off[cell + 1];
      for (uint32_t k = k0; k < k1; k++) {
        const float4 r = g.rects[k];
        if (!hit(r, box)) continue;
        if (!one && (grid_cell(fmaxf(r.x, box.x), g.x0, g.inv_x, g.nx) != cx ||
                     grid_cell(fmaxf(r.y, box.y), g.y0, g.inv_y, g.ny) != cy))
          continue;
        if (WRITE) {
          const uint64_t d = base + n;
          if (d < cap) {
            ob[d] = g.ids[k];
            op[d] = pid;
          }
        }
        n++;
      }
    }
  return n;
}

constexpr int kFilterGridBlock = 512;

// Queries are FP32 rects q[qorder[i]] (pid: i or the row), or, with
// xy, FP64 points already in visiting order whose rects are the points
// rounded to nearest FP32 (the operator's point rects; pid = i).
static __global__ __launch_bounds__(kFilterGridBlock) void k_filter_grid(GridView g, const float4* q,
                                                                         const double2* xy, int64_t nq,
                                                                         const uint32_t* qorder, int emit_rank,
                                                                         uint32_t* ob, uint32_t* op, uint64_t cap,
                                                                         unsigned long long* counter) {
  constexpr int kWaves = kFilterGridBlock / 64;
  __shared__ uint32_t s_wave[kWaves];
  __shared__ unsigned long long s_base;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  bool active = i < nq;
  uint32_t pid = 0;
  float4 box = make_float4(0, 0, 0, 0);
  if (active) {
    if (xy) {
      const double2 p = xy[i];
      const float fx = (float)p.x, fy = (float)p.y;
      box = make_float4(fx, fy, fx, fy);
      pid = (uint32_t)i;
    } else {
      const uint32_t qid = qorder ? qorder[i] : (uint32_t)i;
      pid = emit_rank ? (uint32_t)i : qid;
      box = q[qid];
    }
    active = rect_ok(box);
  }
  uint64_t mask = 0;
  const uint32_t n = active ? grid_visit<false>(g, box, pid, ob, op, 0, 0, &mask) : 0u;
  // exclusive prefix of n within the wavefront, then across wavefronts
  uint32_t incl = n;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  if (tid == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kWaves; w++) {
      const uint32_t c = s_wave[w];
      s_wave[w] = (uint32_t)tot;
      tot += c;
    }
    s_base = tot ? atomicAdd(counter, tot) : 0ull;
  }
  __syncthreads();
  if (active && n) grid_visit<true>(g, box, pid, ob, op, s_base + s_wave[wave] + (incl - n), cap, &mask);
}

inline GridView grid_view(const RectTree& t) {
  GridView g{};
  g.off = t.grid_off.p;
  g.rects = t.grid_rects.p;
  g.ids = t.grid_ids.p;
  g.nx = t.grid_nx;
  g.ny = t.grid_ny;
  g.x0 = t.grid_x0;
  g.y0 = t.grid_y0;
  g.inv_x = t.grid_inv_x;
  g.inv_y = t.grid_inv_y;
  return g;
}

inline TreeView tree_view(const RectTree& t) {
  TreeView v{};
  v.nodes = t.nodes.p;
  v.leaf_rects = t.leaf_rects.p;
  v.leaf_ids = t.leaf_ids.p;
  v.n_valid = t.n_valid;
  v.n_levels = t.n_levels;
  v.coop_extent = t.coop_extent;
  for (int l = 0; l < t.n_levels; l++) {
    v.off[l] = t.level_off[l];
    v.cnt[l] = t.level_cnt[l];
  }
  return v;
}

}  // namespace sjhip
