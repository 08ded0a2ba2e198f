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

inline TreeView tree_view(const RectTree& t) {
  TreeView v{};
  v.nodes = t.nodes.p;
  v.leaf_rects = t.leaf_rects.p;
  v.leaf_ids = t.leaf_ids.p;
  v.n_valid = t.n_valid;
  v.n_levels = t.n_levels;
  for (int l = 0; l < t.n_levels; l++) {
    v.off[l] = t.level_off[l];
    v.cnt[l] = t.level_cnt[l];
  }
  return v;
}

}  // namespace sjhip
