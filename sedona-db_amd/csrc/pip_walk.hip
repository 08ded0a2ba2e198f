// pip_walk.hip -- the certified FP32 walk of point-in-polygon candidates
// whose build row is one polygon with one ring (a PolyDesc record).
//
// The classification is tg_ring_contains_point (c/sedona-tg/src/tg/
// tg.c:2809-2822) as restated in ring_walk.h, in the deferred mode: a
// candidate with any uncertain FP32 test is not decided here but listed
// for the FP64 pass along +x, the reference's own ray.
//
// One lane per candidate.  Any ray from a point off the boundary of a
// closed ring crosses it an odd number of times exactly when the point is
// inside (mirroring and transposing the coordinates are exact, so the
// certified test is the same test on other axes): each lane casts its ray
// toward the nearest side of its ring's rect, where fewer segments and
// tree nodes lie on it.  (A direction shared by the wavefront, or one walk
// per direction present, gave longer rays or serialised walks and was
// slower on the 100M x 1M star join.)
//
// The kernel is kept lean for occupancy: the walk state is a few
// registers, vertex loads are batched (the walk waits on memory more than
// on arithmetic), and the cell-raster read is compiled out when a raster
// pass has already decided the candidates in clean cells.
//
// Rows that are not single rings (multipolygons, polygons with holes) are
// listed in `rest` for the general kernel (or in `redo` when no rest list
// is given).
#include "ring_walk.h"
#include "wave_out.h"

namespace sjhip {

namespace {

constexpr int kWalkBlock = 256;
// candidates a wavefront takes from the queue at a time (64 per step): one
// counter shared by the grid takes one atomic per chunk
constexpr uint32_t kQueueChunk = 1024;
constexpr int kWalkBoxBatch = 4;  // box loads in flight (more raise the register count)
#ifndef SJ_SEG_BATCH
#define SJ_SEG_BATCH 6
#endif
constexpr int kSegBatch = SJ_SEG_BATCH;  // vertex loads in flight

// A lane's walk: ray direction dir (bit 1 set: along the y axis, bit 0
// set: toward minus), the point's coordinates along and across the ray
// (ring-local FP32, the along one mirrored so that the ray runs toward
// + along), and the result st (bit 0: crossing parity, bits 1..: the
// number of uncertain tests).
struct LaneWalk {
  const float4* box;  // the ring's tree, level 0 first (nullptr: none)
  const float2* q32;  // the ring's FP32 ring-local vertices
  float pa, pc;
  float E;            // certification bound (boxes are widened by 2E)
  int ns;
  int dir;
  uint32_t st;
};

__device__ __forceinline__ float along(int dir, float x, float y) {
  const float a = (dir & 2) ? y : x;
  return (dir & 1) ? -a : a;
}
__device__ __forceinline__ float across(int dir, float x, float y) { return (dir & 2) ? x : y; }

// Whether a segment in box b can meet the ray: the point's across
// coordinate within the box's across range and the box's along range not
// entirely behind the point (both widened by the margin m).
__device__ __forceinline__ bool box_hit(const LaneWalk& w, float m, const float4& b) {
  const bool sw = (w.dir & 2) != 0;
  const float c0 = sw ? b.x : b.y, c1 = sw ? b.z : b.w;
  const float a1 = (w.dir & 1) ? -(sw ? b.y : b.x) : (sw ? b.w : b.z);
  return w.pc >= c0 - m && w.pc <= c1 + m && w.pa <= a1 + m;
}

__device__ __forceinline__ void segment_test(LaneWalk& w, uint32_t& tests, float2 a, float2 b) {
  const int rc = raycast_f32_certified(along(w.dir, a.x, a.y), across(w.dir, a.x, a.y), along(w.dir, b.x, b.y),
                                       across(w.dir, b.x, b.y), w.pa, w.pc, w.E);
  if (rc == kRayOut) return;
  tests++;
  w.st ^= rc < 0 ? 0u : 1u;  // certified: kRayIn
  w.st += rc < 0 ? 2u : 0u;
}

// Segments [s0, s1), kSegBatch vertex loads issued before the first test.
// The batch loop itself is not unrolled, which keeps the code small.
__device__ __forceinline__ void seg_scan(LaneWalk& w, uint32_t& tests, const int s0, const int s1) {
  float2 a = w.q32[s0];
  for (int c0 = s0; c0 < s1; c0 += kSegBatch) {
    float2 v[kSegBatch];
#pragma unroll
    for (int k = 0; k < kSegBatch; k++) v[k] = c0 + k + 1 <= s1 ? w.q32[c0 + k + 1] : a;
    segment_test(w, tests, a, v[0]);
#pragma unroll
    for (int k = 1; k < kSegBatch; k++)
      if (c0 + k < s1) segment_test(w, tests, v[k - 1], v[k]);
    a = v[kSegBatch - 1];
  }
}

__device__ __forceinline__ uint32_t hit_mask(const LaneWalk& w, const float4* b, int n) {
  const float m = 2.0f * w.E;
  uint32_t mask = 0;
#pragma unroll
  for (int c0 = 0; c0 < kSegFanout; c0 += kWalkBoxBatch) {
    float4 bx[kWalkBoxBatch];
#pragma unroll
    for (int k = 0; k < kWalkBoxBatch; k++) bx[k] = c0 + k < n ? b[c0 + k] : make_float4(1.f, 1.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < kWalkBoxBatch; k++)
      if (c0 + k < n && box_hit(w, m, bx[k])) mask |= 1u << (c0 + k);
  }
  return mask;
}

// Depth-first walk over the ring's segment tree (leaves of 1 << ls
// segments), the layout walk_tree uses: byte l of `masks` holds the
// children of level l still to visit.
__device__ __forceinline__ void walk_ring(LaneWalk& w, int ls, uint32_t& tests) {
  if (!w.box) {
    seg_scan(w, tests, 0, w.ns);
    return;
  }
  const int top = seg_tree_shape(w.ns, ls).levels - 1;
  int otop, ntop;
  level_geom(w.ns, ls, top, otop, ntop);
  const float4* btop = w.box + otop;
  for (int c0 = 0; c0 < ntop; c0 += kSegFanout) {
    uint64_t masks = (uint64_t)hit_mask(w, btop + c0, min(kSegFanout, ntop - c0)) << (8 * top);
    int lvl = top, base = c0;
    while (masks) {
      const uint32_t m = (uint32_t)(masks >> (8 * lvl)) & 0xffu;
      if (m == 0) {
        lvl++;
        base = (base >> 3) & ~(kSegFanout - 1);
        continue;
      }
      const int k = __ffs(m) - 1;
      masks &= ~(1ull << (8 * lvl + k));
      const int node = base + k;
      if (lvl == 0) {
        const int s0 = node << ls;
        seg_scan(w, tests, s0, min(s0 + (1 << ls), w.ns));
        continue;
      }
      const int cb = node * kSegFanout;
      int loff, lcnt;
      level_geom(w.ns, ls, lvl - 1, loff, lcnt);
      const uint32_t cm = hit_mask(w, w.box + loff + cb, min(kSegFanout, lcnt - cb));
      if (cm) {
        lvl--;
        base = cb;
        masks |= (uint64_t)cm << (8 * lvl);
      }
    }
  }
}

// Appends v to list (count *cnt) for every lane with f set: one atomic
// per wavefront.  Called by all lanes of the wavefront.
__device__ __forceinline__ void wave_push(bool f, uint32_t v, uint32_t* list, unsigned long long* cnt) {
  const uint64_t m = __ballot(f);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(cnt, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  if (f) list[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
}

// RASTER: whether to read the candidates' cell rasters (false when a raster
// pass has already decided every candidate in a clean cell).
template <int WAVES, bool RASTER>
__global__ __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_walk_rings(
    RingIndexView R, const double2* __restrict__ pts, const uint32_t* __restrict__ cb,
    const uint32_t* __restrict__ cp, uint64_t n, const unsigned long long* n_dev, const uint32_t* probe_ids,
    uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter,
    unsigned long long* stats, uint32_t* redo, unsigned long long* n_redo, uint32_t* rest,
    unsigned long long* n_rest, unsigned long long* queue) {
  __shared__ uint32_t wo_b[kWalkBlock / 64][kWaveOutCap], wo_p[kWalkBlock / 64][kWaveOutCap];
  WaveOut wo{wo_b[threadIdx.x >> 6], wo_p[threadIdx.x >> 6], 0};
  if (n_dev) n = min(n, (uint64_t)*n_dev);  // candidate count produced on the device (n: its bound)
  const int ls = R.seg_leaf_shift;
  const uint32_t n32 = (uint32_t)n;  // (the launcher admits at most 2^32 - 1 candidates)
  const int lane = threadIdx.x & 63;
  // Work distribution: with a queue, each wavefront takes the next 64
  // candidates from the shared counter whenever it finishes its last ones
  // (walks differ in length by orders of magnitude); without, a grid-stride
  // loop over the candidates.
  uint32_t chunk_end = 0;  // end of the wavefront's current chunk (queue mode)
  auto next = [&](uint32_t cur) -> uint32_t {
    if (!queue) return cur + gridDim.x * kWalkBlock;
    if (cur + 64 < chunk_end) return cur + 64;
    unsigned long long v = 0;
    if (lane == 0) v = atomicAdd(queue, (unsigned long long)kQueueChunk);
    v = __shfl(v, 0);
    chunk_end = v + kQueueChunk < n32 ? (uint32_t)(v + kQueueChunk) : n32;
    return v < n32 ? (uint32_t)v : n32;
  };
  uint32_t tests = 0, unsure = 0;
  for (uint32_t wbase = queue ? next(0) : blockIdx.x * kWalkBlock + (threadIdx.x & ~63u); wbase < n32;
       wbase = next(wbase)) {
    const uint32_t i = wbase + lane;
    bool keep = false, walk = false, general = false;
    LaneWalk w;
    if (i < n32) {
      const uint32_t b = cb[i];
      // the probe is a point column: null and empty rows hold NaN
      const double2 pt = pts[cp[i]];
      if (pt.x == pt.x || pt.y == pt.y) {
        const PolyDesc d = R.desc[b];
        if (d.npts < 0) {
          general = true;
        } else {
          const int cls = RASTER ? raster_class(R, d, pt) : (rect_covers(d.rect, pt.x, pt.y) ? 0 : 2);
          keep = cls == 1;
          if (cls == 0 && d.npts >= 3) {
            walk = true;
            const double2 o = ring_origin(d.rect);
            const float xl = (float)(pt.x - o.x), yl = (float)(pt.y - o.y);
            const double dl = pt.x - d.rect.x, dr = d.rect.z - pt.x, db = pt.y - d.rect.y, dt = d.rect.w - pt.y;
            w.dir = fmin(db, dt) < fmin(dl, dr) ? (db < dt ? 3 : 2) : (dl < dr ? 1 : 0);
            w.box = d.tree_off >= 0 ? R.seg_tree + d.tree_off : nullptr;
            w.q32 = R.xy32 + d.v0;
            w.E = d.err;
            w.ns = d.npts - 1;
            w.st = 0;
            w.pa = along(w.dir, xl, yl);
            w.pc = across(w.dir, xl, yl);
          }
        }
      }
    }
    if (walk) {
      walk_ring(w, ls, tests);
      keep = w.st == 1u;  // odd crossings, no uncertain test
      unsure += w.st >> 1;
    }
    const bool deferred = walk && w.st > 1u;
    wave_push(deferred || (general && !rest), i, redo, n_redo);
    if (rest) wave_push(general, i, rest, n_rest);
    // (the ids are read again here rather than held through the walk)
    uint32_t b = 0, p = 0;
    if (keep) {
      b = cb[i];
      p = cp[i];
      if (probe_ids) p = probe_ids[p];
    }
    wave_append(wo, keep, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
  for (int o = 32; o > 0; o >>= 1) {
    tests += __shfl_xor(tests, o);
    unsure += __shfl_xor(unsure, o);
  }
  if ((threadIdx.x & 63) == 0 && stats && (tests | unsure)) {
    atomicAdd(&stats[0], (unsigned long long)tests);
    atomicAdd(&stats[1], (unsigned long long)unsure);
  }
}

// ---- ST_DWithin over single-line build rows ---------------------------------
//
// The decision dist(line, point) <= d of rust/sedona-spatial-join/src/
// refine/geo.rs:312-335 (a minimum over the line's segments, so <= d exactly
// when one segment term is) with certified FP32 point-segment terms
// (segment_within_f32) only: a lane stops at its first segment that is
// certainly within d; a candidate with no such segment but an uncertain one
// is listed for the FP64 kernel, which decides it in the reference's
// operation order.  Tree boxes, and segments of rows without a tree, are
// pruned by the query square [p - D, p + D] with D = d (1 + 1e-4) + 4E,
// the bound the FP64 kernel prunes with.

// resident waves per SIMD of the line walk (at 6 it holds its state in 79
// VGPRs without scratch; at 8 it spills)
#ifndef SJ_LW_WAVES
#define SJ_LW_WAVES 6
#endif

struct LineWalk {
  const float4* box;  // the line's tree, level 0 first (nullptr: none)
  const float2* q32;  // the line's FP32 line-local vertices
  float px, py;       // the point, line-local FP32
  float D;            // pruning half-width of the query square
  float E, Ep;        // bounds of the vertices' and the point's FP32 rounding
  int ns;
  uint32_t st;        // bit 0: a segment certainly within d; bits 1..: uncertain tests
};

__device__ __forceinline__ bool line_box_hit(const LineWalk& w, const float4& b) {
  return b.x <= w.px + w.D && b.z >= w.px - w.D && b.y <= w.py + w.D && b.w >= w.py - w.D;
}

__device__ __forceinline__ void line_segment_test(LineWalk& w, uint32_t& tests, float2 a, float2 b, double d) {
  const float gx = fmaxf(fminf(a.x, b.x) - w.px, w.px - fmaxf(a.x, b.x));
  const float gy = fmaxf(fminf(a.y, b.y) - w.py, w.py - fmaxf(a.y, b.y));
  if (fmaxf(gx, gy) > w.D) return;
  tests++;
  const int r = segment_within_f32(a.x, a.y, b.x, b.y, w.px, w.py, w.E, w.Ep, d);
  w.st |= r == 1 ? 1u : 0u;
  w.st += r < 0 ? 2u : 0u;
}

// Segments [s0, s1), kSegBatch vertex loads in flight; stops once a
// segment is certainly within d.
__device__ __forceinline__ void line_seg_scan(LineWalk& w, uint32_t& tests, const int s0, const int s1, double d) {
  float2 a = w.q32[s0];
  for (int c0 = s0; c0 < s1 && !(w.st & 1u); c0 += kSegBatch) {
    float2 v[kSegBatch];
#pragma unroll
    for (int k = 0; k < kSegBatch; k++) v[k] = c0 + k + 1 <= s1 ? w.q32[c0 + k + 1] : a;
    line_segment_test(w, tests, a, v[0], d);
#pragma unroll
    for (int k = 1; k < kSegBatch; k++)
      if (c0 + k < s1) line_segment_test(w, tests, v[k - 1], v[k], d);
    a = v[kSegBatch - 1];
  }
}

__device__ __forceinline__ uint32_t line_hit_mask(const LineWalk& w, const float4* b, int n) {
  uint32_t mask = 0;
#pragma unroll
  for (int c0 = 0; c0 < kSegFanout; c0 += kWalkBoxBatch) {
    float4 bx[kWalkBoxBatch];
#pragma unroll
    for (int k = 0; k < kWalkBoxBatch; k++) bx[k] = c0 + k < n ? b[c0 + k] : make_float4(1.f, 1.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < kWalkBoxBatch; k++)
      if (c0 + k < n && line_box_hit(w, bx[k])) mask |= 1u << (c0 + k);
  }
  return mask;
}

// Depth-first walk over the line's segment tree (the layout of walk_ring),
// ending at the first segment certainly within d.
__device__ __forceinline__ void walk_line(LineWalk& w, int ls, uint32_t& tests, double d) {
  if (!w.box) {
    line_seg_scan(w, tests, 0, w.ns, d);
    return;
  }
  const int top = seg_tree_shape(w.ns, ls).levels - 1;
  int otop, ntop;
  level_geom(w.ns, ls, top, otop, ntop);
  const float4* btop = w.box + otop;
  for (int c0 = 0; c0 < ntop && !(w.st & 1u); c0 += kSegFanout) {
    uint64_t masks = (uint64_t)line_hit_mask(w, btop + c0, min(kSegFanout, ntop - c0)) << (8 * top);
    int lvl = top, base = c0;
    while (masks && !(w.st & 1u)) {
      const uint32_t m = (uint32_t)(masks >> (8 * lvl)) & 0xffu;
      if (m == 0) {
        lvl++;
        base = (base >> 3) & ~(kSegFanout - 1);
        continue;
      }
      const int k = __ffs(m) - 1;
      masks &= ~(1ull << (8 * lvl + k));
      const int node = base + k;
      if (lvl == 0) {
        const int s0 = node << ls;
        line_seg_scan(w, tests, s0, min(s0 + (1 << ls), w.ns), d);
        continue;
      }
      const int cb = node * kSegFanout;
      int loff, lcnt;
      level_geom(w.ns, ls, lvl - 1, loff, lcnt);
      const uint32_t cm = line_hit_mask(w, w.box + loff + cb, min(kSegFanout, lcnt - cb));
      if (cm) {
        lvl--;
        base = cb;
        masks |= (uint64_t)cm << (8 * lvl);
      }
    }
  }
}

// Appends the pair (b, p) to lists (lb, lp) (count *cnt) for every lane
// with f set: one atomic per wavefront.  Called by all lanes.
__device__ __forceinline__ void wave_push_pair(bool f, uint32_t b, uint32_t p, uint32_t* lb, uint32_t* lp,
                                               unsigned long long* cnt) {
  const uint64_t m = __ballot(f);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(cnt, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  if (f) {
    const unsigned long long k = base + __popcll(m & ((1ull << lane) - 1ull));
    lb[k] = b;
    lp[k] = p;
  }
}

// Candidates whose build row is not a single line (POLY: a single-ring
// polygon) with an FP32 frame, and those with an uncertain test and no
// certain one, go to `redo` for the FP64 kernel.  POLY: the ring's
// segments are walked as the line's are (a single-ring polygon is within d
// of the point when a segment is, or when it holds the point); the
// candidates with every segment certainly farther than d go, as pairs, to
// (pb, pp) (count *n_p) for the point-in-polygon walk.
template <int WAVES, bool POLY>
__global__ __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_walk_lines(
    RingIndexView R, const double2* __restrict__ pts, const uint32_t* __restrict__ cb,
    const uint32_t* __restrict__ cp, uint64_t n, const unsigned long long* n_dev, const uint32_t* probe_ids,
    double dist, uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter, unsigned long long* stats,
    uint32_t* redo, unsigned long long* n_redo, unsigned long long* queue, uint32_t* pb, uint32_t* pp,
    unsigned long long* n_p) {
  __shared__ uint32_t wo_b[kWalkBlock / 64][kWaveOutCap], wo_p[kWalkBlock / 64][kWaveOutCap];
  WaveOut wo{wo_b[threadIdx.x >> 6], wo_p[threadIdx.x >> 6], 0};
  if (n_dev) n = min(n, (uint64_t)*n_dev);  // candidate count produced on the device (n: its bound)
  const int ls = R.seg_leaf_shift;
  const uint32_t n32 = (uint32_t)n;  // (the launcher admits at most 2^32 - 65 candidates)
  const int lane = threadIdx.x & 63;
  // the FP64 kernel's pruning bound, without E
  const float Dd = (float)(dist * 1.0001);
  uint32_t chunk_end = 0;
  auto next = [&](uint32_t cur) -> uint32_t {
    if (cur + 64 < chunk_end) return cur + 64;
    unsigned long long v = 0;
    if (lane == 0) v = atomicAdd(queue, (unsigned long long)kQueueChunk);
    v = __shfl(v, 0);
    chunk_end = v + kQueueChunk < n32 ? (uint32_t)(v + kQueueChunk) : n32;
    return v < n32 ? (uint32_t)v : n32;
  };
  uint32_t tests = 0, unsure = 0;
  for (uint32_t wbase = next(0); wbase < n32; wbase = next(wbase)) {
    const uint32_t i = wbase + lane;
    bool keep = false, walk = false, defer = false, inside_test = false;
    LineWalk w;
    if (i < n32) {
      const uint32_t b = cb[i];
      // the probe is a point column: null and empty rows hold NaN
      const double2 pt = pts[cp[i]];
      if (pt.x == pt.x || pt.y == pt.y) {
        const PolyDesc dsc = R.desc[b];
        if (dsc.npts < (POLY ? 4 : 2) || !(dsc.err < INFINITY)) {
          // other rows, lines of fewer than 2 vertices (distance 0), rings
          // of fewer than 4 and rows without an FP32 frame: the FP64 kernel
          defer = true;
        } else {
          walk = true;
          const double2 o = ring_origin(dsc.rect);
          w.px = (float)(pt.x - o.x);
          w.py = (float)(pt.y - o.y);
          w.E = dsc.err;
          w.D = Dd + 4.0f * dsc.err;
          w.Ep = fmaxf(fabsf(w.px), fabsf(w.py)) * 5.9604645e-08f * 1.0001f;
          w.box = dsc.tree_off >= 0 ? R.seg_tree + dsc.tree_off : nullptr;
          w.q32 = R.xy32 + dsc.v0;
          w.ns = dsc.npts - 1;
          w.st = 0;
        }
      }
    }
    if (walk) {
      walk_line(w, ls, tests, dist);
      keep = (w.st & 1u) != 0;
      defer = !keep && w.st > 1u;
      inside_test = POLY && !keep && !defer;
      unsure += w.st >> 1;
    }
    wave_push(defer, i, redo, n_redo);
    if (POLY) wave_push_pair(inside_test, inside_test ? cb[i] : 0u, inside_test ? cp[i] : 0u, pb, pp, n_p);
    uint32_t b = 0, p = 0;
    if (keep) {
      b = cb[i];
      p = cp[i];
      if (probe_ids) p = probe_ids[p];
    }
    wave_append(wo, keep, b, p, ob, op, cap, counter);
  }
  wave_flush(wo, ob, op, cap, counter);
  for (int o = 32; o > 0; o >>= 1) {
    tests += __shfl_xor(tests, o);
    unsure += __shfl_xor(unsure, o);
  }
  if ((threadIdx.x & 63) == 0 && stats && (tests | unsure)) {
    atomicAdd(&stats[0], (unsigned long long)tests);
    atomicAdd(&stats[1], (unsigned long long)unsure);
  }
}

}  // namespace

void launch_line_walk(const RingIndex& ri, const GeomColumn& probe, const uint32_t* cb, const uint32_t* cp,
                      uint64_t n, const unsigned long long* n_dev, const uint32_t* probe_ids, double d, uint32_t* ob,
                      uint32_t* op, uint64_t cap, unsigned long long* counter, unsigned long long* stats,
                      uint32_t* redo, unsigned long long* n_redo, hipStream_t s, unsigned long long* queue,
                      uint32_t* pb, uint32_t* pp, unsigned long long* n_p) {
  if (n == 0) return;
  if (!probe.points_only) throw Error("launch_line_walk: the probe column must hold points only");
  if (ri.desc.p == nullptr) throw Error("launch_line_walk: the ring index has no row records");
  if (!queue) throw Error("launch_line_walk: needs a queue counter");
  if (n >= 0xffffffffull - 64) throw Error("launch_line_walk: 2^32 - 64 candidates or more");
  const uint64_t blocks = (n + kWalkBlock - 1) / kWalkBlock;
  int dev = 0, cus = 0;
  SJ_CHECK(hipGetDevice(&dev));
  SJ_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // the blocks that can be resident at once
  const unsigned g =
      (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)cus * 4 * SJ_LW_WAVES / (kWalkBlock / 64)));
  if (pb) {
    if (!pp || !n_p) throw Error("launch_line_walk: the polygon form needs its pair list");
    hipLaunchKernelGGL((k_walk_lines<SJ_LW_WAVES, true>), g, kWalkBlock, 0, s, ri.view(), probe.xy.p, cb, cp, n,
                       n_dev, probe_ids, d, ob, op, cap, counter, stats, redo, n_redo, queue, pb, pp, n_p);
  } else {
    hipLaunchKernelGGL((k_walk_lines<SJ_LW_WAVES, false>), g, kWalkBlock, 0, s, ri.view(), probe.xy.p, cb, cp, n,
                       n_dev, probe_ids, d, ob, op, cap, counter, stats, redo, n_redo, queue, nullptr, nullptr,
                       nullptr);
  }
  SJ_CHECK(hipGetLastError());
}

void launch_ring_walk(const RingIndex& ri, const GeomColumn& probe, const uint32_t* cb, const uint32_t* cp,
                      uint64_t n, const unsigned long long* n_dev, const uint32_t* probe_ids, bool raster_checked,
                      uint32_t* ob, uint32_t* op, uint64_t cap, unsigned long long* counter,
                      unsigned long long* stats, uint32_t* redo, unsigned long long* n_redo, uint32_t* rest,
                      unsigned long long* n_rest, int waves, hipStream_t s, unsigned long long* queue) {
  if (n == 0) return;
  if (!probe.points_only) throw Error("launch_ring_walk: the probe column must hold points only");
  if (ri.desc.p == nullptr) throw Error("launch_ring_walk: the ring index has no row records");
  if (n >= 0xffffffffull - 64) throw Error("launch_ring_walk: 2^32 - 64 candidates or more");
  const uint64_t blocks = (n + kWalkBlock - 1) / kWalkBlock;
  unsigned g = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 1u << 16));
  if (queue) {
    // a queue launch needs only the blocks that can be resident at once
    int dev = 0, cus = 0;
    SJ_CHECK(hipGetDevice(&dev));
    SJ_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int per_cu = 4 * (waves == 4 ? 4 : waves == 6 ? 6 : 8) / (kWalkBlock / 64);
    g = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)cus * per_cu));
  }
  const RingIndexView R = ri.view();
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, g, kWalkBlock, 0, s, R, probe.xy.p, cb, cp, n, n_dev, probe_ids, ob, op, cap, counter,
                       stats, redo, n_redo, rest, n_rest, queue);
  };
  if (raster_checked) {
    if (waves == 4) go(k_walk_rings<4, false>);
    else if (waves == 6) go(k_walk_rings<6, false>);
    else go(k_walk_rings<8, false>);
  } else {
    if (waves == 4) go(k_walk_rings<4, true>);
    else if (waves == 6) go(k_walk_rings<6, true>);
    else go(k_walk_rings<8, true>);
  }
  SJ_CHECK(hipGetLastError());
}

}  // namespace sjhip
