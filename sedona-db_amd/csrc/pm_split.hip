// pm_split.hip -- polygon-major filter of Morton-ordered point probes, with
// the cell-raster split of the point-in-polygon refine folded in.
//
// The probe-major filter (the cell grid of build rects, one lane per probe)
// emits candidates in probe order, and the walk wants them grouped by
// polygon: that took a stable radix sort of every open candidate by build
// row.  Here the roles are swapped.  The probe points are already sorted by
// their 32-bit Morton keys (16 bits per axis in the probe frame), so every
// aligned square of 2^s x 2^s Morton units -- a cell -- holds one
// contiguous range of them.  Each build rect (in the rect tree's Morton
// order) visits the cells its rect covers and tests their points, 64 at a
// time, so its candidates come out together, in Morton order within a
// cell.
//
// Filter semantics are those of the rect filter (bvh.hip): a pair is a
// candidate when the FP32 build rect and the probe's FP32 point (rounded
// to nearest) meet, bounds inclusive.  The cells a rect covers are found
// with the same monotone FP32 map that made the probes' keys
// (k_query_keys: q(v) = clamp(((0.5v + 0.5v) - v0) * scale, 0, 65535)),
// so every point inside the rect lies in one of them.
//
// Passes (one join):
//   k_pm_ends   one past the last sorted position of each occupied cell
//               (cells are key >> 2s), then a max scan: each cell's start;
//   k_pm_count  per build rect, the 64-point steps of its covered cells;
//   scan        step offsets per rect;
//   k_pm_fill   one record (rect, first point, points <= 64) per step;
//   k_pm_test   per step, the hit lanes, split by the cell raster of the
//               build row into kept (clean inside cell), open (near the
//               boundary, or no raster) and dropped (clean outside cell),
//               as two 64-bit masks per step;
//   scan        kept / open offsets per step (packed in one 64-bit word);
//   k_pm_emit   the kept pairs into the join's output and the open ones
//               into the walk's list, each step's at its scanned offset.
// The output order is deterministic: build rects in Morton order, then
// cells row by row, then points in Morton order.
#include <hipcub/hipcub.hpp>

#include "ring_walk.h"
#include "sjhip.h"

namespace sjhip {

namespace {

constexpr int kPmBlock = 256;
// rects covering more cells than this take the probe-major path
constexpr int64_t kPmMaxCellsPerRect = 1 << 14;

struct PmGrid {
  float x0, y0, sx, sy;  // Morton frame of the probe keys
  int s;                 // cell = 2^s x 2^s Morton units
  int L;                 // cells per axis = 2^L (L = 16 - s)
  double dist;           // ST_DWithin distance (< 0 or NaN: a relation join)
};

__device__ __forceinline__ uint32_t pm_spread16(uint32_t x) {
  x &= 0xffff;
  x = (x | (x << 8)) & 0x00ff00ff;
  x = (x | (x << 4)) & 0x0f0f0f0f;
  x = (x | (x << 2)) & 0x33333333;
  x = (x | (x << 1)) & 0x55555555;
  return x;
}

// The probe keys' coordinate map (k_query_keys applied to a point rect,
// whose centre 0.5v + 0.5v is v itself for normal floats), monotone in v.
__device__ __forceinline__ uint32_t pm_q(float v, float v0, float sc) {
  const float c = 0.5f * v + 0.5f * v;
  return (uint32_t)fminf(fmaxf((c - v0) * sc, 0.f), 65535.f);
}

__device__ __forceinline__ uint32_t rl32(uint32_t v, int j) { return (uint32_t)__builtin_amdgcn_readlane((int)v, j); }
__device__ __forceinline__ uint64_t rl64(uint64_t v, int j) {
  return ((uint64_t)rl32((uint32_t)(v >> 32), j) << 32) | rl32((uint32_t)v, j);
}

// ends[c] = one past the last sorted position of cell c (cells with no
// point keep 0); an inclusive max scan then gives each cell's start.
__global__ void k_pm_ends(const uint32_t* __restrict__ keys, int64_t nq, int sh, uint32_t* __restrict__ ends) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nq) return;
  const uint32_t c = keys[e] >> sh;
  if (e == nq - 1 || (keys[e + 1] >> sh) != c) ends[c] = (uint32_t)(e + 1);
}

__global__ void k_pm_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The cell range that holds every point whose probe rect can meet build
// rect r.  For ST_DWithin the probe rect is the point's outward FP32 box
// widened by dist and rounded outward again (finish_rect), which meets r
// only when the point lies within dist plus a few FP32 ulps of r: r is
// widened by dist + 2^-20 (|r| + dist), and the bounds rounded outward to
// FP32 hold every such point rounded to nearest, which is what its key is
// made from.
__device__ __forceinline__ void pm_range(const PmGrid& g, const float4& r, uint32_t& cx0, uint32_t& cx1,
                                         uint32_t& cy0, uint32_t& cy1) {
  float x0 = r.x, x1 = r.z, y0 = r.y, y1 = r.w;
  if (g.dist >= 0.0) {
    const double mx = g.dist + 0x1p-20 * (fmax(fabs((double)r.x), fabs((double)r.z)) + g.dist) + 1e-30;
    const double my = g.dist + 0x1p-20 * (fmax(fabs((double)r.y), fabs((double)r.w)) + g.dist) + 1e-30;
    const float2 bx = outward((double)r.x - mx, (double)r.z + mx), by = outward((double)r.y - my, (double)r.w + my);
    x0 = bx.x;
    x1 = bx.y;
    y0 = by.x;
    y1 = by.y;
  }
  cx0 = pm_q(x0, g.x0, g.sx) >> g.s;
  cx1 = pm_q(x1, g.x0, g.sx) >> g.s;
  cy0 = pm_q(y0, g.y0, g.sy) >> g.s;
  cy1 = pm_q(y1, g.y0, g.sy) >> g.s;
}

// cnt[i] = 64-point steps over the cells of leaf rect i (cnt[n] = 0), one
// wavefront per rect, lanes over its cells; *flag set when a rect covers
// too many cells or has too many steps.
__global__ __launch_bounds__(kPmBlock) void k_pm_count(PmGrid g, const float4* __restrict__ rects, int64_t n,
                                                       const uint32_t* __restrict__ start, uint32_t* __restrict__ cnt,
                                                       unsigned int* flag) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)(((uint64_t)blockIdx.x * kPmBlock + threadIdx.x) >> 6);
  if (i > n) return;
  if (i == n) {
    if (lane == 0) cnt[n] = 0;
    return;
  }
  const float4 r = rects[i];
  uint64_t steps = 0;
  if (r.x <= r.z && r.y <= r.w) {
    uint32_t cx0, cx1, cy0, cy1;
    pm_range(g, r, cx0, cx1, cy0, cy1);
    const uint32_t w = cx1 - cx0 + 1;
    const int64_t cells = (int64_t)w * (int64_t)(cy1 - cy0 + 1);
    if (cells > kPmMaxCellsPerRect) {
      if (lane == 0) atomicOr(flag, 1u);
    } else {
      for (int64_t idx = lane; idx < cells; idx += 64) {
        const uint32_t cx = cx0 + (uint32_t)(idx % w), cy = cy0 + (uint32_t)(idx / w);
        const uint32_t c = pm_spread16(cx) | (pm_spread16(cy) << 1);
        steps += (start[c + 1] - start[c] + 63u) >> 6;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) steps += __shfl_xor(steps, o);
  if (lane == 0) {
    if (steps > 0x7fffffffull) {
      atomicOr(flag, 2u);
      steps = 0;
    }
    cnt[i] = (uint32_t)steps;
  }
}

// One record per step: leaf rect index, first point, points in the step.
// One wavefront per rect: lanes load the starts of 64 cells at a time (row
// order) and scan their step counts; then each cell's records are written
// by the lanes in turn.
__global__ __launch_bounds__(kPmBlock) void k_pm_fill(PmGrid g, const float4* __restrict__ rects, int64_t n,
                                                      const uint32_t* __restrict__ start,
                                                      const uint64_t* __restrict__ off, uint32_t* __restrict__ rec_i,
                                                      uint32_t* __restrict__ rec_s, uint8_t* __restrict__ rec_n) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)(((uint64_t)blockIdx.x * kPmBlock + threadIdx.x) >> 6);
  if (i >= n) return;
  uint64_t k = off[i];
  const uint64_t end = off[i + 1];
  if (k == end) return;
  const float4 r = rects[i];
  uint32_t cx0, cx1, cy0, cy1;
  pm_range(g, r, cx0, cx1, cy0, cy1);
  const uint32_t w = cx1 - cx0 + 1;
  const int64_t cells = (int64_t)w * (int64_t)(cy1 - cy0 + 1);
  for (int64_t c0 = 0; c0 < cells && k < end; c0 += 64) {
    const int64_t idx = c0 + lane;
    uint32_t a = 0, b = 0;
    if (idx < cells) {
      const uint32_t cx = cx0 + (uint32_t)(idx % w), cy = cy0 + (uint32_t)(idx / w);
      const uint32_t c = pm_spread16(cx) | (pm_spread16(cy) << 1);
      a = start[c];
      b = start[c + 1];
    }
    const uint32_t ns = (b - a + 63u) >> 6;
    uint32_t incl = ns;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o);
      if (lane >= o) incl += v;
    }
    const uint32_t excl = incl - ns;
    const int nc = (int)min<int64_t>(64, cells - c0);
    for (int j = 0; j < nc; j++) {
      const uint32_t nsj = rl32(ns, j);
      if (nsj == 0) continue;
      const uint32_t aj = rl32(a, j), bj = rl32(b, j), ej = rl32(excl, j);
      for (uint32_t t = lane; t < nsj; t += 64) {
        const uint64_t kk = k + ej + t;
        if (kk < end) {
          rec_i[kk] = (uint32_t)i;
          rec_s[kk] = aj + 64u * t;
          rec_n[kk] = (uint8_t)min(64u, bj - aj - 64u * t);
        }
      }
    }
    k += rl32(incl, 63);
  }
}

// Per step (64 per wavefront): the hit lanes split into kept / open masks
// (MODE 1: point in polygon with cell rasters), all hits open (MODE 0), or
// (MODE 2: ST_DWithin) the point's widened rect tested, then by the build
// rows' rasters (draster 1: line rasters drop the hits no segment can be
// within dist of; 2: polygon cell rasters keep the hits in clean inside
// cells, at distance 0; 0: none) ; packed counts (kept << 32 | open) for
// the scan; the wavefront's hits in whits.
template <int MODE>
__global__ __launch_bounds__(kPmBlock) void k_pm_test(const float4* __restrict__ rects,
                                                      const uint32_t* __restrict__ ids, RingIndexView R,
                                                      const double2* __restrict__ xy, const uint32_t* __restrict__ rec_i,
                                                      const uint32_t* __restrict__ rec_s,
                                                      const uint8_t* __restrict__ rec_n, uint64_t S,
                                                      ulonglong2* __restrict__ masks, uint64_t* __restrict__ cnt2,
                                                      uint64_t* __restrict__ whits, double dist, int draster,
                                                      const float4* __restrict__ qrects) {
  const int lane = threadIdx.x & 63;
  const uint64_t w = ((uint64_t)blockIdx.x * kPmBlock + threadIdx.x) >> 6;
  const uint64_t k0 = w * 64;
  if (k0 >= S) return;
  const uint64_t k = k0 + lane;
  const bool have = k < S;
  const uint32_t my_i = have ? rec_i[k] : 0u, my_s = have ? rec_s[k] : 0u, my_n = have ? rec_n[k] : 0u;
  const int steps = (int)min<uint64_t>(64, S - k0);
  uint64_t my_mk = 0, my_mo = 0;
  uint32_t hits = 0;
  uint32_t cur = 0xffffffffu, bid = 0;
  float4 r = make_float4(0, 0, 0, 0);
  PolyDesc d{};
  double2 inv = make_double2(0.0, 0.0);  // the raster's cells per unit of the current row
  // the next step's point (MODE 2: its probe rect) is loaded before this
  // step is tested
  double2 nxt = make_double2(0.0, 0.0);
  float4 nxq = make_float4(0.f, 0.f, 0.f, 0.f);
  if ((uint32_t)lane < rl32(my_n, 0)) {
    if (MODE == 2) nxq = qrects[rl32(my_s, 0) + lane];
    else nxt = xy[rl32(my_s, 0) + lane];
  }
  for (int j = 0; j < steps; j++) {
    const uint32_t i = rl32(my_i, j), n = rl32(my_n, j);
    double2 pt = nxt;
    const float4 q = nxq;
    if (j + 1 < steps && (uint32_t)lane < rl32(my_n, j + 1)) {
      if (MODE == 2) nxq = qrects[rl32(my_s, j + 1) + lane];
      else nxt = xy[rl32(my_s, j + 1) + lane];
    }
    if (i != cur) {
      cur = i;
      r = rects[i];
      bid = ids[i];
      if (MODE == 1 || (MODE == 2 && draster != 0)) {
        d = R.desc[bid];
        if (d.raster_off >= 0)
          inv = cells_per_unit(d.rect, MODE == 2 && draster == 1 ? (double)kLineCells : (double)kRasterCells);
      }
    }
    const bool in = (uint32_t)lane < n;
    bool h;
    if (MODE == 2) {
      // the probe rect of the ST_DWithin filter (compute_rects with the
      // distance, made once per point by k_pm_probe_rects) against the
      // build rect; the point itself only where a raster reads it
      h = in && !(q.x > r.z || q.z < r.x || q.y > r.w || q.w < r.y);
      if (h && draster != 0) pt = xy[rl32(my_s, j) + lane];
    } else {
      const float fx = (float)pt.x, fy = (float)pt.y;
      h = in && !(fx > r.z || fx < r.x || fy > r.w || fy < r.y);
    }
    bool keep = false, open = h;
    if (MODE == 1 && h) {
      const int cls = raster_class_inv(R, d, pt, inv);
      keep = cls == 1;
      open = cls == 0;
    }
    if (MODE == 2 && h && draster == 1) open = line_raster_may_hold_d(R, d, inv, pt, dist);
    if (MODE == 2 && h && draster == 2) {
      keep = raster_class_inv(R, d, pt, inv) == 1;
      open = !keep;
    }
    const uint64_t mk = __ballot(keep), mo = __ballot(open);
    hits += (uint32_t)__popcll(__ballot(h));
    if (lane == j) {
      my_mk = mk;
      my_mo = mo;
    }
  }
  if (have) {
    masks[k] = make_ulonglong2(my_mk, my_mo);
    cnt2[k] = ((uint64_t)__popcll(my_mk) << 32) | (uint64_t)__popcll(my_mo);
  }
  if (lane == 0) whits[w] = hits;  // (one 64-bit word per wavefront: the sum needs no wider accumulator)
}

// The r-th (0-based) set bit of m, which has more than r set bits.
__device__ __forceinline__ uint32_t pm_nth_bit(uint64_t m, uint32_t r) {
  uint32_t pos = 0, v = (uint32_t)m, c = (uint32_t)__popc(v);
  if (r >= c) {
    r -= c;
    v = (uint32_t)(m >> 32);
    pos = 32;
  }
  c = (uint32_t)__popc(v & 0xffffu);
  if (r >= c) { r -= c; v >>= 16; pos += 16; }
  c = (uint32_t)__popc(v & 0xffu);
  if (r >= c) { r -= c; v >>= 8; pos += 8; }
  c = (uint32_t)__popc(v & 0xfu);
  if (r >= c) { r -= c; v >>= 4; pos += 4; }
  c = (uint32_t)__popc(v & 0x3u);
  if (r >= c) { r -= c; v >>= 2; pos += 2; }
  if (r >= (v & 1u)) pos += 1;
  return pos;
}

// Writes the kept pairs (build id, probe_ids[position]) and the open pairs
// (build id, position) of 64 steps per wavefront.  The steps' outputs are
// contiguous (their offsets come from one scan), so each lane takes every
// 64th output of the wavefront's range, finds its step by binary search
// over the step offsets (in LDS) and its point as the n-th set bit of the
// step's mask: every store is coalesced.
template <bool RASTER>
__global__ __launch_bounds__(kPmBlock) void k_pm_emit(const uint32_t* __restrict__ ids,
                                                      const uint32_t* __restrict__ rec_i,
                                                      const uint32_t* __restrict__ rec_s, uint64_t S,
                                                      const ulonglong2* __restrict__ masks,
                                                      const uint64_t* __restrict__ off2,
                                                      const uint32_t* __restrict__ probe_ids, uint32_t* __restrict__ kb,
                                                      uint32_t* __restrict__ kp, uint32_t* __restrict__ ob,
                                                      uint32_t* __restrict__ op) {
  constexpr int kW = kPmBlock / 64;
  __shared__ uint32_t s_ko[kW][64], s_oo[kW][64], s_b[kW][64], s_st[kW][64];
  __shared__ uint64_t s_mk[kW][64], s_mo[kW][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t w = ((uint64_t)blockIdx.x * kPmBlock + threadIdx.x) >> 6;
  const uint64_t k0 = w * 64;
  const int steps = k0 < S ? (int)min<uint64_t>(64, S - k0) : 0;
  const uint64_t k = k0 + lane;
  uint64_t o = 0, o_end = 0;
  if (lane < steps) {
    const ulonglong2 m = masks[k];
    o = off2[k];
    s_mk[wv][lane] = m.x;
    s_mo[wv][lane] = m.y;
    s_b[wv][lane] = ids[rec_i[k]];
    s_st[wv][lane] = rec_s[k];
  }
  if (steps > 0 && lane == 0) o_end = off2[k0 + steps];  // (off2 holds S + 1 entries)
  const uint64_t base = rl64(o, 0), stop = rl64(o_end, 0);
  // offsets relative to the wavefront's first step (below 64 * 64 each)
  if (lane < steps) {
    s_ko[wv][lane] = (uint32_t)((o >> 32) - (base >> 32));
    s_oo[wv][lane] = (uint32_t)((o & 0xffffffffull) - (base & 0xffffffffull));
  }
  __syncthreads();
  if (steps == 0) return;
  for (int list = RASTER ? 0 : 1; list < 2; list++) {
    const uint32_t* so = list == 0 ? s_ko[wv] : s_oo[wv];
    const uint64_t* sm = list == 0 ? s_mk[wv] : s_mo[wv];
    const uint64_t b0 = list == 0 ? (base >> 32) : (base & 0xffffffffull);
    const uint32_t total = (uint32_t)((list == 0 ? (stop >> 32) : (stop & 0xffffffffull)) - b0);
    for (uint32_t x = lane; x < total; x += 64) {
      int j = 0;
      for (int d = 32; d > 0; d >>= 1)
        if (j + d < steps && so[j + d] <= x) j += d;
      const uint32_t pos = s_st[wv][j] + pm_nth_bit(sm[j], x - so[j]);
      if (list == 0) {
        kb[b0 + x] = s_b[wv][j];
        kp[b0 + x] = probe_ids[pos];
      } else {
        ob[b0 + x] = s_b[wv][j];
        op[b0 + x] = pos;
      }
    }
  }
}

__global__ void k_pm_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// The ST_DWithin probe rect of each point (compute_rects' form: the point's
// outward FP32 box widened by dist, rounded outward again).
__global__ void k_pm_probe_rects(const double2* __restrict__ xy, int64_t n, double dist, float4* __restrict__ q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 p = xy[i];
  q[i] = finish_rect(p.x, p.y, p.x, p.y, false, dist);
}

}  // namespace

bool pm_split(const RectTree& t, const RingIndex& ri, const double2* xy, const uint32_t* keys, int64_t nq,
              const MortonFrame& f, const uint32_t* probe_ids, bool raster, PairBuf& out, PairBuf& open,
              PmScratch& sc, PmSplitResult* res, hipStream_t s, double dist, int dist_raster) {
  *res = PmSplitResult{};
  if (t.grid_nx <= 0 || t.n_valid <= 0 || nq <= 0 || nq >= (int64_t)0xffffffffll || !(t.grid_inv_x > 0.f))
    return false;
  const bool dwithin = dist >= 0.0 && dist < INFINITY;
  if (dwithin && dist_raster != 0 && ri.desc.p == nullptr) return false;
  if (dwithin && dist_raster == 1 && ri.n_line_rasters == 0) return false;
  if (dwithin && dist_raster == 2 && ri.n_rasters == 0) return false;
  if (raster && (dwithin || ri.desc.p == nullptr || ri.n_rasters == 0)) return false;
  const int mode = dwithin ? 2 : raster ? 1 : 0;
  // kept pairs: point in polygon with rasters, or ST_DWithin with polygon rasters
  const bool kept_list = raster || (dwithin && dist_raster == 2);
  // cells about half the median build rect wide: the rect grid's cells are
  // the median extent (coarser where the rects are large)
  const double cell_world = 1.0 / (double)t.grid_inv_x;
  const double scale = f.sx > 0.f ? (double)f.sx : (double)f.sy;
  int sh = 15;
  if (scale > 0.0) {
    const double units = 0.5 * cell_world * scale;
    sh = units <= 1.0 ? 0 : (int)std::lround(std::log2(units));
  }
  sh = std::min(15, std::max(4, sh));  // 4^(16 - sh) + 1 cell starts: at most 2^24 + 1
  PmGrid g{f.x0, f.y0, f.sx, f.sy, sh, 16 - sh, dwithin ? dist : -1.0};
  const int64_t ncell = (int64_t)1 << (2 * g.L);
  const int64_t n = t.n_valid;
  fit(sc.start, (size_t)(ncell + 1), s);
  fit(sc.cnt, (size_t)(n + 1), s);
  fit(sc.off, (size_t)(n + 1), s);
  fit(sc.flag, 1, s);
  // cell starts: ends of the occupied cells, then an inclusive max scan
  // (start[0] = 0; start[c + 1] = the largest end of cells 0..c)
  hipLaunchKernelGGL(k_pm_fill_u32, (unsigned)((ncell + 1 + kPmBlock - 1) / kPmBlock), kPmBlock, 0, s, sc.start.p,
                     ncell + 1, 0u);
  fit(sc.ends, (size_t)ncell, s);
  hipLaunchKernelGGL(k_pm_fill_u32, (unsigned)((ncell + kPmBlock - 1) / kPmBlock), kPmBlock, 0, s, sc.ends.p, ncell,
                     0u);
  hipLaunchKernelGGL(k_pm_ends, (unsigned)((nq + kPmBlock - 1) / kPmBlock), kPmBlock, 0, s, keys, nq, 2 * sh,
                     sc.ends.p);
  {
    size_t tb = 0;
    SJ_CHECK(hipcub::DeviceScan::InclusiveScan(nullptr, tb, sc.ends.p, sc.start.p + 1, hipcub::Max(), (int)ncell, s));
    fit(sc.tmp, tb + 16, s);
    SJ_CHECK(hipcub::DeviceScan::InclusiveScan(sc.tmp.p, tb, sc.ends.p, sc.start.p + 1, hipcub::Max(), (int)ncell, s));
  }
  hipLaunchKernelGGL(k_pm_set_u64, 1, 1, 0, s, (unsigned long long*)sc.flag.p, 0ull);
  const unsigned bw = (unsigned)(((n + 1) * 64 + kPmBlock - 1) / kPmBlock);  // one wavefront per rect
  hipLaunchKernelGGL(k_pm_count, bw, kPmBlock, 0, s, g, t.leaf_rects.p, n, sc.start.p, sc.cnt.p,
                     (unsigned int*)sc.flag.p);
  SJ_CHECK(hipGetLastError());
  {
    size_t tb = 0;
    SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, sc.cnt.p, sc.off.p, (int)(n + 1), s));
    fit(sc.tmp, tb + 16, s);
    SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(sc.tmp.p, tb, sc.cnt.p, sc.off.p, (int)(n + 1), s));
  }
  uint64_t h[2] = {0, 0};
  SJ_CHECK(hipMemcpyAsync(&h[0], sc.off.p + n, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipMemcpyAsync(&h[1], sc.flag.p, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  const uint64_t S = h[0];
  if (h[1] != 0 || S >= 0x7fffffffull) return false;
  if (S == 0) {
    out.counter.reserve(1, s);
    hipLaunchKernelGGL(k_pm_set_u64, 1, 1, 0, s, out.counter.p, 0ull);
    SJ_CHECK(hipGetLastError());
    return true;
  }
  fit(sc.rec_i, S, s);
  fit(sc.rec_s, S, s);
  fit(sc.rec_n, S, s);
  fit(sc.masks, S, s);
  fit(sc.cnt2, S + 1, s);
  fit(sc.off2, S + 1, s);
  const uint64_t nw = (S + 63) / 64;
  fit(sc.whits, nw, s);
  fit(sc.hits, 1, s);
  hipLaunchKernelGGL(k_pm_fill, (unsigned)((n * 64 + kPmBlock - 1) / kPmBlock), kPmBlock, 0, s, g, t.leaf_rects.p, n,
                     sc.start.p, sc.off.p, sc.rec_i.p, sc.rec_s.p, sc.rec_n.p);
  const unsigned bs = (unsigned)((S + kPmBlock - 1) / kPmBlock);
  const RingIndexView R = ri.view();
  if (mode == 2) {
    fit(sc.qrect, (size_t)nq, s);
    hipLaunchKernelGGL(k_pm_probe_rects, (unsigned)((nq + kPmBlock - 1) / kPmBlock), kPmBlock, 0, s, xy, nq, dist,
                       sc.qrect.p);
  }
  auto test = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, bs, kPmBlock, 0, s, t.leaf_rects.p, t.leaf_ids.p, R, xy, sc.rec_i.p, sc.rec_s.p,
                       sc.rec_n.p, S, sc.masks.p, sc.cnt2.p, sc.whits.p, dist, dwithin ? dist_raster : 0,
                       mode == 2 ? sc.qrect.p : nullptr);
  };
  if (mode == 1) test(k_pm_test<1>);
  else if (mode == 2) test(k_pm_test<2>);
  else test(k_pm_test<0>);
  SJ_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_pm_set_u64, 1, 1, 0, s, (unsigned long long*)(sc.cnt2.p + S), 0ull);
  {
    size_t tb = 0;
    SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, sc.cnt2.p, sc.off2.p, (int)(S + 1), s));
    size_t tr = 0;
    SJ_CHECK(hipcub::DeviceReduce::Sum(nullptr, tr, sc.whits.p, sc.hits.p, (int)nw, s));
    fit(sc.tmp, std::max(tb, tr) + 16, s);
    SJ_CHECK(hipcub::DeviceScan::ExclusiveSum(sc.tmp.p, tb, sc.cnt2.p, sc.off2.p, (int)(S + 1), s));
    SJ_CHECK(hipcub::DeviceReduce::Sum(sc.tmp.p, tr, sc.whits.p, sc.hits.p, (int)nw, s));
  }
  uint64_t tot = 0, hits = 0;
  SJ_CHECK(hipMemcpyAsync(&tot, sc.off2.p + S, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipMemcpyAsync(&hits, sc.hits.p, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  if (hits >= 0xffffffffull) throw Error("pm_split: 2^32 - 1 candidates or more");
  const uint64_t kept = tot >> 32, n_open = tot & 0xffffffffull;
  res->candidates = hits;
  res->kept = kept;
  res->open = n_open;
  // the walk appends its kept pairs after these (capacity: every candidate)
  if (kept_list) {
    out.counter.reserve(1, s);
    fit(out.b, hits > 0 ? hits : 1, s);
    fit(out.p, hits > 0 ? hits : 1, s);
    hipLaunchKernelGGL(k_pm_set_u64, 1, 1, 0, s, out.counter.p, (unsigned long long)kept);
  }
  fit(open.b, n_open > 0 ? n_open : 1, s);
  fit(open.p, n_open > 0 ? n_open : 1, s);
  if (kept_list)
    hipLaunchKernelGGL(k_pm_emit<true>, bs, kPmBlock, 0, s, t.leaf_ids.p, sc.rec_i.p, sc.rec_s.p, S, sc.masks.p,
                       sc.off2.p, probe_ids, out.b.p, out.p.p, open.b.p, open.p.p);
  else
    hipLaunchKernelGGL(k_pm_emit<false>, bs, kPmBlock, 0, s, t.leaf_ids.p, sc.rec_i.p, sc.rec_s.p, S, sc.masks.p,
                       sc.off2.p, probe_ids, nullptr, nullptr, open.b.p, open.p.p);
  SJ_CHECK(hipGetLastError());
  if (nq > 0) t.cand_per_query_x16.store((uint64_t)((hits * 16) / (uint64_t)nq + 1), std::memory_order_relaxed);
  return true;
}

}  // namespace sjhip
