// decode.hip -- device-side WKB decode of an Arrow geometry column into the
// structure-of-arrays layout (row -> parts -> rings -> points, FP64 xy),
// and per-row FP32 rects.
//
// Replaces the reference's host-thread loader (c/sedona-libgpuspatial/
// libgpuspatial/include/gpuspatial/loader/parallel_wkb_loader.hpp:578-787)
// and the per-row bounds of the operator (rust/sedona-geometry/src/
// bounds.rs:240-249, rust/sedona-spatial-join/src/utils/bounds.rs:147-162,
// rust/sedona-spatial-join-gpu/src/join_provider.rs:94-146).
//
// Passes: (1) scan: one thread per row walks the WKB headers, validates
// them like the reference's parser (c/sedona-tg/src/tg/tg.c:8380-8399:
// lines need 2 positions, rings 3 and must close) and counts parts, rings
// and points; (2) exclusive scans; (3) fill: one thread per row writes the
// offset tables and, per ring, the byte offset of its coordinate run;
// (4) coordinates: one thread per 4 output points reads 16-byte
// coordinate pairs at any byte alignment with aligned dword loads.
// Columns holding only points skip (1)-(4) and decode with one thread per
// row straight into xy[row].
#include <hipcub/hipcub.hpp>

#include <cstring>

#include "../../include/gpuspatial_hip.h"
#include "geom_math.h"
#include "sjhip.h"

namespace sjhip {

namespace {

constexpr int kBlock = 256;

// order-preserving uint -> float (host side of the bounds atomics)
inline float o2f_host(uint32_t u) {
  const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

// Error word: (row << 8) | code; the smallest row wins.
enum DecodeErr : uint32_t {
  kErrTruncated = 1,
  kErrByteOrder = 2,
  kErrType = 3,
  kErrLinePositions = 4,
  kErrRingPositions = 5,
  kErrRingClosed = 6,
};

const char* decode_err_text(uint32_t c) {
  switch (c) {
    case kErrTruncated: return "truncated WKB";
    case kErrByteOrder: return "invalid WKB byte order";
    case kErrType: return "unsupported WKB geometry type (2-D Point..MultiPolygon only)";
    case kErrLinePositions: return "lines must have two or more positions";
    case kErrRingPositions: return "rings must have three or more positions";
    case kErrRingClosed: return "rings must have matching first and last positions";
    default: return "invalid WKB";
  }
}

struct Src {
  const uint8_t* data;   // data buffer (device)
  const void* offsets;   // int32/int64 row offsets (device), already shifted by array offset
  int large;
  const uint8_t* validity;  // bitmap or null
  int64_t bit_offset;
};

__device__ __forceinline__ int64_t row_begin(const Src& s, int64_t r) {
  return s.large ? ((const int64_t*)s.offsets)[r] : (int64_t)((const int32_t*)s.offsets)[r];
}

__device__ __forceinline__ bool row_valid(const Src& s, int64_t r) {
  if (!s.validity) return true;
  int64_t b = r + s.bit_offset;
  return (s.validity[b >> 3] >> (b & 7)) & 1;
}

__device__ __forceinline__ uint32_t ld_u32(const uint8_t* d, int64_t o, bool le) {
  uint32_t b0 = d[o], b1 = d[o + 1], b2 = d[o + 2], b3 = d[o + 3];
  return le ? (b0 | (b1 << 8) | (b2 << 16) | (b3 << 24)) : (b3 | (b2 << 8) | (b1 << 16) | (b0 << 24));
}

// 16 bytes at any byte offset, from aligned dwords (never reads past the
// aligned dword holding byte o+15).
__device__ __forceinline__ void ld_16(const uint8_t* d, int64_t o, uint32_t w[4]) {
  const uint32_t* base = (const uint32_t*)((uintptr_t)(d + o) & ~(uintptr_t)3);
  uint32_t sh = (uint32_t)((uintptr_t)(d + o) & 3);
  uint32_t a0 = base[0], a1 = base[1], a2 = base[2], a3 = base[3];
  if (sh == 0) {
    w[0] = a0; w[1] = a1; w[2] = a2; w[3] = a3;
  } else {
    uint32_t a4 = base[4];
    w[0] = __builtin_amdgcn_alignbyte(a1, a0, sh);
    w[1] = __builtin_amdgcn_alignbyte(a2, a1, sh);
    w[2] = __builtin_amdgcn_alignbyte(a3, a2, sh);
    w[3] = __builtin_amdgcn_alignbyte(a4, a3, sh);
  }
}

__device__ __forceinline__ double2 words_to_xy(const uint32_t w[4], bool le) {
  uint64_t x, y;
  if (le) {
    x = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    y = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
  } else {
    x = ((uint64_t)__builtin_bswap32(w[0]) << 32) | __builtin_bswap32(w[1]);
    y = ((uint64_t)__builtin_bswap32(w[2]) << 32) | __builtin_bswap32(w[3]);
  }
  return make_double2(bitsd(x), bitsd(y));
}

// ORs the geometry types seen by the active lanes into *mask with at most
// one atomic per wavefront.
__device__ __forceinline__ void or_type_mask(unsigned int* mask, uint8_t t) {
  unsigned int m = 0;
#pragma unroll
  for (int k = 1; k <= 6; k++)
    if (__ballot(t == k)) m |= 1u << k;
  uint64_t act = __ballot(1);
  // bits already set need no atomic (most waves of a column see the same
  // types; same-address atomics from every wave serialise in L2)
  if (m && (int)(threadIdx.x & 63) == __ffsll((unsigned long long)act) - 1 &&
      (__atomic_load_n(mask, __ATOMIC_RELAXED) & m) != m)
    atomicOr(mask, m);
}

__device__ __forceinline__ void report(unsigned long long* err, int64_t row, uint32_t code) {
  atomicMin(err, ((unsigned long long)row << 8) | code);
}

// Walks one row; in the fill pass (Fill = true) writes the tables.
// Returns false on error (already reported).
template <bool Fill>
__device__ bool walk_row(const Src& s, int64_t r, int64_t pos, int64_t end, uint8_t* type_out,
                         int64_t* n_parts, int64_t* n_rings, int64_t* n_pts,
                         // fill outputs
                         int64_t part_base, int64_t ring_base, int64_t pt_base, int64_t* ring_off,
                         int64_t* pt_off, int64_t* ring_src, unsigned long long* err) {
  const uint8_t* d = s.data;
  int64_t parts = 0, rings = 0, pts = 0;
  if (pos + 5 > end) { report(err, r, kErrTruncated); return false; }
  uint8_t bo = d[pos];
  if (bo > 1) { report(err, r, kErrByteOrder); return false; }
  bool le = bo == 1;
  uint32_t t = ld_u32(d, pos + 1, le);
  pos += 5;
  if (t < 1 || t > 6) { report(err, r, kErrType); return false; }
  *type_out = (uint8_t)t;
  uint32_t nchild = 1;
  bool multi = t >= 4;
  uint32_t ctype = multi ? t - 3 : t;
  if (multi) {
    if (pos + 4 > end) { report(err, r, kErrTruncated); return false; }
    nchild = ld_u32(d, pos, le);
    pos += 4;
  }
  for (uint32_t c = 0; c < nchild; c++) {
    bool cle = le;
    if (multi) {
      if (pos + 5 > end) { report(err, r, kErrTruncated); return false; }
      uint8_t cbo = d[pos];
      if (cbo > 1) { report(err, r, kErrByteOrder); return false; }
      cle = cbo == 1;
      uint32_t ct = ld_u32(d, pos + 1, cle);
      if (ct != ctype) { report(err, r, kErrType); return false; }
      pos += 5;
    }
    if (ctype == kPoint) {
      if (pos + 16 > end) { report(err, r, kErrTruncated); return false; }
      uint32_t w[4];
      ld_16(d, pos, w);
      double2 xy = words_to_xy(w, cle);
      bool empty = (xy.x != xy.x) && (xy.y != xy.y);
      if (!empty) {
        if (Fill) {
          ring_off[part_base + parts] = ring_base + rings;
          pt_off[ring_base + rings] = pt_base + pts;
          ring_src[ring_base + rings] = pos | (cle ? 0 : (1ll << 62));
        }
        parts++; rings++; pts++;
      }
      pos += 16;
    } else if (ctype == kLine) {
      if (pos + 4 > end) { report(err, r, kErrTruncated); return false; }
      uint32_t n = ld_u32(d, pos, cle);
      pos += 4;
      if (pos + 16 * (int64_t)n > end) { report(err, r, kErrTruncated); return false; }
      if (n == 0) continue;
      if (n < 2) { report(err, r, kErrLinePositions); return false; }
      if (Fill) {
        ring_off[part_base + parts] = ring_base + rings;
        pt_off[ring_base + rings] = pt_base + pts;
        ring_src[ring_base + rings] = pos | (cle ? 0 : (1ll << 62));
      }
      parts++; rings++; pts += n;
      pos += 16 * (int64_t)n;
    } else {  // polygon
      if (pos + 4 > end) { report(err, r, kErrTruncated); return false; }
      uint32_t nr = ld_u32(d, pos, cle);
      pos += 4;
      if (nr == 0) continue;
      if (Fill) ring_off[part_base + parts] = ring_base + rings;
      for (uint32_t k = 0; k < nr; k++) {
        if (pos + 4 > end) { report(err, r, kErrTruncated); return false; }
        uint32_t n = ld_u32(d, pos, cle);
        pos += 4;
        if (pos + 16 * (int64_t)n > end) { report(err, r, kErrTruncated); return false; }
        if (n > 0) {
          if (n < 3) { report(err, r, kErrRingPositions); return false; }
          uint32_t w0[4], w1[4];
          ld_16(d, pos, w0);
          ld_16(d, pos + 16 * (int64_t)(n - 1), w1);
          double2 f = words_to_xy(w0, cle), l = words_to_xy(w1, cle);
          if (!(f.x == l.x && f.y == l.y)) { report(err, r, kErrRingClosed); return false; }
        }
        if (Fill) {
          pt_off[ring_base + rings] = pt_base + pts;
          ring_src[ring_base + rings] = pos | (cle ? 0 : (1ll << 62));
        }
        rings++;
        pts += n;
        pos += 16 * (int64_t)n;
      }
      parts++;
    }
  }
  *n_parts = parts;
  *n_rings = rings;
  *n_pts = pts;
  return true;
}

__global__ void k_scan(Src s, int64_t n, uint8_t* type, int64_t* cp, int64_t* cr, int64_t* cv,
                       unsigned long long* err, unsigned int* type_mask) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint8_t t = 0;
  int64_t np = 0, nr = 0, nv = 0;
  if (row_valid(s, r)) {
    int64_t b = row_begin(s, r), e = row_begin(s, r + 1);
    if (!walk_row<false>(s, r, b, e, &t, &np, &nr, &nv, 0, 0, 0, nullptr, nullptr, nullptr, err)) t = 0;
    if (nv == 0) np = nr = 0;
  }
  type[r] = t;
  cp[r] = np;
  cr[r] = nr;
  cv[r] = nv;
  or_type_mask(type_mask, t);
}

__global__ void k_fill(Src s, int64_t n, const int64_t* part_start, const int64_t* ring_start,
                       const int64_t* pt_start, int64_t* part_off, int64_t* ring_off, int64_t* pt_off,
                       int64_t* ring_src, unsigned long long* err) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  part_off[r] = part_start[r];
  if (!row_valid(s, r)) return;
  if (part_start[r + 1] == part_start[r]) return;
  uint8_t t;
  int64_t np, nr, nv;
  walk_row<true>(s, r, row_begin(s, r), row_begin(s, r + 1), &t, &np, &nr, &nv, part_start[r], ring_start[r],
                 pt_start[r], ring_off, pt_off, ring_src, err);
}

// Points per thread in the coordinate pass.
constexpr int kPtsPerThread = 4;

__global__ void k_coords(const uint8_t* data, const int64_t* pt_off, const int64_t* ring_src, int64_t n_rings,
                         int64_t n_pts, double2* xy) {
  int64_t v0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kPtsPerThread;
  if (v0 >= n_pts) return;
  // ring holding v0: last ring with pt_off[ring] <= v0
  int64_t lo = 0, hi = n_rings - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (pt_off[mid] <= v0) lo = mid;
    else hi = mid - 1;
  }
  int64_t ring = lo;
  int64_t rend = pt_off[ring + 1];
  for (int k = 0; k < kPtsPerThread; k++) {
    int64_t v = v0 + k;
    if (v >= n_pts) break;
    while (v >= rend) {
      ring++;
      rend = pt_off[ring + 1];
    }
    int64_t src = ring_src[ring];
    bool le = (src & (1ll << 62)) == 0;
    int64_t pos = (src & ((1ll << 62) - 1)) + 16 * (v - pt_off[ring]);
    uint32_t w[4];
    ld_16(data, pos, w);
    xy[v] = words_to_xy(w, le);
  }
}

// A 21-byte WKB point record at any byte offset, from six aligned dword
// loads (the last one holds the record's last byte, so nothing past the
// record is read).  w[0] byte 0 is the byte order, bytes 1-4 the type,
// bytes 5-20 the coordinates.
__device__ __forceinline__ void ld_point_record(const uint8_t* d, int64_t o, uint32_t w[6]) {
  const uint32_t* base = (const uint32_t*)((uintptr_t)(d + o) & ~(uintptr_t)3);
  uint32_t sh = (uint32_t)((uintptr_t)(d + o) & 3);
  uint32_t a0 = base[0], a1 = base[1], a2 = base[2], a3 = base[3], a4 = base[4], a5 = base[5];
  w[0] = __builtin_amdgcn_alignbyte(a1, a0, sh);
  w[1] = __builtin_amdgcn_alignbyte(a2, a1, sh);
  w[2] = __builtin_amdgcn_alignbyte(a3, a2, sh);
  w[3] = __builtin_amdgcn_alignbyte(a4, a3, sh);
  w[4] = __builtin_amdgcn_alignbyte(a5, a4, sh);
  w[5] = __builtin_amdgcn_alignbyte(0u, a5, sh);
}

// The fields of a 21-byte point record read by ld_point_record: byte
// order (byte 0), geometry type (bytes 1..4, in that order) and
// coordinates (bytes 5..20).
struct PointRecord {
  uint32_t bo, ty;
  double2 xy;
};

__device__ __forceinline__ PointRecord parse_point_record(const uint32_t w[6]) {
  PointRecord p;
  p.bo = w[0] & 0xff;
  const bool le = p.bo == 1;
  // shift the stream by one byte: c[k] = bytes 1+4k .. 4+4k
  const uint32_t c0 = __builtin_amdgcn_alignbyte(w[1], w[0], 1), c1 = __builtin_amdgcn_alignbyte(w[2], w[1], 1);
  const uint32_t c2 = __builtin_amdgcn_alignbyte(w[3], w[2], 1), c3 = __builtin_amdgcn_alignbyte(w[4], w[3], 1);
  const uint32_t c4 = __builtin_amdgcn_alignbyte(w[5], w[4], 1);
  p.ty = le ? c0 : __builtin_bswap32(c0);
  const uint32_t cw[4] = {c1, c2, c3, c4};
  p.xy = words_to_xy(cw, le);
  return p;
}

// One row of a column of 21-byte WKB points: coordinates (NaN for null
// and empty rows) and type (0 for null rows).
__device__ __forceinline__ void decode_point_row(const Src& s, int64_t r, double2& out, uint8_t& t,
                                                 unsigned long long* err) {
  out = make_double2(NAN, NAN);
  t = 0;
  if (row_valid(s, r)) {
    int64_t b = row_begin(s, r), e = row_begin(s, r + 1);
    if (b + 21 > e) {
      report(err, r, kErrTruncated);
    } else {
      uint32_t w[6];
      ld_point_record(s.data, b, w);
      const PointRecord pr = parse_point_record(w);
      if (pr.bo > 1) report(err, r, kErrByteOrder);
      else if (pr.ty != 1) report(err, r, kErrType);
      else {
        out = pr.xy;
        bool empty = (out.x != out.x) && (out.y != out.y);
        t = kPoint;
        if (empty) out = make_double2(NAN, NAN);
      }
    }
  }
}

// float <-> order-preserving uint (for min / max atomics on floats)
__device__ __forceinline__ uint32_t pt_f2o(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Decodes point rows, one per thread, and accumulates the bounds of the
// valid points rounded to FP32 -- the centres of their point rects -- in
// bounds[0..3] (ordered uints: min x, min y, max x, max y) and the number
// of rows that are not valid points in bounds[4].  A block issues a bounds
// atomic only when it improves on the value it reads (few blocks do once
// the extremes are found) and a count atomic only when it has such rows.
// A valid row that is not a 21-byte point record (of byte order 0 or 1 and
// type 1) sets *not_points and is left empty: the caller then decodes the
// column with the general passes, which report its errors.
__global__ __launch_bounds__(kBlock) void k_points(Src s, int64_t n, double2* xy, uint8_t* type,
                                                   unsigned long long* err, unsigned int* not_points,
                                                   unsigned int* type_mask, unsigned int* bounds) {
  __shared__ unsigned int sb[5][kBlock / 64];
  unsigned int mnx = 0xffffffffu, mny = 0xffffffffu, mxx = 0, mxy = 0, bad = 0;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint8_t t = 0;
  bool other = false;
  if (r < n) {
    double2 out = make_double2(NAN, NAN);
    if (row_valid(s, r)) {
      const int64_t b = row_begin(s, r), e = row_begin(s, r + 1);
      bool ok = false;
      if (e - b == 21) {
        // the record is read once, as aligned dwords
        uint32_t w[6];
        ld_point_record(s.data, b, w);
        const PointRecord pr = parse_point_record(w);
        ok = pr.bo <= 1 && pr.ty == 1;
        if (ok) {
          out = pr.xy;
          t = kPoint;
          if (out.x != out.x && out.y != out.y) out = make_double2(NAN, NAN);  // empty
        }
      }
      other = !ok;
    }
    xy[r] = out;
    type[r] = t;
    const float fx = (float)out.x, fy = (float)out.y;
    if (t != 0 && fx == fx && fy == fy) {
      const float cx = 0.5f * fx + 0.5f * fx, cy = 0.5f * fy + 0.5f * fy;
      mnx = mxx = pt_f2o(cx);
      mny = mxy = pt_f2o(cy);
    } else {
      bad = 1;
    }
  }
  if (__ballot(other) && (threadIdx.x & 63) == 0) atomicOr(not_points, 1u);
  or_type_mask(type_mask, t);
  for (int o = 32; o > 0; o >>= 1) {
    mnx = min(mnx, (unsigned int)__shfl_xor((int)mnx, o));
    mny = min(mny, (unsigned int)__shfl_xor((int)mny, o));
    mxx = max(mxx, (unsigned int)__shfl_xor((int)mxx, o));
    mxy = max(mxy, (unsigned int)__shfl_xor((int)mxy, o));
    bad += (unsigned int)__shfl_xor((int)bad, o);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    sb[0][wv] = mnx; sb[1][wv] = mny; sb[2][wv] = mxx; sb[3][wv] = mxy; sb[4][wv] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; w++) {
      mnx = min(mnx, sb[0][w]); mny = min(mny, sb[1][w]);
      mxx = max(mxx, sb[2][w]); mxy = max(mxy, sb[3][w]);
      bad += sb[4][w];
    }
    if (mnx < __atomic_load_n(&bounds[0], __ATOMIC_RELAXED)) atomicMin(&bounds[0], mnx);
    if (mny < __atomic_load_n(&bounds[1], __ATOMIC_RELAXED)) atomicMin(&bounds[1], mny);
    if (mxx > __atomic_load_n(&bounds[2], __ATOMIC_RELAXED)) atomicMax(&bounds[2], mxx);
    if (mxy > __atomic_load_n(&bounds[3], __ATOMIC_RELAXED)) atomicMax(&bounds[3], mxy);
    if (bad) atomicAdd(&bounds[4], bad);
  }
}

// The same for a column described in device memory, over a fixed capacity
// of rows (rows at or beyond d->n are empty): a launch that a captured
// sequence replays over new inputs.  Rows that are not 21-byte point
// records report an error.
__global__ void k_points_desc(const PointsDesc* d, int64_t cap, double2* xy, uint8_t* type,
                              unsigned long long* err) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= cap) return;
  double2 out = make_double2(NAN, NAN);
  uint8_t t = 0;
  if (r < d->n) {
    Src s;
    s.data = d->data;
    s.offsets = d->offsets;
    s.large = d->large;
    s.validity = nullptr;
    s.bit_offset = 0;
    if (row_begin(s, r + 1) - row_begin(s, r) != 21) report(err, r, kErrType);
    else decode_point_row(s, r, out, t, err);
  }
  xy[r] = out;
  type[r] = t;
}

// Int32 BinaryView -> int64 offsets + packed data (host side of the view
// conversion is in wkb_input_from_arrow).

template <typename T>
T* dev_upload(const T* host, size_t n, DevArray<T>& keep, hipStream_t s) {
  keep.alloc(n, s);
  if (n) SJ_CHECK(hipMemcpyAsync(keep.p, host, n * sizeof(T), hipMemcpyHostToDevice, s));
  return keep.p;
}

void exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s) {
  // out has n+1 entries: out[0] = 0, out[i+1] = sum(in[0..i])
  size_t tmp_bytes = 0;
  SJ_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, in, out + 1, (int)n, s));
  DevArray<uint8_t> tmp(tmp_bytes + 16, s);
  SJ_CHECK(hipMemsetAsync(out, 0, sizeof(int64_t), s));
  if (n) SJ_CHECK(hipcub::DeviceScan::InclusiveSum(tmp.p, tmp_bytes, in, out + 1, (int)n, s));
}

unsigned int blocks_for(int64_t n, int per_thread = 1) {
  int64_t t = (n + per_thread - 1) / per_thread;
  return (unsigned int)((t + kBlock - 1) / kBlock);
}

}  // namespace

WkbInput wkb_input_from_arrow(const struct ArrowSchema* schema, const struct ArrowArray* array, bool on_device) {
  if (!schema || !array) throw Error("null Arrow schema or array");
  const char* f = schema->format;
  if (!f) throw Error("Arrow schema without format");
  WkbInput in;
  in.length = array->length;
  in.offset = array->offset;
  in.on_device = on_device;
  if (array->null_count != 0 && array->n_buffers > 0) in.validity = (const uint8_t*)array->buffers[0];
  if (std::strcmp(f, "z") == 0 || std::strcmp(f, "Z") == 0) {
    if (array->n_buffers != 3) throw Error("Binary array must have 3 buffers");
    in.large = f[0] == 'Z';
    in.offsets = array->buffers[1];
    in.data = (const uint8_t*)array->buffers[2];
  } else if (std::strcmp(f, "vz") == 0) {
    if (on_device) throw Error("BinaryView geometry columns must be passed in host memory");
    if (array->n_buffers < 3) throw Error("BinaryView array must have at least 3 buffers");
    in.view = true;
    in.views = (const uint8_t*)array->buffers[1];
    int64_t nvar = array->n_buffers - 3;
    const int64_t* sizes = (const int64_t*)array->buffers[array->n_buffers - 1];
    for (int64_t i = 0; i < nvar; i++) {
      in.view_buffers.push_back((const uint8_t*)array->buffers[2 + i]);
      in.view_buffer_sizes.push_back(sizes ? sizes[i] : 0);
    }
  } else {
    throw Error(std::string("unsupported geometry storage type '") + f + "' (expected Binary, LargeBinary or BinaryView WKB)");
  }
  return in;
}

void decode_wkb(const WkbInput& in0, GeomColumn& g, hipStream_t s) {
  const int64_t n = in0.length;
  g = GeomColumn();
  g.n_rows = n;
  // Host-side materialisation of the row offsets (int64, rebased) and of a
  // compact data range; device-side inputs keep their buffers.
  DevArray<int64_t> d_off_keep;
  DevArray<uint8_t> d_data_keep, d_valid_keep;
  Src src{};
  src.large = 1;
  src.bit_offset = in0.offset;
  std::vector<int64_t> off64;
  std::vector<uint8_t> packed;
  if (in0.view) {
    // BinaryView: pack into contiguous bytes on the host.
    off64.resize(n + 1);
    off64[0] = 0;
    for (int64_t i = 0; i < n; i++) {
      const uint8_t* v = in0.views + 16 * (in0.offset + i);
      int32_t len;
      std::memcpy(&len, v, 4);
      off64[i + 1] = off64[i] + (len > 0 ? len : 0);
    }
    packed.resize((size_t)off64[n] + 16);
    for (int64_t i = 0; i < n; i++) {
      const uint8_t* v = in0.views + 16 * (in0.offset + i);
      int32_t len;
      std::memcpy(&len, v, 4);
      if (len <= 0) continue;
      if (len <= 12) {
        std::memcpy(&packed[off64[i]], v + 4, len);
      } else {
        int32_t bi, bo;
        std::memcpy(&bi, v + 8, 4);
        std::memcpy(&bo, v + 12, 4);
        if (bi < 0 || (size_t)bi >= in0.view_buffers.size()) throw Error("BinaryView buffer index out of range");
        std::memcpy(&packed[off64[i]], in0.view_buffers[bi] + bo, len);
      }
    }
    src.offsets = dev_upload(off64.data(), off64.size(), d_off_keep, s);
    src.data = dev_upload(packed.data(), packed.size(), d_data_keep, s);
  } else if (!in0.on_device) {
    off64.resize(n + 1);
    if (in0.large) {
      const int64_t* o = (const int64_t*)in0.offsets + in0.offset;
      for (int64_t i = 0; i <= n; i++) off64[i] = o[i] - o[0];
    } else {
      const int32_t* o = (const int32_t*)in0.offsets + in0.offset;
      for (int64_t i = 0; i <= n; i++) off64[i] = (int64_t)o[i] - o[0];
    }
    int64_t base = in0.large ? ((const int64_t*)in0.offsets)[in0.offset] : ((const int32_t*)in0.offsets)[in0.offset];
    src.offsets = dev_upload(off64.data(), off64.size(), d_off_keep, s);
    d_data_keep.alloc((size_t)off64[n] + 32, s);
    if (off64[n]) SJ_CHECK(hipMemcpyAsync(d_data_keep.p, in0.data + base, (size_t)off64[n], hipMemcpyHostToDevice, s));
    SJ_CHECK(hipMemsetAsync(d_data_keep.p + off64[n], 0, 32, s));
    src.data = d_data_keep.p;
  } else {
    src.large = in0.large ? 1 : 0;
    src.offsets = in0.large ? (const void*)((const int64_t*)in0.offsets + in0.offset)
                            : (const void*)((const int32_t*)in0.offsets + in0.offset);
    src.data = in0.data;
  }
  if (in0.validity) {
    if (in0.on_device && !in0.view) {
      src.validity = in0.validity;
    } else {
      int64_t nbytes = (in0.offset + n + 7) / 8;
      src.validity = dev_upload(in0.validity, (size_t)nbytes, d_valid_keep, s);
    }
  }
  if (!in0.view && !in0.on_device) src.bit_offset = in0.offset;

  DevArray<unsigned long long> d_err(1, s);
  // [0] type mask, [1] not-all-points, [2..6] point bounds and non-point rows (k_points)
  DevArray<unsigned int> d_flags(7, s);
  unsigned long long init_err = ~0ull;
  SJ_CHECK(hipMemcpyAsync(d_err.p, &init_err, 8, hipMemcpyHostToDevice, s));
  const unsigned int init_flags[7] = {0u, 0u, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  SJ_CHECK(hipMemcpyAsync(d_flags.p, init_flags, sizeof(init_flags), hipMemcpyHostToDevice, s));
  g.type.alloc(n > 0 ? n : 1, s);

  // a column of point records is decoded in one pass; any other row sends
  // the column through the general passes
  unsigned int flags[2];
  unsigned long long herr = ~0ull;
  g.xy.alloc(n > 0 ? n : 1, s);
  if (n > 0)
    hipLaunchKernelGGL(k_points, blocks_for(n), kBlock, 0, s, src, n, g.xy.p, g.type.p, d_err.p, d_flags.p + 1,
                       d_flags.p, d_flags.p + 2);
  unsigned int pf[7];
  SJ_CHECK(hipMemcpyAsync(&herr, d_err.p, 8, hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipMemcpyAsync(pf, d_flags.p, sizeof(pf), hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  flags[0] = pf[0];
  flags[1] = pf[1];
  if (flags[1] == 0) {
    g.points_only = true;
    g.pt_valid = n - (int64_t)pf[6];
    for (int k = 0; k < 4; k++) g.pt_bounds[k] = g.pt_valid > 0 ? o2f_host(pf[2 + k]) : 0.f;
    g.n_parts = g.n_rings = g.n_pts = n;
  } else {
    g.xy.release();
    // (the type mask k_points gathered is dropped: k_scan gathers it again;
    // k_points reported nothing, it decoded point records only)
    SJ_CHECK(hipMemsetAsync(d_flags.p, 0, 4, s));
    DevArray<int64_t> cp(n, s), cr(n, s), cv(n, s);
    hipLaunchKernelGGL(k_scan, blocks_for(n), kBlock, 0, s, src, n, g.type.p, cp.p, cr.p, cv.p, d_err.p, d_flags.p);
    DevArray<int64_t> sp(n + 1, s), sr(n + 1, s), sv(n + 1, s);
    exclusive_scan_i64(cp.p, sp.p, n, s);
    exclusive_scan_i64(cr.p, sr.p, n, s);
    exclusive_scan_i64(cv.p, sv.p, n, s);
    int64_t tot[3];
    SJ_CHECK(hipMemcpyAsync(&tot[0], sp.p + n, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipMemcpyAsync(&tot[1], sr.p + n, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipMemcpyAsync(&tot[2], sv.p + n, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipMemcpyAsync(&herr, d_err.p, 8, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipMemcpyAsync(flags, d_flags.p, 4, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    if (herr == ~0ull) {
      g.n_parts = tot[0];
      g.n_rings = tot[1];
      g.n_pts = tot[2];
      g.part_off.alloc(n + 1, s);
      g.ring_off.alloc(g.n_parts + 1, s);
      g.pt_off.alloc(g.n_rings + 1, s);
      g.xy.alloc(g.n_pts > 0 ? g.n_pts : 1, s);
      DevArray<int64_t> ring_src(g.n_rings > 0 ? g.n_rings : 1, s);
      hipLaunchKernelGGL(k_fill, blocks_for(n), kBlock, 0, s, src, n, sp.p, sr.p, sv.p, g.part_off.p, g.ring_off.p,
                         g.pt_off.p, ring_src.p, d_err.p);
      SJ_CHECK(hipMemcpyAsync(g.part_off.p + n, &tot[0], 8, hipMemcpyHostToDevice, s));
      SJ_CHECK(hipMemcpyAsync(g.ring_off.p + g.n_parts, &tot[1], 8, hipMemcpyHostToDevice, s));
      SJ_CHECK(hipMemcpyAsync(g.pt_off.p + g.n_rings, &tot[2], 8, hipMemcpyHostToDevice, s));
      if (g.n_pts > 0)
        hipLaunchKernelGGL(k_coords, blocks_for(g.n_pts, kPtsPerThread), kBlock, 0, s, src.data, g.pt_off.p,
                           ring_src.p, g.n_rings, g.n_pts, g.xy.p);
      SJ_CHECK(hipStreamSynchronize(s));
    }
  }
  if (herr != ~0ull) {
    int64_t row = (int64_t)(herr >> 8);
    throw Error("WKB decode failed at row " + std::to_string(row) + ": " + decode_err_text((uint32_t)(herr & 0xff)));
  }
  g.type_mask = flags[0];
}

// ---------------------------------------------------------------------------
// Rects

namespace {

__global__ void k_rects_points(const double2* xy, const uint8_t* type, int64_t n, int nearest, double expand,
                               float4* out) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double2 p = xy[r];
  if (type[r] == 0 || p.x != p.x) {
    out[r] = make_float4(NAN, NAN, NAN, NAN);
    return;
  }
  out[r] = finish_rect(p.x, p.y, p.x, p.y, nearest != 0, expand);
}

// One wavefront per row: min/max over all of the row's points.
__global__ void k_rects_rows(GeomView g, int nearest, double expand, float4* out) {
  int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int lane = threadIdx.x & 63;
  if (r >= g.n_rows) return;
  int64_t p0 = g.part_off[r], p1 = g.part_off[r + 1];
  int64_t v0 = 0, v1 = 0;
  if (p0 < p1) {
    v0 = g.pt_off[g.ring_off[p0]];
    v1 = g.pt_off[g.ring_off[p1]];
  }
  double xl = INFINITY, yl = INFINITY, xh = -INFINITY, yh = -INFINITY;
  for (int64_t v = v0 + lane; v < v1; v += 64) {
    double2 p = g.xy[v];
    xl = fmin(xl, p.x); xh = fmax(xh, p.x);
    yl = fmin(yl, p.y); yh = fmax(yh, p.y);
  }
  for (int o = 32; o > 0; o >>= 1) {
    xl = fmin(xl, __shfl_xor(xl, o));
    yl = fmin(yl, __shfl_xor(yl, o));
    xh = fmax(xh, __shfl_xor(xh, o));
    yh = fmax(yh, __shfl_xor(yh, o));
  }
  if (lane == 0) {
    if (v1 == v0 || g.type[r] == 0) out[r] = make_float4(NAN, NAN, NAN, NAN);
    else out[r] = finish_rect(xl, yl, xh, yh, nearest != 0 && g.type[r] == kPoint, expand);
  }
}

// The reads of xy are random; the order is read and the output written
// once, as streams (non-temporal: they need not stay in the caches, which
// keeps more of them for the random reads).
__global__ void k_gather_points(const double2* xy, const uint32_t* order, int64_t n, double2* out) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint32_t o = __builtin_nontemporal_load(order + i);
    const double2 v = xy[o];
    d2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (d2v*)(out + i));
  }
}

}  // namespace

namespace {
__global__ void k_store_points_desc(PointsDesc d, PointsDesc* dst) { *dst = d; }
}  // namespace

void store_points_desc(const PointsDesc& d, PointsDesc* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_store_points_desc, 1, 1, 0, s, d, dst);
  SJ_CHECK(hipGetLastError());
}

void decode_points_desc(const PointsDesc* desc_dev, int64_t cap, double2* xy, uint8_t* type, unsigned long long* err,
                        hipStream_t s) {
  if (cap > 0) hipLaunchKernelGGL(k_points_desc, blocks_for(cap), kBlock, 0, s, desc_dev, cap, xy, type, err);
  SJ_CHECK(hipGetLastError());
}

void gather_points_into(const double2* xy, const uint32_t* order, int64_t n, double2* out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_points, blocks_for(n), kBlock, 0, s, xy, order, n, out);
  SJ_CHECK(hipGetLastError());
}

void rects_of_points_into(const double2* xy, const uint8_t* type, int64_t n, bool point_nearest, double expand,
                          float4* out, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_rects_points, blocks_for(n), kBlock, 0, s, xy, type, n, point_nearest ? 1 : 0, expand, out);
  SJ_CHECK(hipGetLastError());
}

void gather_points(const GeomColumn& probe, const uint32_t* order, int64_t n, GeomColumn& out, hipStream_t s) {
  if (!probe.points_only) throw Error("gather_points needs a point column");
  out = GeomColumn();
  out.n_rows = out.n_parts = out.n_rings = out.n_pts = n;
  out.points_only = true;
  out.type_mask = probe.type_mask;
  out.xy.alloc(n > 0 ? n : 1, s);
  if (n > 0) hipLaunchKernelGGL(k_gather_points, blocks_for(n), kBlock, 0, s, probe.xy.p, order, n, out.xy.p);
  SJ_CHECK(hipGetLastError());
}

void compute_rects(const GeomColumn& g, const RingIndex* ri, bool point_nearest, double expand,
                   DevArray<float4>& rects, hipStream_t s) {
  (void)ri;
  rects.alloc(g.n_rows > 0 ? g.n_rows : 1, s);
  if (g.n_rows == 0) return;
  if (g.points_only) {
    hipLaunchKernelGGL(k_rects_points, blocks_for(g.n_rows), kBlock, 0, s, g.xy.p, g.type.p, g.n_rows,
                       point_nearest ? 1 : 0, expand, rects.p);
  } else {
    int64_t threads = g.n_rows * 64;
    hipLaunchKernelGGL(k_rects_rows, (unsigned int)((threads + kBlock - 1) / kBlock), kBlock, 0, s, g.view(),
                       point_nearest ? 1 : 0, expand, rects.p);
  }
  SJ_CHECK(hipGetLastError());
}

}  // namespace sjhip
