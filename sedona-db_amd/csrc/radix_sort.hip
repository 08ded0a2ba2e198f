// radix_sort.hip -- a stable LSD radix sort of (u32 key, u32 value) pairs
// made of kernels only: no memset, no allocation and no host
// synchronisation, so that a captured HIP graph can replay it.
//
// The request loop's captured launch sequence orders a request's probe
// points by Morton key with it (up to 65536 keys; any n up to 2^24 is
// accepted).  Every counter a pass reads is written by the same pass, so a
// replay never depends on the previous replay's contents.
//
// One pass per 8-bit digit, three kernels per pass:
//   k_rs_count    one workgroup per tile of kTile keys: the tile's digit
//                 histogram in LDS, then stored (not added) as
//                 hist[d * T + t];
//   k_rs_scan     one workgroup: exclusive prefix sum of hist in place.  In
//                 digit-major order hist[d * T + t] becomes the first output
//                 slot of digit d's keys from tile t;
//   k_rs_scatter  one workgroup per tile.  Each wavefront ranks its 1024
//                 consecutive keys 64 at a time: lanes of one digit are
//                 matched with 8 ballots, a lane's rank is its digit's running
//                 count in LDS plus the matched lanes below it, and the
//                 lowest matched lane advances the count.  Rows go in order
//                 and lanes in order, so equal digits keep their input order
//                 (stable).  Wavefront w's slots follow wavefronts 0..w-1.
#include "sjhip.h"

namespace sjhip {

namespace {

constexpr int kRsThreads = 256;            // 4 wavefronts
constexpr int kRsRows = 16;                // rows of 64 keys per wavefront
constexpr int kTile = kRsThreads * kRsRows;  // 4096 keys per workgroup
constexpr int kRsScanThreads = 1024;

__global__ __launch_bounds__(kRsThreads) void k_rs_count(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                          uint32_t mask, uint32_t* __restrict__ hist, int64_t T) {
  __shared__ uint32_t h[256];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int i = t; i < kTile; i += kRsThreads) {
    const int64_t k = base + i;
    if (k < n) atomicAdd(&h[(keys[k] >> shift) & mask], 1u);
  }
  __syncthreads();
  hist[(int64_t)t * T + blockIdx.x] = h[t];
}

__global__ __launch_bounds__(kRsScanThreads) void k_rs_scan(uint32_t* __restrict__ hist, int64_t m) {
  __shared__ uint32_t part[kRsScanThreads];
  const int t = threadIdx.x;
  const int64_t per = (m + kRsScanThreads - 1) / kRsScanThreads;
  const int64_t b = (int64_t)t * per;
  const int64_t e = b + per < m ? b + per : m;
  uint32_t sum = 0;
  for (int64_t i = b; i < e; i++) sum += hist[i];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < kRsScanThreads; o <<= 1) {
    const uint32_t v = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (int64_t i = b; i < e; i++) {
    const uint32_t v = hist[i];
    hist[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint32_t* __restrict__ kin,
                                                            const uint32_t* __restrict__ vin,
                                                            uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                            int64_t n, int shift, uint32_t mask,
                                                            const uint32_t* __restrict__ hist, int64_t T) {
  __shared__ uint32_t cnt[kRsThreads / 64][256];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  for (int i = t; i < (kRsThreads / 64) * 256; i += kRsThreads) (&cnt[0][0])[i] = 0;
  const uint32_t base = hist[(int64_t)t * T + blockIdx.x];
  __syncthreads();
  const int64_t wbase = (int64_t)blockIdx.x * kTile + (int64_t)w * (kRsRows * 64);
  const uint64_t lanes_below = lane == 0 ? 0ull : ((1ull << lane) - 1ull);
  uint32_t key[kRsRows], val[kRsRows], rank[kRsRows];
#pragma unroll
  for (int r = 0; r < kRsRows; r++) {
    const int64_t k = wbase + r * 64 + lane;
    const bool ok = k < n;
    const uint32_t kk = ok ? kin[k] : 0u;
    const uint32_t vv = ok ? vin[k] : 0u;
    const uint32_t d = (kk >> shift) & mask;
    uint64_t m = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint64_t below = m & lanes_below;
    const uint32_t before = cnt[w][d];
    if (ok && below == 0) cnt[w][d] = before + (uint32_t)__popcll(m);
    key[r] = kk;
    val[r] = vv;
    rank[r] = before + (uint32_t)__popcll(below);
  }
  __syncthreads();
  {
    // thread t owns digit t: wavefront w's first slot for it
    uint32_t run = base;
#pragma unroll
    for (int ww = 0; ww < kRsThreads / 64; ww++) {
      const uint32_t c = cnt[ww][t];
      cnt[ww][t] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRsRows; r++) {
    const int64_t k = wbase + r * 64 + lane;
    if (k < n) {
      const uint32_t dst = cnt[w][(key[r] >> shift) & mask] + rank[r];
      kout[dst] = key[r];
      vout[dst] = val[r];
    }
  }
}

int64_t rs_tiles(int64_t n) { return (n + kTile - 1) / kTile; }

}  // namespace

size_t graph_sort_temp_bytes(int64_t n) {
  const size_t a = (size_t)((n * 4 + 255) / 256 * 256);
  return 2 * a + (size_t)rs_tiles(n) * 256 * 4 + 256;
}

void graph_sort_pairs(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout, int64_t n,
                      int begin_bit, int end_bit, void* temp, size_t temp_bytes, hipStream_t s) {
  if (n <= 0) return;
  if (n > (int64_t(1) << 24)) throw Error("graph_sort_pairs: at most 2^24 keys");
  if (begin_bit < 0 || end_bit > 32 || begin_bit >= end_bit) throw Error("graph_sort_pairs: bad bit range");
  if (temp_bytes < graph_sort_temp_bytes(n)) throw Error("graph_sort_pairs: temporary storage too small");
  const int64_t T = rs_tiles(n);
  const size_t a = (size_t)((n * 4 + 255) / 256 * 256);
  uint8_t* base = (uint8_t*)temp;
  uint32_t* ktmp = (uint32_t*)base;
  uint32_t* vtmp = (uint32_t*)(base + a);
  uint32_t* hist = (uint32_t*)(base + 2 * a);
  const int passes = (end_bit - begin_bit + 7) / 8;
  const uint32_t* ks = kin;
  const uint32_t* vs = vin;
  for (int p = 0; p < passes; p++) {
    const int shift = begin_bit + 8 * p;
    const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
    const uint32_t mask = (1u << bits) - 1u;
    // the last pass writes the output; the ones before alternate so that no
    // pass reads and writes one buffer
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint32_t* kd = to_out ? kout : ktmp;
    uint32_t* vd = to_out ? vout : vtmp;
    hipLaunchKernelGGL(k_rs_count, (unsigned)T, kRsThreads, 0, s, ks, n, shift, mask, hist, T);
    hipLaunchKernelGGL(k_rs_scan, 1, kRsScanThreads, 0, s, hist, T * 256);
    hipLaunchKernelGGL(k_rs_scatter, (unsigned)T, kRsThreads, 0, s, ks, vs, kd, vd, n, shift, mask, hist, T);
    SJ_CHECK(hipGetLastError());
    ks = kd;
    vs = vd;
  }
}

}  // namespace sjhip
