// wave_out.h -- wavefront-local output buffering of (build, probe) pairs.
//
// Kernels that emit pairs append them to a per-wavefront buffer in LDS
// (ballot + lane-mask prefix, no atomics) and flush a full buffer with a
// single global atomicAdd and coalesced stores.  One global atomic per
// kWaveOutCap pairs instead of one per wavefront step: a single output
// counter shared by the whole grid is otherwise the bottleneck.
//
// Every call must be made by all 64 lanes of the wavefront (converged
// control flow); the buffer count is wavefront-uniform.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sjhip {

constexpr int kWaveOutCap = 512;

struct WaveOut {
  uint32_t* sb;  // LDS, kWaveOutCap entries for this wavefront
  uint32_t* sp;
  uint32_t n;    // entries held (uniform across the wavefront)
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void wave_flush(WaveOut& w, uint32_t* ob, uint32_t* op, uint64_t cap,
                                           unsigned long long* counter) {
  if (w.n == 0) return;
  const int lane = threadIdx.x & 63;
  wave_lds_sync();
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(counter, (unsigned long long)w.n);
  base = __shfl(base, 0);
  for (uint32_t k = lane; k < w.n; k += 64) {
    unsigned long long d = base + k;
    if (d < cap) {
      ob[d] = w.sb[k];
      op[d] = w.sp[k];
    }
  }
  wave_lds_sync();
  w.n = 0;
}

// Appends (b, p) for every lane with keep set.
__device__ __forceinline__ void wave_append(WaveOut& w, bool keep, uint32_t b, uint32_t p, uint32_t* ob,
                                            uint32_t* op, uint64_t cap, unsigned long long* counter) {
  uint64_t m = __ballot(keep);
  uint32_t cnt = (uint32_t)__popcll(m);
  if (cnt == 0) return;
  if (w.n + cnt > (uint32_t)kWaveOutCap) wave_flush(w, ob, op, cap, counter);
  const int lane = threadIdx.x & 63;
  if (keep) {
    uint64_t below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
    uint32_t idx = w.n + (uint32_t)__popcll(below);
    w.sb[idx] = b;
    w.sp[idx] = p;
  }
  w.n += cnt;
}

// Appends up to 32 pairs per lane: the build ids ids[base + k] for every set
// bit k of hits, all with probe id p.
__device__ __forceinline__ void wave_append_mask(WaveOut& w, uint32_t hits, const uint32_t* ids, int64_t base,
                                                 uint32_t p, uint32_t* ob, uint32_t* op, uint64_t cap,
                                                 unsigned long long* counter) {
  const int lane = threadIdx.x & 63;
  uint32_t cnt = __popc(hits);
  uint32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  uint32_t total = __shfl(incl, 63);
  if (total == 0) return;
  if (w.n + total > (uint32_t)kWaveOutCap) wave_flush(w, ob, op, cap, counter);
  if (total > (uint32_t)kWaveOutCap) {
    // larger than a buffer: write through with one atomic
    unsigned long long g = 0;
    if (lane == 0) g = atomicAdd(counter, (unsigned long long)total);
    g = __shfl(g, 0);
    unsigned long long d = g + incl - cnt;
    while (hits) {
      int k = __ffs(hits) - 1;
      hits &= hits - 1;
      if (d < cap) {
        ob[d] = ids[base + k];
        op[d] = p;
      }
      d++;
    }
    return;
  }
  uint32_t idx = w.n + incl - cnt;
  while (hits) {
    int k = __ffs(hits) - 1;
    hits &= hits - 1;
    w.sb[idx] = ids[base + k];
    w.sp[idx] = p;
    idx++;
  }
  w.n += total;
}

}  // namespace sjhip
