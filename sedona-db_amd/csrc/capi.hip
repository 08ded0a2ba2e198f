// capi.hip -- the C ABI exporter: runtime, FP32 rect index, refiner and the
// extension entry points declared in include/gpuspatial_hip.h.
//
// Error convention of the reference exporter
// (c/sedona-libgpuspatial/libgpuspatial/src/gpuspatial_c.cc:45-58): every
// entry point returns 0, or EINVAL with the message kept in the object
// (probe errors in the context, gpuspatial_c.cc:191-214).  Every entry
// point makes the runtime's device current (gpuspatial_c.cc:247-253).
#include <errno.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "../../include/gpuspatial_hip.h"
#include "../../include/gpuspatial_hip_testing.h"
#include "sjhip.h"

namespace sjhip {

namespace {

thread_local std::string t_last_error;

template <typename F>
int guarded(std::string* err, F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    if (err) *err = e.what();
    t_last_error = e.what();
    return e.code ? e.code : EINVAL;
  } catch (const std::exception& e) {
    if (err) *err = e.what();
    t_last_error = e.what();
    return EINVAL;
  } catch (...) {
    if (err) *err = "unknown error";
    t_last_error = "unknown error";
    return EINVAL;
  }
}

struct Runtime {
  int device_id = 0;
  bool initialized = false;
  std::string last_error;
};

Runtime* runtime_of(const GpuSpatialRuntime* rt) {
  if (!rt || !rt->private_data) throw Error("runtime is null or released");
  Runtime* r = (Runtime*)rt->private_data;
  if (!r->initialized) throw Error("runtime is not initialized");
  return r;
}

void use_device(const Runtime* r) { SJ_CHECK(hipSetDevice(r->device_id)); }

unsigned char* noalloc_scratch(int dev);

// The Arrow C device interface's sync_event: for a ROCm array, a pointer to
// the producer's hipEvent_t, which must complete before the buffers are
// read.  The host waits for it before the call reads or queues the array.
void wait_sync_event(const ArrowDeviceArray* a) {
  if (a && a->device_type == ARROW_DEVICE_ROCM && a->sync_event)
    SJ_CHECK(hipEventSynchronize(*(hipEvent_t*)a->sync_event));
}

// Pinned host blocks for the host-facing calls (probe results handed to
// the callback, staging of caller arrays).  A call leases a block at least
// as large as it needs; the block returns to the pool when the lease ends
// and is kept for later calls, so that steady-state calls allocate no
// pinned memory (and free none: hipHostFree synchronises the device).
// Blocks are held until the process ends.
class PinnedPool {
 public:
  static PinnedPool& get() {
    static PinnedPool* p = new PinnedPool();  // never destroyed: blocks live until exit
    return *p;
  }
  void* take(size_t bytes, size_t* got) {
    std::lock_guard<std::mutex> g(mu_);
    // the smallest free block that fits
    Block* best = nullptr;
    for (auto& b : blocks_)
      if (!b.used && b.bytes >= bytes && (!best || b.bytes < best->bytes)) best = &b;
    if (!best) {
      size_t sz = std::max<size_t>(bytes, (size_t)1 << 20);
      sz = (sz + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
      void* p = nullptr;
      SJ_CHECK(hipHostMalloc(&p, sz, hipHostMallocDefault));
      blocks_.push_back(Block{p, sz, false});
      best = &blocks_.back();
    }
    best->used = true;
    *got = best->bytes;
    return best->p;
  }
  void give(void* p) {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& b : blocks_)
      if (b.p == p) b.used = false;
  }

 private:
  struct Block {
    void* p;
    size_t bytes;
    bool used;
  };
  std::mutex mu_;
  std::deque<Block> blocks_;  // (a deque keeps element addresses stable)
};

struct PinnedLease {
  void* p = nullptr;
  size_t bytes = 0;
  explicit PinnedLease(size_t n) { p = PinnedPool::get().take(n, &bytes); }
  ~PinnedLease() { PinnedPool::get().give(p); }
  PinnedLease(const PinnedLease&) = delete;
  PinnedLease& operator=(const PinnedLease&) = delete;
};

// Copies n bytes of pageable host memory to the device through pinned
// blocks: the CPU fills one chunk while the DMA of the previous one runs.
// Returns when every chunk has been transferred (the caller may then reuse
// or free src).
void upload_staged(void* dst, const void* src, size_t n, hipStream_t s) {
  if (n == 0) return;
  constexpr size_t kChunk = (size_t)8 << 20;
  if (n <= kChunk) {
    PinnedLease l(n);
    std::memcpy(l.p, src, n);
    SJ_CHECK(hipMemcpyAsync(dst, l.p, n, hipMemcpyHostToDevice, s));
    SJ_CHECK(hipStreamSynchronize(s));
    return;
  }
  PinnedLease a(kChunk), b(kChunk);
  void* buf[2] = {a.p, b.p};
  hipEvent_t done[2];
  SJ_CHECK(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
  SJ_CHECK(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
  bool pending[2] = {false, false};
  try {
    int k = 0;
    for (size_t off = 0; off < n; off += kChunk, k ^= 1) {
      const size_t len = std::min(kChunk, n - off);
      if (pending[k]) SJ_CHECK(hipEventSynchronize(done[k]));
      std::memcpy(buf[k], (const uint8_t*)src + off, len);
      SJ_CHECK(hipMemcpyAsync((uint8_t*)dst + off, buf[k], len, hipMemcpyHostToDevice, s));
      SJ_CHECK(hipEventRecord(done[k], s));
      pending[k] = true;
    }
    SJ_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(done[0]);
    (void)hipEventDestroy(done[1]);
    throw;
  }
  SJ_CHECK(hipEventDestroy(done[0]));
  SJ_CHECK(hipEventDestroy(done[1]));
}

// A stream owned for the duration of one call.
struct ScopedStream {
  hipStream_t s = nullptr;
  ScopedStream() { SJ_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~ScopedStream() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

// ---------------------------------------------------------------------------
// Runtime

int rt_init(GpuSpatialRuntime* self, GpuSpatialRuntimeConfig* cfg) {
  Runtime* r = (Runtime*)self->private_data;
  return guarded(&r->last_error, [&] {
    if (!cfg) throw Error("null runtime config");
    if (cfg->ptx_root && cfg->ptx_root[0]) {
      struct stat st;
      if (stat(cfg->ptx_root, &st) != 0)
        throw Error(std::string("cannot open shader directory '") + cfg->ptx_root + "': " + std::strerror(errno));
    }
    int n = 0;
    SJ_CHECK(hipGetDeviceCount(&n));
    if (cfg->device_id < 0 || cfg->device_id >= n)
      throw Error("device_id " + std::to_string(cfg->device_id) + " out of range (" + std::to_string(n) + " devices)");
    r->device_id = cfg->device_id;
    SJ_CHECK(hipSetDevice(r->device_id));
    (void)noalloc_scratch(r->device_id);
    if (cfg->use_cuda_memory_pool) {
      hipMemPool_t pool;
      SJ_CHECK(hipDeviceGetDefaultMemPool(&pool, r->device_id));
      size_t free_b = 0, total_b = 0;
      SJ_CHECK(hipMemGetInfo(&free_b, &total_b));
      int pct = std::min(100, std::max(0, cfg->cuda_memory_pool_init_precent));
      uint64_t keep = (uint64_t)((double)total_b * pct / 100.0);
      SJ_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    }
    r->initialized = true;
  });
}

void rt_release(GpuSpatialRuntime* self) {
  if (!self) return;
  delete (Runtime*)self->private_data;
  self->private_data = nullptr;
}

const char* rt_err(GpuSpatialRuntime* self) {
  return self && self->private_data ? ((Runtime*)self->private_data)->last_error.c_str() : "runtime released";
}

// ---------------------------------------------------------------------------
// FP32 rect index

// Join parameters chosen by SjhipAutotune, per kind of refine: 0 point in
// polygon, 1 distance, 2 other relations.  refine_order < 0: not tuned.
struct Tuning {
  int refine_order[3] = {-1, -1, -1};
  int pip_waves = 8;
};

struct IndexImpl {
  Runtime* rt = nullptr;
  std::string last_error;
  std::vector<float> host_rects;
  hipStream_t stream = nullptr;
  DevArray<float4> rects;
  RectTree tree;
  uint64_t tree_gen = 0;  // bumped whenever the tree is rebuilt (captured plans compare it)
  Tuning tuning;
  bool built = false;
  std::mutex mu;
};

struct IndexCtx {
  std::string last_error;
  std::vector<uint32_t> hb, hp;
};

IndexImpl* index_of(SedonaFloatIndex2D* self) {
  if (!self || !self->private_data) throw Error("index is null or released");
  return (IndexImpl*)self->private_data;
}

int ix_clear(SedonaFloatIndex2D* self) {
  IndexImpl* ix = (IndexImpl*)self->private_data;
  return guarded(&ix->last_error, [&] {
    use_device(ix->rt);
    ix->host_rects.clear();
    ix->tree = RectTree();
    ix->tree_gen++;
    ix->rects.release();
    ix->built = false;
  });
}

void ix_create_context(SedonaSpatialIndexContext* ctx) { ctx->private_data = new IndexCtx(); }
void ix_destroy_context(SedonaSpatialIndexContext* ctx) {
  delete (IndexCtx*)ctx->private_data;
  ctx->private_data = nullptr;
}

int ix_push_build(SedonaFloatIndex2D* self, const float* buf, uint32_t n) {
  IndexImpl* ix = (IndexImpl*)self->private_data;
  return guarded(&ix->last_error, [&] {
    if (ix->built) throw Error("push_build after finish_building; call clear first");
    if (n && !buf) throw Error("null rect buffer");
    ix->host_rects.insert(ix->host_rects.end(), buf, buf + 4 * (size_t)n);
  });
}

void index_finish(IndexImpl* ix) {
  use_device(ix->rt);
  int64_t n = (int64_t)(ix->host_rects.size() / 4);
  ix->rects.alloc(n > 0 ? n : 1, ix->stream);
  if (n)
    SJ_CHECK(hipMemcpyAsync(ix->rects.p, ix->host_rects.data(), n * sizeof(float4), hipMemcpyHostToDevice,
                            ix->stream));
  build_rect_tree(ix->rects.p, n, ix->tree, ix->stream);
  ix->tree_gen++;
  SJ_CHECK(hipStreamSynchronize(ix->stream));
  ix->built = true;
}

int ix_finish(SedonaFloatIndex2D* self) {
  IndexImpl* ix = (IndexImpl*)self->private_data;
  return guarded(&ix->last_error, [&] { index_finish(ix); });
}

int ix_probe(SedonaFloatIndex2D* self, SedonaSpatialIndexContext* context, const float* buf, uint32_t n,
             int (*cb)(const uint32_t*, const uint32_t*, uint32_t, void*), void* ud) {
  IndexImpl* ix = (IndexImpl*)self->private_data;
  IndexCtx* cx = context ? (IndexCtx*)context->private_data : nullptr;
  std::string* err = cx ? &cx->last_error : &ix->last_error;
  int cb_rc = 0;
  int rc = guarded(err, [&] {
    if (!ix->built) throw Error("probe before finish_building");
    if (!cb) throw Error("null probe callback");
    use_device(ix->rt);
    if (n == 0) return;
    ScopedStream ss;
    DevArray<float4> q(n, ss.s);
    upload_staged(q.p, buf, (size_t)n * sizeof(float4), ss.s);
    DevArray<uint32_t> order;
    int64_t nv = morton_order(q.p, n, order, ss.s);
    // point rects (the operator's rects of point rows, nulls as NaN) take
    // the grid filter
    bool points = true;
    for (uint32_t i = 0; i < n && points; i++) {
      const float* r = buf + 4 * (size_t)i;
      points = (r[0] == r[2] && r[1] == r[3]) || !(r[0] <= r[2] && r[1] <= r[3]);
    }
    PairBuf pb;
    uint64_t cnt = filter_pairs(ix->tree, q.p, nv, order.p, pb, ss.s, nullptr, false, points);
    // the pairs come back into pinned blocks that the callback reads in place
    PinnedLease lb(cnt * 4 + 4), lp(cnt * 4 + 4);
    uint32_t* hb = (uint32_t*)lb.p;
    uint32_t* hp = (uint32_t*)lp.p;
    if (cnt) {
      SJ_CHECK(hipMemcpyAsync(hb, pb.b.p, cnt * 4, hipMemcpyDeviceToHost, ss.s));
      SJ_CHECK(hipMemcpyAsync(hp, pb.p.p, cnt * 4, hipMemcpyDeviceToHost, ss.s));
    }
    SJ_CHECK(hipStreamSynchronize(ss.s));
    // deliver in chunks that fit the callback's u32 length
    const uint64_t chunk = 1ull << 30;
    for (uint64_t off = 0; off < cnt || (cnt == 0 && off == 0); off += chunk) {
      uint32_t len = (uint32_t)std::min<uint64_t>(chunk, cnt - off);
      cb_rc = cb(hb + off, hp + off, len, ud);
      if (cb_rc != 0 || cnt == 0) break;
    }
  });
  if (rc) return rc;
  return cb_rc;
}

const char* ix_err(SedonaFloatIndex2D* self) {
  return self && self->private_data ? ((IndexImpl*)self->private_data)->last_error.c_str() : "index released";
}
const char* ix_ctx_err(SedonaSpatialIndexContext* ctx) {
  return ctx && ctx->private_data ? ((IndexCtx*)ctx->private_data)->last_error.c_str() : "context released";
}

void ix_release(SedonaFloatIndex2D* self) {
  if (!self || !self->private_data) return;
  IndexImpl* ix = (IndexImpl*)self->private_data;
  try {
    use_device(ix->rt);
    ix->tree = RectTree();
    ix->tree_gen++;
    ix->rects.release();
    if (ix->stream) {
      (void)hipStreamSynchronize(ix->stream);
      (void)hipStreamDestroy(ix->stream);
    }
  } catch (...) {
  }
  delete ix;
  self->private_data = nullptr;
}

// ---------------------------------------------------------------------------
// Refiner

struct RefinerImpl {
  Runtime* rt = nullptr;
  std::string last_error;
  std::string build_format;
  hipStream_t stream = nullptr;
  // pushed build batches, decoded on the device as they arrive
  std::vector<GeomColumn> batches;
  GeomColumn build;
  RingIndex ring_index;
  KnnIndex knn;  // object tree of a point build side, built by the first KNN call
  std::mutex knn_mu;
  bool built = false;
  std::mutex mu;
  // device scratch of the joins against this build side, one workspace per
  // join running at a time, kept between joins
  std::mutex ws_mu;
  std::vector<std::unique_ptr<Workspace>> ws_free;
  uint64_t ring_gen = 0;  // bumped whenever the ring index is rebuilt (captured plans compare it)
  // refine precision of the ABI's refine calls: 0 certified FP32 with FP64
  // fallback, 1 the all-FP64 plain kernels
  std::atomic<int> precision{0};
};

// A join's workspace: taken from the refiner's free list (or new) for the
// join's stream, and given back when the join ends, once the stream has
// drained (on failure too), with the detached stream as its buffers' owner.
struct WorkspaceLease {
  RefinerImpl* rf;
  hipStream_t s;
  std::unique_ptr<Workspace> ws;
  WorkspaceLease(RefinerImpl* r, hipStream_t stream) : rf(r), s(stream) {
    {
      std::lock_guard<std::mutex> g(rf->ws_mu);
      if (!rf->ws_free.empty()) {
        ws = std::move(rf->ws_free.back());
        rf->ws_free.pop_back();
      }
    }
    if (!ws) ws.reset(new Workspace());
    ws->set_stream(s);
  }
  ~WorkspaceLease() {
    (void)hipStreamSynchronize(s);
    ws->set_stream(detached_stream());
    std::lock_guard<std::mutex> g(rf->ws_mu);
    rf->ws_free.push_back(std::move(ws));
  }
  WorkspaceLease(const WorkspaceLease&) = delete;
  WorkspaceLease& operator=(const WorkspaceLease&) = delete;
};

RefinerImpl* refiner_of(SedonaSpatialRefiner* self) {
  if (!self || !self->private_data) throw Error("refiner is null or released");
  return (RefinerImpl*)self->private_data;
}

int rf_clear(SedonaSpatialRefiner* self) {
  RefinerImpl* rf = (RefinerImpl*)self->private_data;
  return guarded(&rf->last_error, [&] {
    use_device(rf->rt);
    rf->batches.clear();
    rf->build = GeomColumn();
    rf->ring_index = RingIndex();
    rf->ring_gen++;
    rf->knn = KnnIndex();
    rf->built = false;
  });
}

int rf_init_schema(SedonaSpatialRefiner* self, const ArrowSchema* schema) {
  RefinerImpl* rf = (RefinerImpl*)self->private_data;
  return guarded(&rf->last_error, [&] {
    if (!schema || !schema->format) throw Error("null build schema");
    std::string f = schema->format;
    if (f != "z" && f != "Z" && f != "vz")
      throw Error("unsupported build storage type '" + f + "' (expected Binary, LargeBinary or BinaryView WKB)");
    rf->build_format = f;
  });
}

void refiner_push(RefinerImpl* rf, const WkbInput& in) {
  use_device(rf->rt);
  if (rf->built) throw Error("push_build after finish_building; call clear first");
  GeomColumn g;
  decode_wkb(in, g, rf->stream);
  rf->batches.push_back(std::move(g));
}

int rf_push_build(SedonaSpatialRefiner* self, const ArrowArray* array) {
  RefinerImpl* rf = (RefinerImpl*)self->private_data;
  return guarded(&rf->last_error, [&] {
    if (rf->build_format.empty()) throw Error("init_build_schema must be called before push_build");
    ArrowSchema sch{};
    sch.format = rf->build_format.c_str();
    refiner_push(rf, wkb_input_from_arrow(&sch, array, false));
  });
}

// Concatenation of decoded batches (build ids are global in push order).
__global__ void k_shift_copy(const int64_t* in, int64_t n, int64_t delta, int64_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] + delta;
}
__global__ void k_iota64(int64_t n, int64_t base, int64_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = base + i;
}

void concat_columns(std::vector<GeomColumn>& parts, GeomColumn& out, hipStream_t s) {
  if (parts.size() == 1) {
    out = std::move(parts[0]);
    parts.clear();
    return;
  }
  out = GeomColumn();
  bool all_points = true;
  for (auto& g : parts) {
    out.n_rows += g.n_rows;
    all_points &= g.points_only;
    out.type_mask |= g.type_mask;
  }
  out.type.alloc(out.n_rows > 0 ? out.n_rows : 1, s);
  int64_t r = 0;
  for (auto& g : parts) {
    if (g.n_rows) SJ_CHECK(hipMemcpyAsync(out.type.p + r, g.type.p, g.n_rows, hipMemcpyDeviceToDevice, s));
    r += g.n_rows;
  }
  if (all_points) {
    out.points_only = true;
    out.n_parts = out.n_rings = out.n_pts = out.n_rows;
    out.xy.alloc(out.n_rows > 0 ? out.n_rows : 1, s);
    r = 0;
    for (auto& g : parts) {
      if (g.n_rows) SJ_CHECK(hipMemcpyAsync(out.xy.p + r, g.xy.p, g.n_rows * 16, hipMemcpyDeviceToDevice, s));
      r += g.n_rows;
    }
  } else {
    // points-only batches are viewed as one part / ring / point per row;
    // their NaN (empty) rows keep one NaN point, which every predicate
    // treats as empty
    for (auto& g : parts) {
      out.n_parts += g.n_parts;
      out.n_rings += g.n_rings;
      out.n_pts += g.n_pts;
    }
    out.part_off.alloc(out.n_rows + 1, s);
    out.ring_off.alloc(out.n_parts + 1, s);
    out.pt_off.alloc(out.n_rings + 1, s);
    out.xy.alloc(out.n_pts > 0 ? out.n_pts : 1, s);
    int64_t rr = 0, pp = 0, gg = 0, vv = 0;
    for (auto& g : parts) {
      unsigned int b;
      if (g.points_only) {
        b = (unsigned int)((g.n_rows + 255) / 256);
        if (g.n_rows) {
          hipLaunchKernelGGL(k_iota64, b, 256, 0, s, g.n_rows, pp, out.part_off.p + gg);
          hipLaunchKernelGGL(k_iota64, b, 256, 0, s, g.n_rows, rr, out.ring_off.p + pp);
          hipLaunchKernelGGL(k_iota64, b, 256, 0, s, g.n_rows, vv, out.pt_off.p + rr);
        }
      } else {
        if (g.n_rows)
          hipLaunchKernelGGL(k_shift_copy, (unsigned int)((g.n_rows + 255) / 256), 256, 0, s, g.part_off.p, g.n_rows,
                             pp, out.part_off.p + gg);
        if (g.n_parts)
          hipLaunchKernelGGL(k_shift_copy, (unsigned int)((g.n_parts + 255) / 256), 256, 0, s, g.ring_off.p,
                             g.n_parts, rr, out.ring_off.p + pp);
        if (g.n_rings)
          hipLaunchKernelGGL(k_shift_copy, (unsigned int)((g.n_rings + 255) / 256), 256, 0, s, g.pt_off.p, g.n_rings,
                             vv, out.pt_off.p + rr);
      }
      if (g.n_pts) SJ_CHECK(hipMemcpyAsync(out.xy.p + vv, g.xy.p, g.n_pts * 16, hipMemcpyDeviceToDevice, s));
      gg += g.n_rows;
      pp += g.n_parts;
      rr += g.n_rings;
      vv += g.n_pts;
    }
    SJ_CHECK(hipMemcpyAsync(out.part_off.p + out.n_rows, &out.n_parts, 8, hipMemcpyHostToDevice, s));
    SJ_CHECK(hipMemcpyAsync(out.ring_off.p + out.n_parts, &out.n_rings, 8, hipMemcpyHostToDevice, s));
    SJ_CHECK(hipMemcpyAsync(out.pt_off.p + out.n_rings, &out.n_pts, 8, hipMemcpyHostToDevice, s));
  }
  SJ_CHECK(hipGetLastError());
  SJ_CHECK(hipStreamSynchronize(s));
  parts.clear();
}

void refiner_finish(RefinerImpl* rf) {
  use_device(rf->rt);
  if (rf->batches.empty()) {
    rf->build = GeomColumn();
    rf->build.type.alloc(1, rf->stream);
    rf->build.points_only = true;
    rf->build.xy.alloc(1, rf->stream);
  } else {
    concat_columns(rf->batches, rf->build, rf->stream);
  }
  RingIndexParams prm;
  rf->knn = KnnIndex();
  build_ring_index(rf->build, rf->ring_index, prm, rf->stream);
  rf->ring_gen++;
  SJ_CHECK(hipStreamSynchronize(rf->stream));
  rf->built = true;
}

int rf_finish(SedonaSpatialRefiner* self) {
  RefinerImpl* rf = (RefinerImpl*)self->private_data;
  return guarded(&rf->last_error, [&] { refiner_finish(rf); });
}

// Shared body of refine / refine-distance over host candidate arrays.
void refine_host(RefinerImpl* rf, const ArrowSchema* probe_schema, const ArrowArray* probe_array, int pred,
                 double distance, uint32_t* bidx, uint32_t* pidx, uint32_t n, uint32_t* new_n) {
  if (!rf->built) throw Error("refine before finish_building");
  if (!new_n) throw Error("null new_indices_size");
  use_device(rf->rt);
  if (n == 0) {
    *new_n = 0;
    return;
  }
  if (!bidx || !pidx) throw Error("null candidate index arrays");
  ScopedStream ss, ss2;
  hipStream_t s = ss.s;
  // two slots: a helper thread uploads the candidate arrays on its own
  // stream while this thread decodes the probe batch
  DevArray<uint32_t> cb, cp;
  std::exception_ptr up_err;
  std::thread up([&] {
    try {
      use_device(rf->rt);
      cb.alloc(n, ss2.s);
      cp.alloc(n, ss2.s);
      upload_staged(cb.p, bidx, (size_t)n * 4, ss2.s);  // (returns with ss2's work complete)
      upload_staged(cp.p, pidx, (size_t)n * 4, ss2.s);
    } catch (...) {
      up_err = std::current_exception();
    }
  });
  GeomColumn probe;
  std::exception_ptr dec_err;
  try {
    decode_wkb(wkb_input_from_arrow(probe_schema, probe_array, false), probe, s);
  } catch (...) {
    dec_err = std::current_exception();
  }
  up.join();
  // the arrays are released on the call's stream (ss2 ends first)
  cb.s = cp.s = s;
  if (dec_err) std::rethrow_exception(dec_err);
  if (up_err) std::rethrow_exception(up_err);
  if (pred != 100) check_relation_types(rf->build.type_mask, probe.type_mask, pred);
  // range checks on the device, before any kernel indexes with them
  check_pair_ranges(cb.p, cp.p, n, (uint64_t)rf->build.n_rows, (uint64_t)probe.n_rows, s);
  PairBuf out;
  uint64_t kept;
  const int precision = rf->precision.load();
  if (pred == 100)
    kept = refine_distance(rf->build, rf->ring_index, probe, distance, cb.p, cp.p, n, out, s, nullptr, precision);
  else
    kept = refine_relation(rf->build, rf->ring_index, probe, pred, cb.p, cp.p, n, precision, out, nullptr, s);
  sort_pairs(out.b.p, out.p.p, kept, s);
  if (kept) {
    // through pinned memory, then into the caller's arrays
    PinnedLease l(kept * 8);
    uint32_t* h = (uint32_t*)l.p;
    SJ_CHECK(hipMemcpyAsync(h, out.b.p, kept * 4, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipMemcpyAsync(h + kept, out.p.p, kept * 4, hipMemcpyDeviceToHost, s));
    SJ_CHECK(hipStreamSynchronize(s));
    std::memcpy(bidx, h, kept * 4);
    std::memcpy(pidx, h + kept, kept * 4);
  }
  SJ_CHECK(hipStreamSynchronize(s));
  *new_n = (uint32_t)kept;
}

int rf_refine(SedonaSpatialRefiner* self, const ArrowSchema* probe_schema, const ArrowArray* probe_array,
              SedonaSpatialRelationPredicate pred, uint32_t* bidx, uint32_t* pidx, uint32_t n, uint32_t* new_n) {
  RefinerImpl* rf = (RefinerImpl*)self->private_data;
  std::string err;
  int rc = guarded(&err, [&] { refine_host(rf, probe_schema, probe_array, (int)pred, 0.0, bidx, pidx, n, new_n); });
  if (rc) {
    std::lock_guard<std::mutex> g(rf->mu);
    rf->last_error = err;
  }
  return rc;
}

const char* rf_err(SedonaSpatialRefiner* self) {
  return self && self->private_data ? ((RefinerImpl*)self->private_data)->last_error.c_str() : "refiner released";
}

void rf_release(SedonaSpatialRefiner* self) {
  if (!self || !self->private_data) return;
  RefinerImpl* rf = (RefinerImpl*)self->private_data;
  try {
    use_device(rf->rt);
    rf->batches.clear();
    rf->build = GeomColumn();
    rf->ring_index = RingIndex();
    rf->ring_gen++;
    rf->knn = KnnIndex();
    if (rf->stream) {
      (void)hipStreamSynchronize(rf->stream);
      (void)hipStreamDestroy(rf->stream);
    }
  } catch (...) {
  }
  delete rf;
  self->private_data = nullptr;
}

// ---------------------------------------------------------------------------
// Device join result

struct JoinResultImpl {
  PairBuf pairs;
  int device_id = 0;
};

void jr_release(SjhipJoinResult* r) {
  if (!r || !r->private_data) return;
  JoinResultImpl* j = (JoinResultImpl*)r->private_data;
  (void)hipSetDevice(j->device_id);
  delete j;
  r->private_data = nullptr;
  r->build_indices = r->probe_indices = nullptr;
  r->size = 0;
}

// Plain 16-byte vector copy (bandwidth calibration).
__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

// Fused device join used by the extension entry point and the request loop.
struct JoinEngine {
  static void run(IndexImpl* ix, RefinerImpl* rf, const WkbInput& in, const SjhipJoinOptions& o, SjhipJoinResult* out);
};

void JoinEngine::run(IndexImpl* ix, RefinerImpl* rf, const WkbInput& in, const SjhipJoinOptions& o,
                     SjhipJoinResult* out) {
  if (!ix->built) throw Error("index is not built");
  if (!rf->built) throw Error("refiner is not built");
  use_device(rf->rt);
  std::unique_ptr<ScopedStream> own;
  hipStream_t s = (hipStream_t)o.stream;
  if (!s) {
    own.reset(new ScopedStream());
    s = own->s;
  }
  // stage events: decode+rects, Morton order, filter kernel, refine
  // kernels, output sort
  struct Events {
    hipEvent_t e[10] = {};
    bool on = false;
    ~Events() {
      for (auto& x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  ev.on = o.profile != 0;
  if (ev.on)
    for (auto& x : ev.e) SJ_CHECK(hipEventCreate(&x));
  auto mark = [&](int i) {
    if (ev.on) SJ_CHECK(hipEventRecord(ev.e[i], s));
  };
  KernelTimer ft, rt;
  if (ev.on) {
    ft.start = ev.e[4]; ft.stop = ev.e[5];
    rt.start = ev.e[6]; rt.stop = ev.e[7];
  }
  mark(0);
  GeomColumn probe;
  decode_wkb(in, probe, s);
  const bool dwithin = o.predicate == 100;
  if (!dwithin) check_relation_types(rf->build.type_mask, probe.type_mask, o.predicate);
  DevArray<float4> q;
  // refine order: the caller's, or (refine_order < 0) the tuned one, else
  // the default of the refine kind
  const bool pip_kind = !dwithin && pip_refine_applies(rf->build, rf->ring_index, probe, o.predicate);
  const int kind = dwithin ? 1 : pip_kind && o.sort_probes ? 0 : 2;
  int order_used = o.refine_order;
  if (order_used < 0) order_used = ix->tuning.refine_order[kind];
  // ST_DWithin of Morton-ordered point probes in build-row-major order:
  // the build-rect pass (with the build side's rasters, if any)
  const bool dist_pm_ok = dwithin && o.sort_probes && probe.pt_valid >= 0 && probe.points_only &&
                          ix->tree.filter_mode == 0 && o.precision != 1 && o.distance >= 0 &&
                          o.distance < INFINITY && !rf->build.points_only && rf->ring_index.desc.p != nullptr;
  // default: build-row-major; DWithin in filter order when neither the
  // build-rect pass nor line rasters (whose split leaves few enough
  // candidates to sort) apply
  if (order_used < 0) order_used = dwithin && !dist_pm_ok && rf->ring_index.n_line_rasters == 0 ? 0 : 1;
  const bool dist_pm = dist_pm_ok && order_used == 1;
  const int dist_raster =
      dist_pm ? dist_raster_kind(rf->build, rf->ring_index, probe, o.distance, o.precision) : 0;
  // point probes of the point-in-polygon refine (and of dist_pm) are
  // ordered by the Morton keys of their point rects, computed from the
  // points (the rects are made below only where a filter reads them)
  const bool point_keys = o.sort_probes && probe.pt_valid >= 0 && (pip_kind || dist_pm);
  // probe rects: GPU operator semantics (points nearest) for relation
  // predicates; outward + distance expansion for DWithin
  auto make_rects = [&] {
    if (!q.p) compute_rects(probe, nullptr, !dwithin, dwithin ? o.distance : NAN, q, s);
  };
  if (!point_keys) make_rects();
  mark(1);
  DevArray<uint32_t> order, mkeys;
  MortonFrame frame;
  int64_t nq = probe.n_rows;
  const uint32_t* ord = nullptr;
  if (o.sort_probes) {
    nq = point_keys ? morton_order_points(probe, order, s, &frame, &mkeys)
                    : morton_order(q.p, probe.n_rows, order, s, &frame, &mkeys);
    ord = order.p;
  }
  // Point probes refined by the point-in-polygon kernel are gathered into
  // Morton order: candidates then name positions in that order, so that the
  // points of one polygon's candidates lie together in memory.  The refine
  // maps kept positions back to probe rows.
  const bool ranked = ord != nullptr && pip_kind;
  GeomColumn rprobe;
  if (ranked || (dist_pm && ord)) gather_points(probe, ord, nq, rprobe, s);
  const GeomColumn& rp = ranked ? rprobe : probe;
  mark(2);
  std::unique_ptr<JoinResultImpl> jr(new JoinResultImpl());
  jr->device_id = rf->rt->device_id;
  RefineStats st;
  uint64_t kept = 0, nc = 0;
  // refine_order 2 (DWithin over point probes): filter and distance refine
  // in one kernel, candidates never written out
  const bool fused = dwithin && probe.points_only && o.precision != 1 && order_used == 2;
  if (fused) {
    kept = filter_refine_distance(ix->tree, q.p, nq, ord, rf->build, rf->ring_index, probe, o.distance, jr->pairs, &nc,
                                  &st, s, ev.on ? &ft : nullptr);
    mark(3);
    mark(6);
    mark(7);
  } else {
    // candidates and the refine's output live in the workspace; the
    // result gets the kept pairs at their exact size
    WorkspaceLease lease(rf, s);
    Workspace* ws = lease.ws.get();
    uint64_t n_open = 0;
    bool split = false, grouped = false;
    PmSplitResult pr;
    // point in polygon over Morton-ordered points, polygon-major order:
    // each build rect tests the points of the Morton cells it covers, so
    // the candidates come out grouped by build row, and the cell-raster
    // split is decided in the same pass (no regrouping sort)
    if (order_used == 1 && ranked && ix->tree.filter_mode == 0) {
      const RingIndex& ri = rf->ring_index;
      const bool pm_raster = ri.n_rasters > 0 && o.precision == 0 && !rf->build.points_only && ri.ring_rect.p;
      if (ev.on) ft.record_start(s);
      grouped = pm_split(ix->tree, ri, rprobe.xy.p, mkeys.p, nq, frame, ord, pm_raster, ws->out,
                         pm_raster ? ws->open : ws->cand, ws->pm, &pr, s);
      if (ev.on) ft.record_stop(s);
      if (grouped) {
        nc = pr.candidates;
        split = pm_raster;
        n_open = pr.open;
      }
    } else if (dist_pm && nq > 0) {
      // the line rasters' drops decided in the pass; the open candidates
      // (build row, Morton position) grouped by line
      if (ev.on) ft.record_start(s);
      grouped = pm_split(ix->tree, rf->ring_index, rprobe.xy.p, mkeys.p, nq, frame, ord, false, ws->out, ws->open,
                         ws->pm, &pr, s, o.distance, dist_raster);
      if (ev.on) ft.record_stop(s);
      if (grouped) {
        nc = pr.candidates;
        split = true;
        n_open = pr.open;
      }
    }
    if (!grouped) {
      make_rects();
      nc = filter_pairs(ix->tree, q.p, nq, ord, ws->cand, s, ev.on ? &ft : nullptr, ranked,
                        probe.points_only && !dwithin, ranked ? rprobe.xy.p : nullptr);
    }
    mark(3);
    // point in polygon, polygon-major order: candidates in clean raster
    // cells are decided first, so that the regrouping sorts only the rest
    // (for DWithin over lines, the line rasters drop candidates first)
    if (!grouped)
      split = order_used == 1 &&
              (dwithin ? raster_split_distance(rf->build, rf->ring_index, probe, o.distance, o.precision,
                                               ws->cand.b.p, ws->cand.p.p, nc, ws->open, ws->mask, &n_open, s)
                       : raster_split(rf->build, rf->ring_index, rp, o.predicate, o.precision, ws->cand.b.p,
                                      ws->cand.p.p, nc, ranked ? ord : nullptr, ws->out, ws->open, ws->mask, &n_open,
                                      s));
    PairBuf& cand = split ? ws->open : ws->cand;
    const uint64_t n_cand = split ? n_open : nc;
    // polygon-major order: one wavefront then walks few distinct rings
    if (order_used == 1 && !grouped) sort_pairs_by_build(cand, n_cand, rf->build.n_rows, s);
    if (dwithin)
      kept = refine_distance(rf->build, rf->ring_index, grouped ? rprobe : probe, o.distance, cand.b.p, cand.p.p,
                             n_cand, ws->out, s, ev.on ? &rt : nullptr, o.precision, &st, ws,
                             split && (!grouped || dist_raster == 1), grouped ? ord : nullptr,
                             grouped && dist_raster == 2);
    else
      kept = refine_relation(rf->build, rf->ring_index, rp, o.predicate, cand.b.p, cand.p.p, n_cand, o.precision,
                             ws->out, &st, s, ev.on ? &rt : nullptr, ranked ? ord : nullptr, ix->tuning.pip_waves,
                             ws, split);
    mark(8);
    if (kept > 0 && kept * 4 >= ws->out.cap()) {
      // the kept pairs fill a good part of the workspace's output: the
      // result takes those buffers (the workspace gets new ones from the
      // pool on its next join) instead of a copy
      jr->pairs.b = std::move(ws->out.b);
      jr->pairs.p = std::move(ws->out.p);
    } else {
      jr->pairs.b.alloc(kept > 0 ? kept : 1, s);
      jr->pairs.p.alloc(kept > 0 ? kept : 1, s);
      if (kept) {
        SJ_CHECK(hipMemcpyAsync(jr->pairs.b.p, ws->out.b.p, kept * 4, hipMemcpyDeviceToDevice, s));
        SJ_CHECK(hipMemcpyAsync(jr->pairs.p.p, ws->out.p.p, kept * 4, hipMemcpyDeviceToDevice, s));
      }
    }
  }
  if (fused) mark(8);
  if (o.sort_output) sort_pairs(jr->pairs.b.p, jr->pairs.p.p, kept, s);
  mark(9);
  SJ_CHECK(hipStreamSynchronize(s));
  if (ev.on) {
    float ms;
    SJ_CHECK(hipEventElapsedTime(&ms, ev.e[0], ev.e[1])); out->stage_ms[0] = ms;
    SJ_CHECK(hipEventElapsedTime(&ms, ev.e[1], ev.e[2])); out->stage_ms[1] = ms;
    if (nc || nq) { SJ_CHECK(hipEventElapsedTime(&ms, ev.e[4], ev.e[5])); out->stage_ms[2] = ms; }
    if (nc) { SJ_CHECK(hipEventElapsedTime(&ms, ev.e[6], ev.e[7])); out->stage_ms[3] = ms; }
    SJ_CHECK(hipEventElapsedTime(&ms, ev.e[8], ev.e[9])); out->stage_ms[4] = ms;
    if (nc) { SJ_CHECK(hipEventElapsedTime(&ms, ev.e[3], ev.e[6])); out->stage_ms[5] = ms; }
  }
  // the result outlives the call's stream: all its work is complete, so its
  // buffers are released on the detached stream later
  jr->pairs.b.s = jr->pairs.p.s = jr->pairs.counter.s = detached_stream();
  out->build_indices = jr->pairs.b.p;
  out->probe_indices = jr->pairs.p.p;
  out->size = kept;
  out->candidates = nc;
  out->segment_tests = st.segment_tests;
  out->fallback_tests = st.fallback_tests;
  out->refine_candidates = st.walked;
  out->release = jr_release;
  out->private_data = jr.release();
}


// ---------------------------------------------------------------------------
// Request loop over a resident index (continuous batching)
//
// `slots` worker threads, each with its own stream, take requests in
// arrival order.  A request is admitted when a worker is free and its
// device-memory estimate fits the budget next to the requests already
// running (a request larger than the whole budget runs alone).  Admitted
// requests run their join concurrently on the device; a finished request
// frees its slot and its share of the budget for the next.

// A captured per-request launch sequence for point probes of one size
// bucket (capacity B rows, C candidate pairs), replayed for every request
// of at most B rows.  The request's column is described in device memory
// (PointsDesc), the counts are read on the device, every buffer has a
// fixed capacity, and the counters are cleared by a kernel, so the
// sequence has no host synchronisation, allocation or memset.  After the
// replay the host reads the counters: a decode error, a row that is not a
// point record or more candidates than C send the request down the eager
// path (JoinEngine::run) instead.
struct GraphPlan {
  static constexpr int64_t kMinRows = 4096, kMaxRows = 1 << 16;
  int kind = 0;  // 0 point in polygon, 1 distance
  int predicate = 0;
  double distance = 0;
  int pip_waves = 8;
  int leaf = 8;
  uint64_t tree_gen = 0;  // generations of the index tree and of the ring index the plan was captured over
  uint64_t ring_gen = 0;
  int device_id = 0;
  int64_t B = 0;
  uint64_t C = 0;
  DevArray<PointsDesc> desc;
  GeomColumn pcol, rcol;  // the decoded probe points, and in Morton order
  DevArray<float4> rects;
  DevArray<uint32_t> keys, ids, keys_sorted, order;
  DevArray<uint8_t> sort_tmp;
  size_t sort_bytes = 0;
  DevArray<uint32_t> cb, cp, redo, ob, op;
  DevArray<unsigned long long> cnt;  // [0] error, [1] candidates, [2] deferred, [3] kept, [4..5] refine stats
  PointsDesc* host_desc = nullptr;   // pinned
  unsigned long long* host_cnt = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~GraphPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (host_desc) (void)hipHostFree(host_desc);
    if (host_cnt) (void)hipHostFree(host_cnt);
  }
};

// The kind of fixed launch sequence (kernels over fixed-capacity buffers,
// no host synchronisation) that can join a probe batch: 0 point in polygon,
// 1 distance, or -1 when the batch needs the general engine (host or
// BinaryView buffers, nulls, the all-FP64 kernels, other predicates or
// build types).
int fixed_sequence_kind(const RefinerImpl* rf, const WkbInput& in, const SjhipJoinOptions& o) {
  if (!in.on_device || in.view || in.validity || in.length <= 0 || o.precision != 0) return -1;
  const GeomColumn& b = rf->build;
  if (o.predicate == 100) return b.n_rows > 0 ? 1 : -1;
  const uint32_t polys = (1u << kPolygon) | (1u << kMultiPolygon);
  const bool pip = !b.points_only && rf->ring_index.ring_rect.p && rf->ring_index.desc.p &&
                   (b.type_mask & ~polys) == 0 &&
                   (o.predicate == SedonaSpatialPredicateIntersects || o.predicate == SedonaSpatialPredicateCovers ||
                    o.predicate == SedonaSpatialPredicateContains);
  return pip ? 0 : -1;
}

// Captures are made one at a time in the process: a plan's pinned host
// buffers are allocated, and its sequence captured and instantiated, while
// no other thread captures.
std::mutex& capture_mutex() {
  static std::mutex mu;
  return mu;
}

// Throws unless every node of a captured graph is a kernel or a copy.  A
// memset node does not reliably clear its memory again when a graph is
// replayed, and an allocation node would make replays allocate.
void check_graph_nodes(hipGraph_t graph) {
  size_t nn = 0;
  SJ_CHECK(hipGraphGetNodes(graph, nullptr, &nn));
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn) SJ_CHECK(hipGraphGetNodes(graph, nodes.data(), &nn));
  for (size_t i = 0; i < nn; i++) {
    hipGraphNodeType ty;
    SJ_CHECK(hipGraphNodeGetType(nodes[i], &ty));
    if (ty != hipGraphNodeTypeKernel && ty != hipGraphNodeTypeMemcpy && ty != hipGraphNodeTypeEmpty)
      throw Error("captured request sequence holds a node of type " + std::to_string((int)ty) +
                  " (only kernels and copies may be replayed)");
  }
}

// Builds and captures a plan on stream s (the stream's work is complete on
// return; the plan's buffers are freed on the detached stream).
std::unique_ptr<GraphPlan> capture_plan(IndexImpl* ix, RefinerImpl* rf, int kind, const SjhipJoinOptions& o,
                                        int64_t B, uint64_t C, hipStream_t s) {
  std::lock_guard<std::mutex> capture_lock(capture_mutex());
  std::unique_ptr<GraphPlan> g(new GraphPlan());
  g->kind = kind;
  g->predicate = o.predicate;
  g->distance = o.distance;
  g->pip_waves = ix->tuning.pip_waves;
  g->leaf = ix->tree.leaf;
  g->tree_gen = ix->tree_gen;
  g->ring_gen = rf->ring_gen;
  g->device_id = rf->rt->device_id;
  g->B = B;
  g->C = C;
  g->desc.alloc(1, s);
  g->pcol.n_rows = g->pcol.n_parts = g->pcol.n_rings = g->pcol.n_pts = B;
  g->pcol.points_only = true;
  g->pcol.type_mask = 1u << kPoint;
  g->pcol.xy.alloc(B, s);
  g->pcol.type.alloc(B, s);
  g->rects.alloc(B, s);
  g->keys.alloc(B, s);
  g->ids.alloc(B, s);
  g->keys_sorted.alloc(B, s);
  g->order.alloc(B, s);
  g->sort_bytes = morton_sort_temp_bytes(B);
  g->sort_tmp.alloc(g->sort_bytes + 16, s);
  if (kind == 0) {
    g->rcol.n_rows = g->rcol.n_parts = g->rcol.n_rings = g->rcol.n_pts = B;
    g->rcol.points_only = true;
    g->rcol.type_mask = 1u << kPoint;
    g->rcol.xy.alloc(B, s);
    g->redo.alloc(C, s);
  }
  g->cb.alloc(C, s);
  g->cp.alloc(C, s);
  g->ob.alloc(C, s);
  g->op.alloc(C, s);
  g->cnt.alloc(8, s);
  SJ_CHECK(hipHostMalloc((void**)&g->host_desc, sizeof(PointsDesc), hipHostMallocDefault));
  SJ_CHECK(hipHostMalloc((void**)&g->host_cnt, 8 * sizeof(unsigned long long), hipHostMallocDefault));
  SJ_CHECK(hipStreamSynchronize(s));
  const bool dwithin = kind == 1;
  SJ_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    fill_u64(g->cnt.p, 1, ~0ull, s);
    fill_u64(g->cnt.p + 1, 7, 0ull, s);
    decode_points_desc(g->desc.p, B, g->pcol.xy.p, g->pcol.type.p, g->cnt.p, s);
    rects_of_points_into(g->pcol.xy.p, g->pcol.type.p, B, !dwithin, dwithin ? o.distance : NAN, g->rects.p, s);
    morton_order_fixed(ix->tree, g->rects.p, B, g->keys.p, g->ids.p, g->keys_sorted.p, g->order.p, g->sort_tmp.p,
                       g->sort_bytes, s);
    if (!dwithin) {
      gather_points_into(g->pcol.xy.p, g->order.p, B, g->rcol.xy.p, s);
      launch_filter(ix->tree, g->rects.p, B, g->order.p, true, g->cb.p, g->cp.p, C, g->cnt.p + 1, s, true);
      launch_pip_refine_dev(rf->build, rf->ring_index, g->rcol, o.predicate, g->cb.p, g->cp.p, g->cnt.p + 1, C,
                            g->order.p, g->redo.p, g->cnt.p + 2, g->ob.p, g->op.p, C, g->cnt.p + 3, g->cnt.p + 4,
                            g->pip_waves, s);
    } else {
      launch_filter(ix->tree, g->rects.p, B, g->order.p, false, g->cb.p, g->cp.p, C, g->cnt.p + 1, s);
      launch_distance_refine_dev(rf->build, rf->ring_index, g->pcol, o.distance, g->cb.p, g->cp.p, g->cnt.p + 1, C,
                                 g->ob.p, g->op.p, C, g->cnt.p + 3, g->cnt.p + 4, s);
    }
    SJ_CHECK(hipMemcpyAsync(g->host_cnt, g->cnt.p, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  } catch (...) {
    hipGraph_t dead = nullptr;
    (void)hipStreamEndCapture(s, &dead);
    if (dead) (void)hipGraphDestroy(dead);
    throw;
  }
  SJ_CHECK(hipStreamEndCapture(s, &g->graph));
  check_graph_nodes(g->graph);
  SJ_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  // the plan outlives the capturing stream's use: free on the detached stream
  hipStream_t ds = detached_stream();
  for (auto* a : {&g->keys, &g->ids, &g->keys_sorted, &g->order, &g->cb, &g->cp, &g->redo, &g->ob, &g->op}) a->s = ds;
  g->desc.s = ds;
  g->pcol.xy.s = ds;
  g->pcol.type.s = ds;
  g->rcol.xy.s = ds;
  g->rects.s = ds;
  g->sort_tmp.s = ds;
  g->cnt.s = ds;
  return g;
}

// Runs one request through a plan.  Returns false (nothing produced) when
// the request must take the eager path instead.
bool run_plan(GraphPlan& g, const WkbInput& in, SjhipJoinResult* out, hipStream_t s) {
  g.host_desc->data = in.data;
  // an Arrow array's row offset moves the start of its offsets buffer
  g.host_desc->offsets = (const uint8_t*)in.offsets + in.offset * (in.large ? 8 : 4);
  g.host_desc->n = in.length;
  g.host_desc->large = in.large ? 1 : 0;
  g.host_desc->pad = 0;
  SJ_CHECK(hipMemcpyAsync(g.desc.p, g.host_desc, sizeof(PointsDesc), hipMemcpyHostToDevice, s));
  SJ_CHECK(hipGraphLaunch(g.exec, s));
  SJ_CHECK(hipStreamSynchronize(s));
  const unsigned long long* h = g.host_cnt;
  if (h[0] != ~0ull || h[1] > g.C) return false;
  const uint64_t kept = h[3];
  std::unique_ptr<JoinResultImpl> jr(new JoinResultImpl());
  jr->device_id = g.device_id;
  jr->pairs.b.alloc(kept > 0 ? kept : 1, s);
  jr->pairs.p.alloc(kept > 0 ? kept : 1, s);
  if (kept) {
    SJ_CHECK(hipMemcpyAsync(jr->pairs.b.p, g.ob.p, kept * 4, hipMemcpyDeviceToDevice, s));
    SJ_CHECK(hipMemcpyAsync(jr->pairs.p.p, g.op.p, kept * 4, hipMemcpyDeviceToDevice, s));
  }
  SJ_CHECK(hipStreamSynchronize(s));
  jr->pairs.b.s = jr->pairs.p.s = detached_stream();
  std::memset(out, 0, sizeof(*out));
  out->build_indices = jr->pairs.b.p;
  out->probe_indices = jr->pairs.p.p;
  out->size = kept;
  out->candidates = h[1];
  out->segment_tests = h[4];
  out->fallback_tests = h[5];
  out->refine_candidates = h[1];  // the captured sequence walks every candidate
  out->release = jr_release;
  out->private_data = jr.release();
  return true;
}

// ---------------------------------------------------------------------------
// The join without allocation (SjhipJoinDeviceNoAlloc)
//
// The launch sequence of a captured request plan, over a file-scope device
// array instead of allocated buffers.  Each GPU holds its own copy of the
// array for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns on that device.

constexpr int64_t kNoAllocMaxRows = SJHIP_NOALLOC_MAX_ROWS;
constexpr uint64_t kNoAllocMaxCandidates = SJHIP_NOALLOC_MAX_CANDIDATES;

constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
// bytes of the radix sort's temporary storage for n keys (graph_sort_temp_bytes)
constexpr size_t na_sort_bytes(int64_t n) {
  return 2 * na_align((size_t)n * 4) + (size_t)((n + 4095) / 4096) * 256 * 4 + 256;
}

// Scratch layout, in order, each part 256-byte aligned: the probe
// descriptor, 8 counters, decoded points (double2) and types, rects,
// Morton keys, ids, sorted keys and order, the sort's storage, the points
// in Morton order, then per candidate pair its build id, probe position
// and deferred-list slot.
struct NoAllocLayout {
  static constexpr size_t desc = 0;
  static constexpr size_t cnt = desc + na_align(sizeof(PointsDesc));
  static constexpr size_t xy = cnt + na_align(8 * 8);
  static constexpr size_t type = xy + na_align(16 * kNoAllocMaxRows);
  static constexpr size_t rects = type + na_align(kNoAllocMaxRows);
  static constexpr size_t keys = rects + na_align(16 * kNoAllocMaxRows);
  static constexpr size_t ids = keys + na_align(4 * kNoAllocMaxRows);
  static constexpr size_t keys_sorted = ids + na_align(4 * kNoAllocMaxRows);
  static constexpr size_t order = keys_sorted + na_align(4 * kNoAllocMaxRows);
  static constexpr size_t sort = order + na_align(4 * kNoAllocMaxRows);
  static constexpr size_t rxy = sort + na_align(na_sort_bytes(kNoAllocMaxRows));
  static constexpr size_t cb = rxy + na_align(16 * kNoAllocMaxRows);
  static constexpr size_t cp = cb + na_align(4 * kNoAllocMaxCandidates);
  static constexpr size_t redo = cp + na_align(4 * kNoAllocMaxCandidates);
  static constexpr size_t total = redo + na_align(4 * kNoAllocMaxCandidates);
};
static_assert(NoAllocLayout::total == SJHIP_NOALLOC_SCRATCH_BYTES, "scratch size declared in the header");

}  // namespace sjhip

// The scratch array (file scope, one copy per GPU).  Host code must not
// pass its name to a kernel: it is a host-side stand-in, and the device
// address comes from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) unsigned char scratch[sjhip::NoAllocLayout::total];

namespace sjhip {
namespace {

struct NoAllocDevice {
  std::mutex mu;  // calls on one device run one at a time (they share the array)
  unsigned char* ptr = nullptr;
};

NoAllocDevice& noalloc_device(int dev) {
  static NoAllocDevice devs[64];
  if (dev < 0 || dev >= 64) throw Error("device id out of range for the no-allocation join");
  return devs[dev];
}

// The device address of `scratch` on device `dev`, resolved once per device
// with that device current (the runtime's init does it; a first call on
// another device does it then).
unsigned char* noalloc_scratch(int dev) {
  NoAllocDevice& d = noalloc_device(dev);
  static std::mutex resolve_mu;
  std::lock_guard<std::mutex> g(resolve_mu);
  if (!d.ptr) {
    int cur = -1;
    SJ_CHECK(hipGetDevice(&cur));
    if (cur != dev) throw Error("no-allocation scratch resolved with another device current");
    void* p = nullptr;
    SJ_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error("no-allocation scratch has no device address");
    d.ptr = (unsigned char*)p;
  }
  return d.ptr;
}

// A geometry column over borrowed device memory: its buffers are detached
// before the column is destroyed, so nothing is freed.
struct BorrowedPoints {
  GeomColumn col;
  BorrowedPoints(double2* xy, uint8_t* type, int64_t n) {
    col.n_rows = col.n_parts = col.n_rings = col.n_pts = n;
    col.points_only = true;
    col.type_mask = 1u << kPoint;
    col.xy.p = xy;
    col.xy.n = (size_t)n;
    col.type.p = type;
    col.type.n = type ? (size_t)n : 0;
  }
  ~BorrowedPoints() {
    col.xy.p = nullptr;
    col.xy.n = 0;
    col.type.p = nullptr;
    col.type.n = 0;
  }
};

void join_no_alloc(IndexImpl* ix, RefinerImpl* rf, const WkbInput& in, const SjhipJoinOptions& o, uint32_t* ob,
                   uint32_t* op, uint64_t out_cap, uint64_t* out_size) {
  if (!ix->built || !rf->built) throw Error("index and refiner must be built");
  if (o.sort_output) throw Error("the no-allocation join returns pairs in refine order: sort_output must be 0");
  if (in.length > kNoAllocMaxRows)
    throw Error("the no-allocation join takes at most " + std::to_string(kNoAllocMaxRows) + " probe rows, not " +
                std::to_string(in.length));
  if (in.length == 0) {
    *out_size = 0;
    return;
  }
  const int kind = fixed_sequence_kind(rf, in, o);
  if (kind < 0)
    throw Error("the no-allocation join takes device-resident Binary/LargeBinary point columns without nulls, "
                "certified precision (0), and ST_Intersects / ST_Contains / ST_Covers over a polygon build side or "
                "ST_DWithin");
  if (out_cap > 0 && (!ob || !op)) throw Error("null output arrays");
  NoAllocDevice& dev = noalloc_device(rf->rt->device_id);
  std::lock_guard<std::mutex> call_lock(dev.mu);
  unsigned char* sp = noalloc_scratch(rf->rt->device_id);
  using L = NoAllocLayout;
  PointsDesc* desc = (PointsDesc*)(sp + L::desc);
  unsigned long long* cnt = (unsigned long long*)(sp + L::cnt);
  double2* xy = (double2*)(sp + L::xy);
  uint8_t* type = sp + L::type;
  float4* rects = (float4*)(sp + L::rects);
  uint32_t* keys = (uint32_t*)(sp + L::keys);
  uint32_t* ids = (uint32_t*)(sp + L::ids);
  uint32_t* keys_sorted = (uint32_t*)(sp + L::keys_sorted);
  uint32_t* order = (uint32_t*)(sp + L::order);
  void* sort_tmp = sp + L::sort;
  double2* rxy = (double2*)(sp + L::rxy);
  uint32_t* cb = (uint32_t*)(sp + L::cb);
  uint32_t* cp = (uint32_t*)(sp + L::cp);
  uint32_t* redo = (uint32_t*)(sp + L::redo);
  const int64_t n = in.length;
  const uint64_t C = kNoAllocMaxCandidates;
  if (graph_sort_temp_bytes(n) > na_sort_bytes(kNoAllocMaxRows)) throw Error("no-allocation sort storage too small");
  hipStream_t s = (hipStream_t)o.stream;
  PointsDesc d;
  d.data = in.data;
  d.offsets = (const uint8_t*)in.offsets + in.offset * (in.large ? 8 : 4);
  d.n = n;
  d.large = in.large ? 1 : 0;
  d.pad = 0;
  const bool dwithin = kind == 1;
  // every counter the sequence reads is set here, by kernels: the array
  // keeps the previous call's contents
  store_points_desc(d, desc, s);
  fill_u64(cnt, 1, ~0ull, s);
  fill_u64(cnt + 1, 7, 0ull, s);
  decode_points_desc(desc, n, xy, type, cnt, s);
  rects_of_points_into(xy, type, n, !dwithin, dwithin ? o.distance : NAN, rects, s);
  morton_order_fixed(ix->tree, rects, n, keys, ids, keys_sorted, order, sort_tmp, na_sort_bytes(kNoAllocMaxRows), s);
  if (!dwithin) {
    BorrowedPoints rcol(rxy, nullptr, n);
    gather_points_into(xy, order, n, rxy, s);
    launch_filter(ix->tree, rects, n, order, true, cb, cp, C, cnt + 1, s, true);
    launch_pip_refine_dev(rf->build, rf->ring_index, rcol.col, o.predicate, cb, cp, cnt + 1, C, order, redo, cnt + 2,
                          ob, op, out_cap, cnt + 3, cnt + 4, ix->tuning.pip_waves, s);
  } else {
    BorrowedPoints pcol(xy, type, n);
    launch_filter(ix->tree, rects, n, order, false, cb, cp, C, cnt + 1, s);
    launch_distance_refine_dev(rf->build, rf->ring_index, pcol.col, o.distance, cb, cp, cnt + 1, C, ob, op, out_cap,
                               cnt + 3, cnt + 4, s);
  }
  unsigned long long h[8];
  SJ_CHECK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
  SJ_CHECK(hipStreamSynchronize(s));
  if (h[0] != ~0ull)
    throw Error("no-allocation join: probe row " + std::to_string(h[0] & 0xffffffffull) +
                " is not a 21-byte WKB point record");
  if (h[1] > C)
    throw Error("no-allocation join: " + std::to_string(h[1]) + " candidate pairs exceed the scratch limit of " +
                std::to_string(C) + " (split the batch)");
  *out_size = h[3];
  if (h[3] > out_cap)
    throw Error("no-allocation join: " + std::to_string(h[3]) + " pairs exceed the output capacity of " +
                std::to_string(out_cap) + " (*out_size holds the count)");
}

}  // namespace

struct ServerReq {
  uint64_t id = 0;
  WkbInput in;
  SjhipJoinOptions opt{};
  uint64_t estimate = 0;
  int status = 0;  // 0 queued, 1 running, 2 done, 3 failed
  int rc = 0;
  std::string error;
  SjhipJoinResult res{};
  double t_submit = 0, t_admit = 0, t_done = 0;
};

struct Server {
  IndexImpl* ix = nullptr;
  RefinerImpl* rf = nullptr;
  bool graphs = false;
  uint64_t budget = 0;
  uint64_t reserved = 0;
  int running = 0;
  bool stop = false;
  uint64_t next_id = 1;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<ServerReq*> queue;
  std::map<uint64_t, std::unique_ptr<ServerReq>> reqs;
  std::vector<std::thread> workers;
  // Plans that are no longer replayed.  They are destroyed only once every
  // worker has joined: their pinned host buffers are freed with
  // hipHostFree, which synchronises the device and must not run while
  // another worker is capturing a graph on its stream.
  std::vector<std::unique_ptr<GraphPlan>> retired;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // Device bytes a request holds at its peak: the decoded probe column,
  // rects, Morton keys and order, the ordered copy of points, and per
  // candidate the pair arrays, their regrouping copies and keys, the
  // deferred list, the raster pass's list and marks, and the kept pairs.  Candidates per probe come from the
  // index's last probe (8 before any).
  uint64_t estimate(const WkbInput& in, const SjhipJoinOptions& o) const {
    uint64_t n = (uint64_t)in.length;
    uint64_t per16 = ix->tree.cand_per_query_x16.load(std::memory_order_relaxed);
    double cpq = per16 ? per16 / 16.0 : 8.0;
    if (o.predicate == 100) cpq *= 2.0;  // expanded probe rects
    uint64_t cand = (uint64_t)(cpq * (double)n * 1.125) + 1024;
    uint64_t wkb = in.on_device ? 0 : (uint64_t)(in.large ? 8 : 4) * (n + 1) + 32 * n;
    return wkb + 96 * n + 49 * cand + (1u << 20);
  }

  // The kind of captured sequence a request can replay (0 point in
  // polygon, 1 distance), or -1: eager path.
  int graph_kind(const WkbInput& in, const SjhipJoinOptions& o) const {
    if (!graphs || in.length > GraphPlan::kMaxRows) return -1;
    return fixed_sequence_kind(rf, in, o);
  }

  void work() {
    use_device(rf->rt);
    hipStream_t s;
    SJ_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    // this worker's captured sequences, by (kind, predicate, bucket rows)
    std::map<std::tuple<int, int, int64_t>, std::unique_ptr<GraphPlan>> plans;
    std::vector<std::unique_ptr<GraphPlan>> old;  // replaced plans (see `retired`)
    auto run_graph = [&](const WkbInput& in, const SjhipJoinOptions& o, SjhipJoinResult* res) {
      const int kind = graph_kind(in, o);
      if (kind < 0) return false;
      int64_t B = GraphPlan::kMinRows;
      while (B < in.length) B <<= 1;
      auto key = std::make_tuple(kind, (int)o.predicate, B);
      auto& plan = plans[key];
      if (plan && (plan->distance != o.distance || plan->leaf != ix->tree.leaf ||
                   plan->tree_gen != ix->tree_gen || plan->ring_gen != rf->ring_gen ||
                   plan->pip_waves != ix->tuning.pip_waves))
        old.push_back(std::move(plan));
      if (!plan) {
        uint64_t per16 = ix->tree.cand_per_query_x16.load(std::memory_order_relaxed);
        double cpq = per16 ? per16 / 16.0 : 8.0;
        if (kind == 1) cpq *= 2.0;
        plan = capture_plan(ix, rf, kind, o, B, (uint64_t)(cpq * 1.5 * (double)B) + 65536, s);
      }
      if (run_plan(*plan, in, res, s)) {
        if (o.sort_output && res->size) {
          sort_pairs(res->build_indices, res->probe_indices, res->size, s);
          SJ_CHECK(hipStreamSynchronize(s));
        }
        return true;
      }
      // too many candidates for the plan: capture a larger one next time
      if (plan->host_cnt[0] == ~0ull && plan->host_cnt[1] > plan->C) {
        uint64_t C = plan->host_cnt[1] + plan->host_cnt[1] / 4 + 65536;
        old.push_back(std::move(plan));
        plan = capture_plan(ix, rf, kind, o, B, C, s);
        if (run_plan(*plan, in, res, s)) {
          if (o.sort_output && res->size) {
            sort_pairs(res->build_indices, res->probe_indices, res->size, s);
            SJ_CHECK(hipStreamSynchronize(s));
          }
          return true;
        }
      }
      return false;
    };
    for (;;) {
      ServerReq* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] {
          if (queue.empty()) return stop;
          ServerReq* h = queue.front();
          return running == 0 || reserved + h->estimate <= budget;
        });
        if (queue.empty()) break;
        r = queue.front();
        queue.pop_front();
        reserved += r->estimate;
        running++;
        r->status = 1;
        r->t_admit = now_ms();
      }
      SjhipJoinOptions o = r->opt;
      o.stream = s;
      SjhipJoinResult res{};
      std::string err;
      int rc = guarded(&err, [&] {
        if (!run_graph(r->in, o, &res)) JoinEngine::run(ix, rf, r->in, o, &res);
      });
      {
        std::lock_guard<std::mutex> lk(mu);
        reserved -= r->estimate;
        running--;
        r->rc = rc;
        r->error = err;
        r->res = res;
        r->status = rc ? 3 : 2;
        r->t_done = now_ms();
      }
      cv.notify_all();
    }
    (void)hipStreamSynchronize(s);
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& kv : plans)
        if (kv.second) retired.push_back(std::move(kv.second));
      for (auto& g : old) retired.push_back(std::move(g));
    }
    (void)hipStreamDestroy(s);
  }
};

namespace {

__global__ void k_add_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] += v;
}

}  // namespace

// The join of a host-resident probe column in batches of batch_rows rows,
// `slots` batches in flight: a worker thread per slot takes every
// slots-th batch, copies its rows in through pinned blocks (the CPU fills
// one chunk while the DMA of the previous one runs), joins it on its own
// stream and copies the pairs back into a pinned block, so that one
// batch's copies overlap another batch's join.  The callback gets each
// batch's pairs, probe indices global to the column, in batch order.
// Follows the reference's two-slot refine pipeline
// (c/sedona-libgpuspatial/libgpuspatial/src/rt_spatial_refiner.cu:244-436)
// and its pinned staging (parallel_wkb_loader.hpp:505-534).
int join_host_stream(IndexImpl* ix, RefinerImpl* rf, const WkbInput& in, const SjhipJoinOptions& opts,
                     int64_t batch_rows, int slots, int (*cb)(const uint32_t*, const uint32_t*, uint32_t, void*),
                     void* ud) {
  if (in.on_device) throw Error("the streaming join reads a host column (use SjhipJoinDevice for device arrays)");
  if (in.view) throw Error("the streaming join takes Binary or LargeBinary columns (BinaryView: SjhipJoinDevice)");
  if (batch_rows <= 0 || batch_rows > (int64_t)1 << 31) throw Error("batch_rows must be in 1 .. 2^31");
  if (slots < 1 || slots > 8) throw Error("slots must be in 1 .. 8");
  if (!cb) throw Error("null callback");
  if (opts.stream) throw Error("the streaming join runs on streams of its own (options.stream must be NULL)");
  const int64_t n = in.length;
  const int64_t nb = (n + batch_rows - 1) / batch_rows;
  const size_t ow = in.large ? 8 : 4;
  auto off_at = [&](int64_t row) -> int64_t {  // (row counts the array offset in)
    return in.large ? ((const int64_t*)in.offsets)[row] : (int64_t)((const int32_t*)in.offsets)[row];
  };
  std::mutex mu;
  std::condition_variable turn;
  int64_t next = 0;  // the batch whose pairs go to the callback next
  bool stop = false;
  int cb_rc = 0;
  std::exception_ptr err;
  auto worker = [&](int w) {
    try {
      use_device(rf->rt);
      ScopedStream ss;
      DevArray<uint8_t> d_off, d_data, d_valid;
      for (int64_t k = w; k < nb; k += slots) {
        {
          std::lock_guard<std::mutex> g(mu);
          if (stop) break;
        }
        const int64_t r0 = k * batch_rows, cnt = std::min(batch_rows, n - r0);
        // rows [first, first + lead + cnt) of the array, first a multiple of
        // 8 so that the validity bytes start at a byte boundary
        const int64_t abs0 = in.offset + r0;
        const int64_t lead = abs0 & 7, first = abs0 - lead;
        const int64_t o0 = off_at(first), o1 = off_at(abs0 + cnt);
        d_off.reserve((size_t)(lead + cnt + 1) * ow, ss.s);
        d_data.reserve((size_t)(o1 - o0) + 32, ss.s);
        upload_staged(d_off.p, (const uint8_t*)in.offsets + first * ow, (size_t)(lead + cnt + 1) * ow, ss.s);
        upload_staged(d_data.p, in.data + o0, (size_t)(o1 - o0), ss.s);
        SJ_CHECK(hipMemsetAsync(d_data.p + (o1 - o0), 0, 32, ss.s));  // (the host decode's zero padding)
        WkbInput bin;
        bin.length = cnt;
        bin.offset = lead;
        bin.offsets = d_off.p;
        bin.large = in.large;
        // the offsets keep their values: the data pointer is moved back by o0
        bin.data = d_data.p - o0;
        bin.on_device = true;
        if (in.validity) {
          const size_t vb = (size_t)((lead + cnt + 7) / 8);
          d_valid.reserve(vb, ss.s);
          upload_staged(d_valid.p, in.validity + first / 8, vb, ss.s);
          bin.validity = d_valid.p;
        }
        SjhipJoinOptions o = opts;
        o.stream = ss.s;
        SjhipJoinResult res;
        std::memset(&res, 0, sizeof(res));
        JoinEngine::run(ix, rf, bin, o, &res);
        std::unique_ptr<SjhipJoinResult, void (*)(SjhipJoinResult*)> hold(&res, [](SjhipJoinResult* r) {
          if (r->release) r->release(r);
        });
        const uint64_t m = res.size;
        if (m > 0xffffffffull) throw Error("a batch gave more than 2^32 - 1 pairs: use a smaller batch_rows");
        PinnedLease l(m * 8 + 8);
        uint32_t* hb = (uint32_t*)l.p;
        uint32_t* hp = hb + m;
        if (m) {
          if (r0) hipLaunchKernelGGL(k_add_u32, (unsigned)std::min<uint64_t>((m + 255) / 256, 4096), 256, 0, ss.s,
                                     res.probe_indices, m, (uint32_t)r0);
          SJ_CHECK(hipGetLastError());
          SJ_CHECK(hipMemcpyAsync(hb, res.build_indices, m * 4, hipMemcpyDeviceToHost, ss.s));
          SJ_CHECK(hipMemcpyAsync(hp, res.probe_indices, m * 4, hipMemcpyDeviceToHost, ss.s));
        }
        SJ_CHECK(hipStreamSynchronize(ss.s));
        std::unique_lock<std::mutex> lk(mu);
        turn.wait(lk, [&] { return next == k || stop; });
        if (stop) break;
        const int rc = cb(hb, hp, (uint32_t)m, ud);
        if (rc) {
          cb_rc = rc;
          stop = true;
        }
        next++;
        turn.notify_all();
      }
    } catch (...) {
      std::lock_guard<std::mutex> g(mu);
      if (!err) err = std::current_exception();
      stop = true;
      turn.notify_all();
    }
  };
  const int nw = (int)std::min<int64_t>(slots, std::max<int64_t>(nb, 1));
  std::vector<std::thread> th;
  for (int w = 0; w < nw; w++) th.emplace_back(worker, w);
  for (auto& t : th) t.join();
  if (err) std::rethrow_exception(err);
  if (nb == 0 && !cb_rc) cb_rc = cb(nullptr, nullptr, 0, ud);
  return cb_rc;
}

namespace {

__global__ void k_map_u32(uint32_t* p, uint64_t n, const uint32_t* table) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = table[p[i]];
}

}  // namespace

// Estimated device bytes of a build row once resident (decoded FP64
// vertices, FP32 ring-local copies, segment trees, cell raster, row
// records and offsets) from its WKB size: a polygon's WKB holds 16 bytes
// per vertex, its resident form about 30 (16 FP64 + 8 FP32 + tree boxes
// and raster), plus a fixed share per row.
uint64_t resident_bytes_of_row(int64_t wkb_bytes) { return 2 * (uint64_t)wkb_bytes + 256; }

// Splits the rows ids[b, e) into partitions of at most `budget` estimated
// resident bytes, K-D-B style (rust/sedona-spatial-join/src/partitioning/
// kdb.rs:36-40, 302-360): along the longer axis of the rows' rect minima,
// ordered by (min on that axis, min on the other, row id), at the row
// where the bytes before it reach half (at least one row on each side).
// Every row lands in exactly one partition.
void kdb_split(std::vector<uint32_t>& ids, size_t b, size_t e, const std::vector<float4>& rects,
               const std::vector<uint64_t>& bytes, uint64_t budget, std::vector<std::vector<uint32_t>>& out) {
  uint64_t total = 0;
  for (size_t i = b; i < e; i++) total += bytes[ids[i]];
  if (total <= budget || e - b <= 1) {
    out.emplace_back(ids.begin() + b, ids.begin() + e);
    return;
  }
  float x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (size_t i = b; i < e; i++) {
    const float4 r = rects[ids[i]];
    x0 = std::min(x0, r.x); x1 = std::max(x1, r.x);
    y0 = std::min(y0, r.y); y1 = std::max(y1, r.y);
  }
  const bool by_x = (x1 - x0) >= (y1 - y0);
  std::sort(ids.begin() + b, ids.begin() + e, [&](uint32_t u, uint32_t v) {
    const float4 a = rects[u], c = rects[v];
    const float a0 = by_x ? a.x : a.y, c0 = by_x ? c.x : c.y;
    if (a0 != c0) return a0 < c0;
    const float a1 = by_x ? a.y : a.x, c1 = by_x ? c.y : c.x;
    if (a1 != c1) return a1 < c1;
    return u < v;
  });
  size_t k = b;
  uint64_t acc = 0;
  while (k < e - 1 && (acc + bytes[ids[k]]) * 2 <= total) acc += bytes[ids[k++]];
  if (k == b) k = b + 1;
  kdb_split(ids, b, k, rects, bytes, budget, out);
  kdb_split(ids, k, e, rects, bytes, budget, out);
}

}  // namespace sjhip

using namespace sjhip;

extern "C" {

void GpuSpatialRuntimeCreate(struct GpuSpatialRuntime* runtime) {
  runtime->init = rt_init;
  runtime->release = rt_release;
  runtime->get_last_error = rt_err;
  runtime->private_data = new Runtime();
}

int GpuSpatialIndexFloat2DCreate(struct SedonaFloatIndex2D* index, const struct GpuSpatialIndexConfig* config) {
  std::string err;
  return guarded(&err, [&] {
    if (!index || !config) throw Error("null index or config");
    Runtime* rt = runtime_of(config->runtime);
    use_device(rt);
    IndexImpl* ix = new IndexImpl();
    ix->rt = rt;
    SJ_CHECK(hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking));
    index->clear = ix_clear;
    index->create_context = ix_create_context;
    index->destroy_context = ix_destroy_context;
    index->push_build = ix_push_build;
    index->finish_building = ix_finish;
    index->probe = ix_probe;
    index->get_last_error = ix_err;
    index->context_get_last_error = ix_ctx_err;
    index->release = ix_release;
    index->private_data = ix;
  });
}

int GpuSpatialRefinerCreate(struct SedonaSpatialRefiner* refiner, const struct GpuSpatialRefinerConfig* config) {
  std::string err;
  return guarded(&err, [&] {
    if (!refiner || !config) throw Error("null refiner or config");
    Runtime* rt = runtime_of(config->runtime);
    use_device(rt);
    RefinerImpl* rf = new RefinerImpl();
    rf->rt = rt;
    SJ_CHECK(hipStreamCreateWithFlags(&rf->stream, hipStreamNonBlocking));
    refiner->clear = rf_clear;
    refiner->init_build_schema = rf_init_schema;
    refiner->push_build = rf_push_build;
    refiner->finish_building = rf_finish;
    refiner->refine = rf_refine;
    refiner->get_last_error = rf_err;
    refiner->release = rf_release;
    refiner->private_data = rf;
  });
}

int SjhipExtVersion(void) { return SJHIP_EXT_VERSION; }

int SjhipTestingGraphSortPairs(const uint32_t* keys, const uint32_t* values, uint32_t* keys_out,
                               uint32_t* values_out, int64_t n, int begin_bit, int end_bit, int device_id) {
  return guarded(nullptr, [&] {
    if (n < 0 || (n > 0 && (!keys || !values || !keys_out || !values_out))) throw Error("null buffer");
    SJ_CHECK(hipSetDevice(device_id));
    ScopedStream ss;
    {
      DevArray<uint8_t> temp(graph_sort_temp_bytes(n), ss.s);
      graph_sort_pairs(keys, values, keys_out, values_out, n, begin_bit, end_bit, temp.p, temp.n, ss.s);
    }
    SJ_CHECK(hipStreamSynchronize(ss.s));
  });
}

const char* SjhipLastError(void) { return t_last_error.c_str(); }

int SjhipRefinerRefineDistance(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                               const struct ArrowArray* probe_array, double distance, uint32_t* build_indices,
                               uint32_t* probe_indices, uint32_t indices_size, uint32_t* new_indices_size) {
  RefinerImpl* rf = self ? (RefinerImpl*)self->private_data : nullptr;
  std::string err;
  int rc = guarded(&err, [&] {
    if (!rf) throw Error("refiner is null or released");
    if (!(distance == distance)) throw Error("distance is NaN");
    refine_host(rf, probe_schema, probe_array, 100, distance, build_indices, probe_indices, indices_size,
                new_indices_size);
  });
  if (rc && rf) {
    std::lock_guard<std::mutex> g(rf->mu);
    rf->last_error = err;
  }
  return rc;
}

int SjhipRefinerKnn(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                    const struct ArrowArray* probe_array, uint32_t k, uint32_t* out_build_indices,
                    double* out_distances) {
  RefinerImpl* rf = self ? (RefinerImpl*)self->private_data : nullptr;
  std::string err;
  int rc = guarded(&err, [&] {
    if (!rf) throw Error("refiner is null or released");
    if (!rf->built) throw Error("knn before finish_building");
    if (!out_build_indices || !out_distances) throw Error("null knn output arrays");
    use_device(rf->rt);
    ScopedStream ss;
    GeomColumn probe;
    decode_wkb(wkb_input_from_arrow(probe_schema, probe_array, false), probe, ss.s);
    size_t m = (size_t)probe.n_rows * k;
    if (m == 0) return;
    DevArray<uint32_t> ids(m, ss.s);
    DevArray<double> d(m, ss.s);
    {
      std::lock_guard<std::mutex> g(rf->knn_mu);
      knn_points(rf->build, rf->knn, probe, (int)k, ids.p, d.p, ss.s);
    }
    SJ_CHECK(hipMemcpyAsync(out_build_indices, ids.p, m * 4, hipMemcpyDeviceToHost, ss.s));
    SJ_CHECK(hipMemcpyAsync(out_distances, d.p, m * 8, hipMemcpyDeviceToHost, ss.s));
    SJ_CHECK(hipStreamSynchronize(ss.s));
  });
  if (rc && rf) {
    std::lock_guard<std::mutex> g(rf->mu);
    rf->last_error = err;
  }
  return rc;
}

int SjhipRefinerPairDistances(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                              const struct ArrowArray* probe_array, const uint32_t* build_indices,
                              const uint32_t* probe_indices, uint64_t n, double* out_distances) {
  RefinerImpl* rf = self ? (RefinerImpl*)self->private_data : nullptr;
  std::string err;
  int rc = guarded(&err, [&] {
    if (!rf) throw Error("refiner is null or released");
    if (!rf->built) throw Error("distances before finish_building");
    if (n > 0 && (!build_indices || !probe_indices || !out_distances)) throw Error("null distance arrays");
    use_device(rf->rt);
    ScopedStream ss;
    GeomColumn probe;
    decode_wkb(wkb_input_from_arrow(probe_schema, probe_array, false), probe, ss.s);
    for (uint64_t i = 0; i < n; i++) {
      if (build_indices[i] >= (uint64_t)rf->build.n_rows)
        throw Error("build index out of range: " + std::to_string(build_indices[i]));
      if (probe_indices[i] >= (uint64_t)probe.n_rows)
        throw Error("probe index out of range: " + std::to_string(probe_indices[i]));
    }
    if (n == 0) return;
    DevArray<uint32_t> cb(n, ss.s), cp(n, ss.s);
    DevArray<double> d(n, ss.s);
    SJ_CHECK(hipMemcpyAsync(cb.p, build_indices, n * 4, hipMemcpyHostToDevice, ss.s));
    SJ_CHECK(hipMemcpyAsync(cp.p, probe_indices, n * 4, hipMemcpyHostToDevice, ss.s));
    pair_distances(rf->build, probe, cb.p, cp.p, n, d.p, ss.s);
    SJ_CHECK(hipMemcpyAsync(out_distances, d.p, n * 8, hipMemcpyDeviceToHost, ss.s));
    SJ_CHECK(hipStreamSynchronize(ss.s));
  });
  if (rc && rf) {
    std::lock_guard<std::mutex> g(rf->mu);
    rf->last_error = err;
  }
  return rc;
}

int SjhipRefinerKnnTies(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                        const struct ArrowArray* probe_array, uint32_t k, int32_t include_tie_breakers,
                        uint64_t* out_offsets, uint32_t* out_build_indices, double* out_distances, uint64_t capacity,
                        uint64_t* out_total) {
  RefinerImpl* rf = self ? (RefinerImpl*)self->private_data : nullptr;
  std::string err;
  int rc = guarded(&err, [&] {
    if (!rf) throw Error("refiner is null or released");
    if (!rf->built) throw Error("knn before finish_building");
    if (!out_offsets || !out_total) throw Error("null knn output arrays");
    *out_total = 0;
    use_device(rf->rt);
    ScopedStream ss;
    GeomColumn probe;
    decode_wkb(wkb_input_from_arrow(probe_schema, probe_array, false), probe, ss.s);
    DevArray<unsigned long long> off;
    DevArray<uint32_t> ids;
    DevArray<double> d;
    uint64_t total;
    {
      std::lock_guard<std::mutex> g(rf->knn_mu);
      // DISTANCE_TOLERANCE of the reference (index/spatial_index.rs:32)
      total = knn_with_ties(rf->build, rf->knn, probe, (int)k, include_tie_breakers ? 1e-9 : -1.0, off, ids, d, ss.s);
    }
    *out_total = total;
    if (total > capacity)
      throw Error("knn: " + std::to_string(total) + " neighbours exceed the output capacity of " +
                  std::to_string(capacity) + " (*out_total holds the count)");
    if (total > 0 && (!out_build_indices || !out_distances)) throw Error("null knn output arrays");
    SJ_CHECK(hipMemcpyAsync(out_offsets, off.p, (probe.n_rows + 1) * 8, hipMemcpyDeviceToHost, ss.s));
    if (total) {
      SJ_CHECK(hipMemcpyAsync(out_build_indices, ids.p, total * 4, hipMemcpyDeviceToHost, ss.s));
      SJ_CHECK(hipMemcpyAsync(out_distances, d.p, total * 8, hipMemcpyDeviceToHost, ss.s));
    }
    SJ_CHECK(hipStreamSynchronize(ss.s));
  });
  if (rc && rf) {
    std::lock_guard<std::mutex> g(rf->mu);
    rf->last_error = err;
  }
  return rc;
}

int SjhipRefinerKnnDevice(struct SedonaSpatialRefiner* self, const struct ArrowSchema* probe_schema,
                          const struct ArrowDeviceArray* probe_array, uint32_t k, uint32_t* out_build_indices,
                          double* out_distances, double* kernel_ms) {
  RefinerImpl* rf = self ? (RefinerImpl*)self->private_data : nullptr;
  std::string err;
  int rc = guarded(&err, [&] {
    if (!rf) throw Error("refiner is null or released");
    if (!rf->built) throw Error("knn before finish_building");
    if (!probe_array || !out_build_indices || !out_distances) throw Error("null knn argument");
    bool dev = probe_array->device_type == ARROW_DEVICE_ROCM;
    if (!dev && probe_array->device_type != ARROW_DEVICE_CPU)
      throw Error("probe array must be in ROCm device memory or CPU memory");
    use_device(rf->rt);
    wait_sync_event(probe_array);
    ScopedStream ss;
    GeomColumn probe;
    decode_wkb(wkb_input_from_arrow(probe_schema, &probe_array->array, dev), probe, ss.s);
    KernelTimer kt;
    if (kernel_ms) {
      SJ_CHECK(hipEventCreate(&kt.start));
      SJ_CHECK(hipEventCreate(&kt.stop));
    }
    {
      std::lock_guard<std::mutex> g(rf->knn_mu);
      knn_points(rf->build, rf->knn, probe, (int)k, out_build_indices, out_distances, ss.s, kernel_ms ? &kt : nullptr);
    }
    if (kernel_ms) {
      float ms = 0;
      if (probe.n_rows) SJ_CHECK(hipEventElapsedTime(&ms, kt.start, kt.stop));
      *kernel_ms = ms;
      (void)hipEventDestroy(kt.start);
      (void)hipEventDestroy(kt.stop);
    }
  });
  if (rc && rf) {
    std::lock_guard<std::mutex> g(rf->mu);
    rf->last_error = err;
  }
  return rc;
}

int SjhipJoinDevice(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                    const struct ArrowSchema* probe_schema, const struct ArrowDeviceArray* probe_array,
                    const struct SjhipJoinOptions* options, struct SjhipJoinResult* out) {
  return guarded(nullptr, [&] {
    if (!index || !index->private_data || !refiner || !refiner->private_data) throw Error("null index or refiner");
    if (!probe_array || !options || !out) throw Error("null argument");
    std::memset(out, 0, sizeof(*out));
    bool dev = probe_array->device_type == ARROW_DEVICE_ROCM;
    if (!dev && probe_array->device_type != ARROW_DEVICE_CPU)
      throw Error("probe array must be in ROCm device memory or CPU memory");
    WkbInput in = wkb_input_from_arrow(probe_schema, &probe_array->array, dev);
    wait_sync_event(probe_array);
    JoinEngine::run((IndexImpl*)index->private_data, (RefinerImpl*)refiner->private_data, in, *options, out);
  });
}

int SjhipPartitionedJoin(struct GpuSpatialRuntime* runtime, const struct ArrowSchema* build_schema,
                         const struct ArrowArray* build_array, const struct ArrowSchema* probe_schema,
                         const struct ArrowDeviceArray* probe_array, const struct SjhipJoinOptions* options,
                         uint64_t device_budget_bytes,
                         int (*callback)(const uint32_t*, const uint32_t*, uint32_t, void*), void* user_data,
                         struct SjhipPartitionStats* stats) {
  int cb_rc = 0;
  int rc = guarded(nullptr, [&] {
    if (!build_array || !probe_array || !options || !callback) throw Error("null argument");
    if (device_budget_bytes < ((uint64_t)1 << 20)) throw Error("device_budget_bytes must be at least 1 MiB");
    if (options->stream) throw Error("options.stream must be NULL (each partition's join runs on its own stream)");
    Runtime* rt = runtime_of(runtime);
    use_device(rt);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const auto t_start = std::chrono::steady_clock::now();
    WkbInput in = wkb_input_from_arrow(build_schema, build_array, false);
    if (in.view) throw Error("the partitioned join takes a Binary or LargeBinary build column");
    const int64_t n = in.length;
    if (n >= 0xffffffffll) throw Error("at most 2^32 - 2 build rows");
    auto off_at = [&](int64_t row) -> int64_t {
      return in.large ? ((const int64_t*)in.offsets)[row] : (int64_t)((const int32_t*)in.offsets)[row];
    };
    // the rows' rects (the index's semantics), decoded on the device in
    // chunks of at most a quarter of the budget of WKB
    std::vector<float4> rects((size_t)std::max<int64_t>(n, 1));
    std::vector<uint64_t> bytes((size_t)std::max<int64_t>(n, 1));
    {
      ScopedStream ss;
      const uint64_t chunk_bytes = std::max<uint64_t>(device_budget_bytes / 4, 1);
      int64_t r0 = 0;
      while (r0 < n) {
        int64_t r1 = r0 + 1;
        const int64_t base = off_at(in.offset + r0);
        while (r1 < n && (uint64_t)(off_at(in.offset + r1 + 1) - base) <= chunk_bytes) r1++;
        WkbInput sub = in;
        sub.offset = in.offset + r0;
        sub.length = r1 - r0;
        GeomColumn g;
        decode_wkb(sub, g, ss.s);
        DevArray<float4> r;
        compute_rects(g, nullptr, true, NAN, r, ss.s);
        SJ_CHECK(hipMemcpyAsync(rects.data() + r0, r.p, (size_t)(r1 - r0) * sizeof(float4), hipMemcpyDeviceToHost,
                                ss.s));
        SJ_CHECK(hipStreamSynchronize(ss.s));
        r0 = r1;
      }
    }
    // rows with an empty rect (null or empty geometry) meet nothing
    std::vector<uint32_t> ids;
    for (int64_t i = 0; i < n; i++) {
      bytes[i] = resident_bytes_of_row(off_at(in.offset + i + 1) - off_at(in.offset + i));
      const float4 r = rects[i];
      if (r.x <= r.z && r.y <= r.w) ids.push_back((uint32_t)i);
    }
    std::vector<std::vector<uint32_t>> parts;
    if (!ids.empty()) kdb_split(ids, 0, ids.size(), rects, bytes, device_budget_bytes, parts);
    if (stats) stats->partitions = (uint32_t)parts.size();
    // one partition at a time: its rows gathered into a LargeBinary column,
    // built (refiner + index), joined with the whole probe column, released
    const char* fmt = "Z";
    for (const auto& rows : parts) {
      const auto t0 = std::chrono::steady_clock::now();
      std::vector<int64_t> offs(rows.size() + 1, 0);
      uint64_t est = 0;
      for (size_t k = 0; k < rows.size(); k++) {
        const int64_t a = off_at(in.offset + rows[k]), e = off_at(in.offset + rows[k] + 1);
        offs[k + 1] = offs[k] + (e - a);
        est += bytes[rows[k]];
      }
      std::vector<uint8_t> data((size_t)std::max<int64_t>(offs.back(), 1));
      for (size_t k = 0; k < rows.size(); k++) {
        const int64_t a = off_at(in.offset + rows[k]);
        std::memcpy(data.data() + offs[k], in.data + a, (size_t)(offs[k + 1] - offs[k]));
      }
      const void* bufs[3] = {nullptr, offs.data(), data.data()};
      ArrowArray arr;
      std::memset(&arr, 0, sizeof(arr));
      arr.length = (int64_t)rows.size();
      arr.n_buffers = 3;
      arr.buffers = bufs;
      arr.release = [](ArrowArray* a) { a->release = nullptr; };
      ArrowSchema sch;
      std::memset(&sch, 0, sizeof(sch));
      sch.format = fmt;
      sch.release = [](ArrowSchema* x) { x->release = nullptr; };
      SedonaSpatialRefiner rf;
      SedonaFloatIndex2D ix;
      std::memset(&rf, 0, sizeof(rf));
      std::memset(&ix, 0, sizeof(ix));
      struct Holder {
        SedonaSpatialRefiner* rf;
        SedonaFloatIndex2D* ix;
        ~Holder() {
          if (ix->release) ix->release(ix);
          if (rf->release) rf->release(rf);
        }
      } hold{&rf, &ix};
      GpuSpatialRefinerConfig rcfg{runtime, 1, false, 1};
      GpuSpatialIndexConfig icfg{runtime, 1};
      if (GpuSpatialRefinerCreate(&rf, &rcfg)) throw Error(t_last_error);
      auto rf_check = [&](int r) {
        if (r) throw Error(std::string("partition build: ") + rf.get_last_error(&rf));
      };
      rf_check(rf.init_build_schema(&rf, &sch));
      rf_check(rf.push_build(&rf, &arr));
      rf_check(rf.finish_building(&rf));
      data.clear();
      data.shrink_to_fit();
      if (GpuSpatialIndexFloat2DCreate(&ix, &icfg)) throw Error(t_last_error);
      if (SjhipIndexBuildFromRefiner(&ix, &rf)) throw Error(t_last_error);
      const auto t1 = std::chrono::steady_clock::now();
      SjhipJoinResult res;
      std::memset(&res, 0, sizeof(res));
      if (SjhipJoinDevice(&ix, &rf, probe_schema, probe_array, options, &res)) throw Error(t_last_error);
      std::unique_ptr<SjhipJoinResult, void (*)(SjhipJoinResult*)> hres(&res, [](SjhipJoinResult* r) {
        if (r->release) r->release(r);
      });
      const uint64_t m = res.size;
      if (m > 0xffffffffull) throw Error("a partition gave more than 2^32 - 1 pairs");
      PinnedLease l(m * 8 + 8);
      uint32_t* hb = (uint32_t*)l.p;
      uint32_t* hp = hb + m;
      if (m) {
        // partition rows -> build rows of the whole column
        ScopedStream ss;
        DevArray<uint32_t> table(rows.size(), ss.s);
        SJ_CHECK(hipMemcpyAsync(table.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, ss.s));
        hipLaunchKernelGGL(k_map_u32, (unsigned)std::min<uint64_t>((m + 255) / 256, 4096), 256, 0, ss.s,
                           res.build_indices, m, table.p);
        SJ_CHECK(hipGetLastError());
        SJ_CHECK(hipMemcpyAsync(hb, res.build_indices, m * 4, hipMemcpyDeviceToHost, ss.s));
        SJ_CHECK(hipMemcpyAsync(hp, res.probe_indices, m * 4, hipMemcpyDeviceToHost, ss.s));
        SJ_CHECK(hipStreamSynchronize(ss.s));
      }
      const auto t2 = std::chrono::steady_clock::now();
      if (stats) {
        stats->pairs += m;
        stats->max_partition_rows = std::max<uint64_t>(stats->max_partition_rows, rows.size());
        stats->max_partition_bytes = std::max<uint64_t>(stats->max_partition_bytes, est);
        stats->build_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->join_ms += std::chrono::duration<double, std::milli>(t2 - t1).count();
      }
      cb_rc = callback(hb, hp, (uint32_t)m, user_data);
      if (cb_rc) return;
    }
    if (parts.empty()) cb_rc = callback(nullptr, nullptr, 0, user_data);
    if (stats) stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  });
  return rc ? rc : cb_rc;
}

int SjhipJoinHostStream(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                        const struct ArrowSchema* probe_schema, const struct ArrowArray* probe_array,
                        const struct SjhipJoinOptions* options, int64_t batch_rows, int32_t slots,
                        int (*callback)(const uint32_t*, const uint32_t*, uint32_t, void*), void* user_data) {
  int cb_rc = 0;
  int rc = guarded(nullptr, [&] {
    if (!index || !index->private_data || !refiner || !refiner->private_data) throw Error("null index or refiner");
    if (!probe_array || !options) throw Error("null argument");
    WkbInput in = wkb_input_from_arrow(probe_schema, probe_array, false);
    cb_rc = join_host_stream((IndexImpl*)index->private_data, (RefinerImpl*)refiner->private_data, in, *options,
                             batch_rows, slots, callback, user_data);
  });
  return rc ? rc : cb_rc;
}

int SjhipJoinDeviceNoAlloc(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                           const struct ArrowSchema* probe_schema, const struct ArrowDeviceArray* probe_array,
                           const struct SjhipJoinOptions* options, uint32_t* out_build_indices,
                           uint32_t* out_probe_indices, uint64_t out_capacity, uint64_t* out_size) {
  return guarded(nullptr, [&] {
    if (!index || !index->private_data || !refiner || !refiner->private_data) throw Error("null index or refiner");
    if (!probe_array || !options || !out_size) throw Error("null argument");
    *out_size = 0;
    if (probe_array->device_type != ARROW_DEVICE_ROCM)
      throw Error("the no-allocation join needs the probe column in ROCm device memory");
    RefinerImpl* rf = (RefinerImpl*)refiner->private_data;
    use_device(rf->rt);
    WkbInput in = wkb_input_from_arrow(probe_schema, &probe_array->array, true);
    wait_sync_event(probe_array);
    join_no_alloc((IndexImpl*)index->private_data, rf, in, *options, out_build_indices, out_probe_indices,
                  out_capacity, out_size);
  });
}

int SjhipRefinerSetSegmentLeaf(struct SedonaSpatialRefiner* refiner, int32_t segments_per_leaf) {
  return guarded(nullptr, [&] {
    if (!refiner || !refiner->private_data) throw Error("null refiner");
    RefinerImpl* rf = (RefinerImpl*)refiner->private_data;
    if (!rf->built) throw Error("refiner must be built");
    if (segments_per_leaf != 16 && segments_per_leaf != 32) throw Error("segments_per_leaf must be 16 or 32");
    use_device(rf->rt);
    std::lock_guard<std::mutex> g(rf->mu);
    RingIndexParams prm;
    prm.seg_leaf_shift = segments_per_leaf == 32 ? 5 : 4;
    if (rf->ring_index.ring_rect.p == nullptr || rf->ring_index.seg_leaf_shift == prm.seg_leaf_shift) return;
    RingIndex ri;
    build_ring_index(rf->build, ri, prm, rf->stream);
    SJ_CHECK(hipStreamSynchronize(rf->stream));
    rf->ring_index = std::move(ri);
    rf->ring_gen++;
  });
}

int SjhipIndexSetFilterMode(struct SedonaFloatIndex2D* index, int32_t mode) {
  return guarded(nullptr, [&] {
    if (!index || !index->private_data) throw Error("null index");
    if (mode < 0 || mode > 3) throw Error("filter mode must be 0, 1, 2 or 3");
    IndexImpl* ix = (IndexImpl*)index->private_data;
    std::lock_guard<std::mutex> g(ix->mu);
    ix->tree.filter_mode = mode;
  });
}

int SjhipRefinerSetPrecision(struct SedonaSpatialRefiner* refiner, int32_t precision) {
  return guarded(nullptr, [&] {
    if (!refiner || !refiner->private_data) throw Error("null refiner");
    if (precision != 0 && precision != 1) throw Error("precision must be 0 or 1");
    ((RefinerImpl*)refiner->private_data)->precision.store(precision);
  });
}

int SjhipIndexSetTuning(struct SedonaFloatIndex2D* index, int32_t kind, int32_t refine_order, int32_t pip_waves) {
  return guarded(nullptr, [&] {
    if (!index || !index->private_data) throw Error("null index");
    IndexImpl* ix = (IndexImpl*)index->private_data;
    if (kind < 0 || kind > 2) throw Error("kind must be 0, 1 or 2");
    if (refine_order < -1 || refine_order > 2) throw Error("refine_order must be -1, 0, 1 or 2");
    if (pip_waves != 0 && pip_waves != 4 && pip_waves != 6 && pip_waves != 8)
      throw Error("pip_waves must be 0, 4, 6 or 8");
    std::lock_guard<std::mutex> g(ix->mu);
    ix->tuning.refine_order[kind] = refine_order;
    if (pip_waves) ix->tuning.pip_waves = pip_waves;
  });
}

int SjhipAutotune(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                  const struct ArrowSchema* probe_schema, const struct ArrowDeviceArray* probe_sample,
                  const struct SjhipJoinOptions* options, struct SjhipTuneResult* out) {
  return guarded(nullptr, [&] {
    if (!index || !index->private_data || !refiner || !refiner->private_data) throw Error("null index or refiner");
    if (!probe_sample || !options || !out) throw Error("null argument");
    IndexImpl* ix = (IndexImpl*)index->private_data;
    RefinerImpl* rf = (RefinerImpl*)refiner->private_data;
    if (!ix->built || !rf->built) throw Error("index and refiner must be built");
    if (!ix->rects.p) throw Error("autotune needs the index's build rects");
    bool dev = probe_sample->device_type == ARROW_DEVICE_ROCM;
    if (!dev && probe_sample->device_type != ARROW_DEVICE_CPU)
      throw Error("probe sample must be in ROCm device memory or CPU memory");
    use_device(rf->rt);
    WkbInput in = wkb_input_from_arrow(probe_schema, &probe_sample->array, dev);
    wait_sync_event(probe_sample);
    const bool dwithin = options->predicate == 100;
    const int64_t n_rects = (int64_t)(ix->rects.n);
    SjhipJoinOptions o = *options;
    o.profile = 0;
    o.stream = nullptr;
    o.sort_output = 0;
    int evals = 0;
    // best of two runs, wall clock of the whole join
    auto time_join = [&]() {
      double best = 1e300;
      for (int r = 0; r < 2; r++) {
        SjhipJoinResult res{};
        auto t0 = std::chrono::steady_clock::now();
        JoinEngine::run(ix, rf, in, o, &res);
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (res.release) res.release(&res);
        best = std::min(best, ms);
      }
      evals++;
      return best;
    };
    // best of three, HIP-event time of the refine stage alone (for choices
    // that change only the refine, which the join's jitter would hide)
    auto time_refine = [&]() {
      double best = 1e300;
      o.profile = 1;
      for (int r = 0; r < 3; r++) {
        SjhipJoinResult res{};
        JoinEngine::run(ix, rf, in, o, &res);
        double ms = res.stage_ms[3];
        if (res.release) res.release(&res);
        best = std::min(best, ms);
      }
      o.profile = 0;
      evals++;
      return best;
    };
    auto rebuild = [&](int leaf) {
      if (ix->tree.leaf == leaf) return;
      RectTree t;
      build_rect_tree(ix->rects.p, n_rects, t, ix->stream, leaf);
      SJ_CHECK(hipStreamSynchronize(ix->stream));
      t.cand_per_query_x16.store(ix->tree.cand_per_query_x16.load());
      t.kept_per_query_x16.store(ix->tree.kept_per_query_x16.load());
      ix->tree = std::move(t);
      ix->tree_gen++;
    };
    // the kind of refine this probe takes (as JoinEngine::run decides); the
    // decoded column is released before its stream
    int kind = 1;
    if (!dwithin) {
      ScopedStream ss;
      GeomColumn probe_types;
      decode_wkb(in, probe_types, ss.s);
      kind = pip_refine_applies(rf->build, rf->ring_index, probe_types, o.predicate) ? 0 : 2;
    }
    o.refine_order = ix->tuning.refine_order[kind] >= 0 ? ix->tuning.refine_order[kind]
                                                        : (dwithin && rf->ring_index.n_line_rasters == 0 ? 0 : 1);
    (void)time_join();  // warm-up: code objects, memory pool, output-size hints
    // 1. rect tree leaf size
    double best = 1e300;
    int best_leaf = ix->tree.leaf;
    for (int leaf : {4, 8, 16}) {
      rebuild(leaf);
      double ms = time_join();
      if (ms < best) { best = ms; best_leaf = leaf; }
    }
    rebuild(best_leaf);
    // 2. candidate order before the refine
    //    The engine's default order is kept unless another is faster by
    //    more than 5 % on the sample: near a tie the sample's timing noise
    //    would decide, and at full size the default has been the faster.
    std::vector<int> orders = {0, 1};
    if (dwithin) orders.push_back(2);
    const int default_order = dwithin && rf->ring_index.n_line_rasters == 0 ? 0 : 1;
    std::vector<double> order_ms;
    for (int ro : orders) {
      o.refine_order = ro;
      order_ms.push_back(time_join());
    }
    int best_order = default_order;
    best = order_ms[default_order];
    for (size_t k = 0; k < orders.size(); k++)
      if (order_ms[k] < 0.95 * order_ms[default_order] && order_ms[k] < best) {
        best = order_ms[k];
        best_order = orders[k];
      }
    o.refine_order = best_order;
    ix->tuning.refine_order[kind] = best_order;
    // 3. resident waves of the point-in-polygon refine
    if (kind == 0) {
      int best_w = ix->tuning.pip_waves;
      best = 1e300;
      for (int w : {4, 6, 8}) {
        ix->tuning.pip_waves = w;
        double ms = time_join();
        if (ms < best) { best = ms; best_w = w; }
      }
      ix->tuning.pip_waves = best_w;
    }
    // 4. segments per leaf of the build rings' trees (16 or 32; the ring
    //    index is rebuilt for the other size and kept when faster).  The
    //    trees serve every refine against this build side; the point-in-
    //    polygon walk is the one that prefers larger leaves.
    if (kind == 0) {
      double best_refine = time_refine();
      for (int ls : {4, 5}) {
        if (ls == rf->ring_index.seg_leaf_shift) continue;
        RingIndexParams prm;
        prm.seg_leaf_shift = ls;
        RingIndex other;
        build_ring_index(rf->build, other, prm, ix->stream);
        SJ_CHECK(hipStreamSynchronize(ix->stream));
        other.detach_streams();
        std::swap(rf->ring_index, other);
        rf->ring_gen++;
        double ms = time_refine();
        if (ms < best_refine) {
          best_refine = ms;
        } else {
          std::swap(rf->ring_index, other);
          rf->ring_gen++;
        }
      }
      best = time_join();
    }
    // give back the pool memory the sample joins left cached, so that the
    // first full-size joins do not find it fragmented
    SJ_CHECK(hipStreamSynchronize(ix->stream));
    SJ_CHECK(hipStreamSynchronize(detached_stream()));
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, rf->rt->device_id) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
    out->leaf = ix->tree.leaf;
    out->refine_order = best_order;
    out->pip_waves = ix->tuning.pip_waves;
    out->seg_leaf = 1 << rf->ring_index.seg_leaf_shift;
    out->evaluations = evals;
    out->best_ms = best;
  });
}

int SjhipServerCreate(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner,
                      const struct SjhipServerConfig* config, void** out_server) {
  return guarded(nullptr, [&] {
    if (!index || !index->private_data || !refiner || !refiner->private_data) throw Error("null index or refiner");
    if (!config || !out_server) throw Error("null argument");
    if (config->slots < 1 || config->slots > 64) throw Error("slots must be in 1..64");
    *out_server = nullptr;
    std::unique_ptr<Server> sv(new Server());
    sv->ix = (IndexImpl*)index->private_data;
    sv->rf = (RefinerImpl*)refiner->private_data;
    sv->graphs = config->graphs != 0;
    if (!sv->ix->built || !sv->rf->built) throw Error("index and refiner must be built");
    use_device(sv->rf->rt);
    sv->budget = config->memory_budget;
    if (sv->budget == 0) {
      size_t free_b = 0, total_b = 0;
      SJ_CHECK(hipMemGetInfo(&free_b, &total_b));
      sv->budget = (uint64_t)(0.9 * (double)free_b);
    }
    Server* p = sv.get();
    for (uint32_t k = 0; k < config->slots; k++)
      sv->workers.emplace_back([p] {
        try {
          p->work();
        } catch (...) {
        }
      });
    *out_server = sv.release();
  });
}

int SjhipServerSubmit(void* server, const struct ArrowSchema* probe_schema, const struct ArrowDeviceArray* probe_array,
                      const struct SjhipJoinOptions* options, uint64_t* out_request_id) {
  return guarded(nullptr, [&] {
    Server* sv = (Server*)server;
    if (!sv || !probe_array || !options || !out_request_id) throw Error("null argument");
    bool dev = probe_array->device_type == ARROW_DEVICE_ROCM;
    if (!dev && probe_array->device_type != ARROW_DEVICE_CPU)
      throw Error("probe array must be in ROCm device memory or CPU memory");
    std::unique_ptr<ServerReq> r(new ServerReq());
    r->in = wkb_input_from_arrow(probe_schema, &probe_array->array, dev);
    wait_sync_event(probe_array);
    r->opt = *options;
    r->estimate = sv->estimate(r->in, r->opt);
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (sv->stop) throw Error("server is shutting down");
      r->id = sv->next_id++;
      r->t_submit = sv->now_ms();
      *out_request_id = r->id;
      sv->queue.push_back(r.get());
      sv->reqs[r->id] = std::move(r);
    }
    sv->cv.notify_all();
  });
}

int SjhipServerWait(void* server, uint64_t request_id, int64_t timeout_ms, struct SjhipJoinResult* out,
                    struct SjhipRequestTiming* timing) {
  int req_rc = 0;
  std::string req_err;
  int rc = guarded(nullptr, [&] {
    Server* sv = (Server*)server;
    if (!sv || !out) throw Error("null argument");
    std::memset(out, 0, sizeof(*out));
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(request_id);
    if (it == sv->reqs.end()) throw Error("unknown request id " + std::to_string(request_id));
    ServerReq* r = it->second.get();
    auto done = [&] { return r->status >= 2; };
    if (timeout_ms < 0) sv->cv.wait(lk, done);
    else if (timeout_ms > 0) sv->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
    if (!done()) {
      req_rc = EAGAIN;
      return;
    }
    if (timing) {
      timing->submit_ms = r->t_submit;
      timing->admit_ms = r->t_admit;
      timing->done_ms = r->t_done;
      timing->estimate_bytes = r->estimate;
    }
    if (r->status == 2) {
      *out = r->res;
    } else {
      req_rc = r->rc ? r->rc : EINVAL;
      req_err = r->error;
    }
    sv->reqs.erase(it);
  });
  if (rc) return rc;
  if (req_rc && req_rc != EAGAIN) t_last_error = req_err;
  return req_rc;
}

int SjhipServerDestroy(void* server) {
  return guarded(nullptr, [&] {
    Server* sv = (Server*)server;
    if (!sv) return;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->stop = true;
    }
    sv->cv.notify_all();
    for (auto& t : sv->workers) t.join();
    sv->retired.clear();  // no worker is capturing any more
    for (auto& kv : sv->reqs)
      if (kv.second->status == 2 && kv.second->res.release) kv.second->res.release(&kv.second->res);
    delete sv;
  });
}

int SjhipDeviceAlloc(uint64_t bytes, void** out_ptr) {
  return guarded(nullptr, [&] {
    if (!out_ptr) throw Error("null output pointer");
    *out_ptr = nullptr;
    hipStream_t ds = detached_stream();
    SJ_CHECK(hipMallocAsync(out_ptr, bytes ? bytes : 1, ds));
    SJ_CHECK(hipStreamSynchronize(ds));
  });
}

int SjhipDeviceFree(void* ptr) {
  return guarded(nullptr, [&] {
    if (ptr) {
      hipStream_t ds = detached_stream();
      SJ_CHECK(hipFreeAsync(ptr, ds));
      SJ_CHECK(hipStreamSynchronize(ds));
    }
  });
}

int SjhipCopyToDevice(void* dst, const void* src, uint64_t bytes) {
  return guarded(nullptr, [&] {
    if (bytes) {
      ScopedStream ss;
      SJ_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ss.s));
      SJ_CHECK(hipStreamSynchronize(ss.s));
    }
  });
}

int SjhipCopyToHost(void* dst, const void* src, uint64_t bytes) {
  return guarded(nullptr, [&] {
    if (bytes) {
      ScopedStream ss;
      SJ_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ss.s));
      SJ_CHECK(hipStreamSynchronize(ss.s));
    }
  });
}

int SjhipSetDevice(int device_id) {
  return guarded(nullptr, [&] { SJ_CHECK(hipSetDevice(device_id)); });
}

int SjhipCopyKernel(void* dst, const void* src, uint64_t bytes, int reps, double* ms) {
  return guarded(nullptr, [&] {
    if (!dst || !src || !ms || reps < 1 || bytes == 0 || bytes % 16) throw Error("bad copy arguments");
    hipStream_t s;
    SJ_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    SJ_CHECK(hipEventCreate(&e0));
    SJ_CHECK(hipEventCreate(&e1));
    uint64_t n = bytes / 16;
    SJ_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) {
      hipLaunchKernelGGL(k_copy16, 256 * 32, 256, 0, s, (uint4*)dst, (const uint4*)src, n);
      SJ_CHECK(hipGetLastError());
    }
    SJ_CHECK(hipEventRecord(e1, s));
    SJ_CHECK(hipEventSynchronize(e1));
    float t = 0;
    SJ_CHECK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
  });
}

int SjhipJoinResultCopyToHost(const struct SjhipJoinResult* r, uint32_t* b, uint32_t* p) {
  return guarded(nullptr, [&] {
    if (!r) throw Error("null result");
    if (r->size == 0) return;
    ScopedStream ss;
    SJ_CHECK(hipMemcpyAsync(b, r->build_indices, r->size * 4, hipMemcpyDeviceToHost, ss.s));
    SJ_CHECK(hipMemcpyAsync(p, r->probe_indices, r->size * 4, hipMemcpyDeviceToHost, ss.s));
    SJ_CHECK(hipStreamSynchronize(ss.s));
  });
}

// Builds the index's tree from the refiner's decoded build geometry (rects
// computed on the device with the operator's semantics) instead of from
// host rects: the device-resident counterpart of
// rust/sedona-spatial-join-gpu/src/index/gpu_spatial_index_builder.rs:193-259.
int SjhipIndexBuildFromRefiner(struct SedonaFloatIndex2D* index, struct SedonaSpatialRefiner* refiner) {
  return guarded(nullptr, [&] {
    IndexImpl* ix = index_of(index);
    RefinerImpl* rf = refiner_of(refiner);
    if (!rf->built) throw Error("refiner is not built");
    use_device(ix->rt);
    compute_rects(rf->build, &rf->ring_index, true, NAN, ix->rects, ix->stream);
    build_rect_tree(ix->rects.p, rf->build.n_rows, ix->tree, ix->stream);
    ix->tree_gen++;
    SJ_CHECK(hipStreamSynchronize(ix->stream));
    ix->host_rects.clear();
    ix->built = true;
  });
}

// Device-resident build push: decodes an Arrow array whose buffers are in
// device (or host) memory.
int SjhipRefinerPushBuildDevice(struct SedonaSpatialRefiner* refiner, const struct ArrowSchema* schema,
                                const struct ArrowDeviceArray* array) {
  return guarded(nullptr, [&] {
    RefinerImpl* rf = refiner_of(refiner);
    if (!array) throw Error("null array");
    bool dev = array->device_type == ARROW_DEVICE_ROCM;
    wait_sync_event(array);
    refiner_push(rf, wkb_input_from_arrow(schema, &array->array, dev));
  });
}

int SjhipComputeRects(const struct ArrowSchema* schema, const struct ArrowArray* array, int point_nearest,
                      double expand_distance, float* out_rects) {
  return guarded(nullptr, [&] {
    if (!out_rects) throw Error("null output");
    WkbInput in = wkb_input_from_arrow(schema, array, false);
    int n_dev = 0;
    SJ_CHECK(hipGetDevice(&n_dev));
    ScopedStream ss;
    GeomColumn g;
    decode_wkb(in, g, ss.s);
    DevArray<float4> r;
    compute_rects(g, nullptr, point_nearest != 0, expand_distance, r, ss.s);
    if (g.n_rows)
      SJ_CHECK(hipMemcpyAsync(out_rects, r.p, g.n_rows * sizeof(float4), hipMemcpyDeviceToHost, ss.s));
    SJ_CHECK(hipStreamSynchronize(ss.s));
  });
}

}  // extern "C"
