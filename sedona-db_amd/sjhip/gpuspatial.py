"""Safe host-side wrappers over the C ABI, mirroring the reference's Rust
FFI layer (c/sedona-libgpuspatial/src/lib.rs:41-191,
c/sedona-libgpuspatial/src/libgpuspatial.rs:239-526,
c/sedona-libgpuspatial/src/options.rs:19-32,
c/sedona-libgpuspatial/src/predicate.rs:21-46): a process-global runtime
per device, ``GpuSpatialIndex`` (FP32 rect filter) and ``GpuSpatialRefiner``
(exact refine), errors raised with the library's message.
"""
import ctypes
import errno
import enum
import threading

import numpy as np

from . import _lib as L


class GpuSpatialError(RuntimeError):
    pass


class GpuSpatialRelationPredicate(enum.IntEnum):
    """predicate.rs:21-46 / gpuspatial_c.h:142-151"""
    Equals = 0
    Disjoint = 1
    Touches = 2
    Contains = 3
    Covers = 4
    Intersects = 5
    Within = 6
    CoveredBy = 7


PREDICATE_BY_NAME = {p.name.lower(): p for p in GpuSpatialRelationPredicate}


class GpuSpatialOptions:
    """c/sedona-libgpuspatial/src/options.rs:19-32.  The pool percentage
    defaults to the operator's 50 (rust/sedona-spatial-join-gpu/src/options.rs:39):
    the library keeps that share of HBM cached in its stream-ordered pool
    between joins instead of returning it to the driver."""

    def __init__(self, concurrency=1, device_id=0, compress_bvh=False, pipeline_batches=1,
                 cuda_use_memory_pool=True, cuda_memory_pool_init_percent=50):
        self.concurrency = concurrency
        self.device_id = device_id
        self.compress_bvh = compress_bvh
        self.pipeline_batches = pipeline_batches
        self.cuda_use_memory_pool = cuda_use_memory_pool
        self.cuda_memory_pool_init_percent = cuda_memory_pool_init_percent


# ---------------------------------------------------------------------------
# Process-global runtime (lib.rs:41-90): one per process; a different device
# id than the first one is an error.

_runtime_lock = threading.Lock()
_runtime = None
_runtime_device = None


def _check(rc, msg_fn):
    if rc != 0:
        raise GpuSpatialError(msg_fn())


def global_runtime(options):
    global _runtime, _runtime_device
    with _runtime_lock:
        if _runtime is not None:
            if _runtime_device != options.device_id:
                raise GpuSpatialError(f"GPU runtime already initialised on device {_runtime_device}, "
                                      f"requested device {options.device_id}")
            return _runtime
        lib = L.lib()
        rt = L.GpuSpatialRuntime()
        lib.GpuSpatialRuntimeCreate(ctypes.byref(rt))
        cfg = L.GpuSpatialRuntimeConfig(None, options.device_id, options.cuda_use_memory_pool,
                                        options.cuda_memory_pool_init_percent)
        rc = rt.init(ctypes.byref(rt), ctypes.byref(cfg))
        if rc:
            msg = rt.get_last_error(ctypes.byref(rt)).decode()
            rt.release(ctypes.byref(rt))
            raise GpuSpatialError(f"runtime init failed: {msg}")
        _runtime = rt
        _runtime_device = options.device_id
        return rt


class Runtime:
    """A standalone runtime (the C-ABI object itself), for ABI tests."""

    def __init__(self, ptx_root=None, device_id=0, use_pool=False, pool_percent=10):
        self.lib = L.lib()
        self.rt = L.GpuSpatialRuntime()
        self.lib.GpuSpatialRuntimeCreate(ctypes.byref(self.rt))
        self.cfg = L.GpuSpatialRuntimeConfig(ptx_root.encode() if ptx_root else None, device_id, use_pool,
                                             pool_percent)

    def init(self):
        return self.rt.init(ctypes.byref(self.rt), ctypes.byref(self.cfg))

    def last_error(self):
        return self.rt.get_last_error(ctypes.byref(self.rt)).decode()

    def release(self):
        if self.rt.private_data:
            self.rt.release(ctypes.byref(self.rt))


# ---------------------------------------------------------------------------
# Arrow C data export


class ExportedArrow:
    """A pyarrow array exported through the Arrow C data interface; the
    library only borrows it for the duration of each call."""

    def __init__(self, arr):
        import pyarrow as pa
        if isinstance(arr, pa.ChunkedArray):
            arr = arr.combine_chunks()
        if not isinstance(arr, pa.Array):
            arr = arr.to_arrow()  # datagen.WkbColumn
        self.arr = arr
        self.schema = L.ArrowSchema()
        self.array = L.ArrowArray()
        arr._export_to_c(ctypes.addressof(self.array), ctypes.addressof(self.schema))

    def __del__(self):
        try:
            if self.array.release:
                self.array.release(ctypes.byref(self.array))
            if self.schema.release:
                self.schema.release(ctypes.byref(self.schema))
        except Exception:
            pass


_noop_release = ctypes.CFUNCTYPE(None, ctypes.POINTER(L.ArrowArray))(lambda p: None)


class DeviceWkb:
    """An Arrow LargeBinary WKB column whose offsets (int64) and data (uint8)
    live in device memory (e.g. torch tensors on cuda), passed as an
    ArrowDeviceArray with device_type ROCm."""

    def __init__(self, offsets_ptr, data_ptr, length, device_id=0, keepalive=None):
        self._keep = keepalive
        self.bufs = (ctypes.c_void_p * 3)(None, offsets_ptr, data_ptr)
        self.schema = L.ArrowSchema()
        self.schema.format = b"Z"
        self.dev = L.ArrowDeviceArray()
        a = self.dev.array
        a.length = length
        a.null_count = 0
        a.offset = 0
        a.n_buffers = 3
        a.n_children = 0
        a.buffers = ctypes.cast(self.bufs, ctypes.POINTER(ctypes.c_void_p))
        a.release = _noop_release
        self.dev.device_id = device_id
        self.dev.device_type = L.ARROW_DEVICE_ROCM

    def slice(self, start, stop):
        """Rows [start, stop) as a DeviceWkb over the same device buffers
        (Arrow offsets need not start at 0)."""
        start, stop = int(start), int(stop)
        if not (0 <= start <= stop <= self.dev.array.length):
            raise ValueError("slice out of range")
        return DeviceWkb(self.bufs[1] + 8 * start, self.bufs[2], stop - start, device_id=self.dev.device_id,
                         keepalive=self)

    def arrow_slice(self, start, stop):
        """Rows [start, stop) as an Arrow slice: the same buffers with the
        array's `offset` set to start (the form pyarrow's slice exports)."""
        start, stop = int(start), int(stop)
        if not (0 <= start <= stop <= self.dev.array.length):
            raise ValueError("slice out of range")
        out = DeviceWkb(self.bufs[1], self.bufs[2], stop - start, device_id=self.dev.device_id, keepalive=self)
        out.dev.array.offset = self.dev.array.offset + start
        return out

    @classmethod
    def from_host(cls, col, device_id=0):
        """Copies a datagen.WkbColumn (int64 offsets + data) into device
        buffers owned by the returned object."""
        off = DeviceBuffer.upload(np.ascontiguousarray(col.offsets, dtype=np.int64), device_id)
        data = DeviceBuffer.upload(np.ascontiguousarray(col.data, dtype=np.uint8), device_id)
        return cls(off.ptr, data.ptr, len(col), device_id=device_id, keepalive=(off, data))


class DeviceBuffer:
    """Device memory owned through the library's plumbing entry points."""

    def __init__(self, nbytes, device_id=0):
        lib = L.lib()
        rc = lib.SjhipSetDevice(device_id)
        if rc:
            raise GpuSpatialError(lib.SjhipLastError().decode())
        p = ctypes.c_void_p()
        rc = lib.SjhipDeviceAlloc(int(nbytes), ctypes.byref(p))
        if rc:
            raise GpuSpatialError(lib.SjhipLastError().decode())
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def upload(cls, arr, device_id=0):
        b = cls(max(arr.nbytes, 1), device_id)
        rc = L.lib().SjhipCopyToDevice(b.ptr, arr.ctypes.data, arr.nbytes)
        if rc:
            raise GpuSpatialError(L.lib().SjhipLastError().decode())
        return b

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        rc = L.lib().SjhipCopyToHost(out.ctypes.data, self.ptr, out.nbytes)
        if rc:
            raise GpuSpatialError(L.lib().SjhipLastError().decode())
        return out

    def free(self):
        if self.ptr:
            L.lib().SjhipDeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_kernel_ms(nbytes, reps=10, device_id=0):
    """Mean time (ms) of the library's plain copy kernel over nbytes
    (2 * nbytes of HBM traffic per launch)."""
    nbytes = int(nbytes) // 16 * 16
    src, dst = DeviceBuffer(nbytes, device_id), DeviceBuffer(nbytes, device_id)
    ms = ctypes.c_double()
    rc = L.lib().SjhipCopyKernel(dst.ptr, src.ptr, nbytes, int(reps), ctypes.byref(ms))
    src.free()
    dst.free()
    if rc:
        raise GpuSpatialError(L.lib().SjhipLastError().decode())
    return ms.value


def host_device_array(exported):
    """Wraps an ExportedArrow as an ArrowDeviceArray with device_type CPU."""
    d = L.ArrowDeviceArray()
    d.array = exported.array
    d.device_id = -1
    d.device_type = L.ARROW_DEVICE_CPU
    return d


def compute_rects(arr, point_nearest=True, expand_distance=float("nan")):
    """Device-computed FP32 rects of a WKB column (n x 4, NaN for empty)."""
    ex = ExportedArrow(arr)
    out = np.zeros((ex.array.length, 4), np.float32)
    rc = L.lib().SjhipComputeRects(ctypes.byref(ex.schema), ctypes.byref(ex.array), 1 if point_nearest else 0,
                                   float(expand_distance), out.ctypes.data)
    if rc:
        raise GpuSpatialError(L.lib().SjhipLastError().decode())
    return out


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr32(a):
    return a.ctypes.data_as(L.c_u32p)


# ---------------------------------------------------------------------------
# Index (lib.rs:94-134)


class GpuSpatialIndex:
    def __init__(self, options=None, runtime=None):
        self.options = options or GpuSpatialOptions()
        self.lib = L.lib()
        rt = runtime.rt if runtime is not None else global_runtime(self.options)
        self._rt = rt
        self.cfg = L.GpuSpatialIndexConfig(ctypes.pointer(rt), self.options.concurrency)
        self.ix = L.SedonaFloatIndex2D()
        rc = self.lib.GpuSpatialIndexFloat2DCreate(ctypes.byref(self.ix), ctypes.byref(self.cfg))
        if rc:
            raise GpuSpatialError(f"index create failed: {self.lib.SjhipLastError().decode()}")

    try_new = classmethod(lambda cls, options=None: cls(options))

    def _err(self):
        return self.ix.get_last_error(ctypes.byref(self.ix)).decode()

    def clear(self):
        _check(self.ix.clear(ctypes.byref(self.ix)), self._err)

    def push_build(self, rects):
        r = np.ascontiguousarray(rects, dtype=np.float32).reshape(-1, 4)
        _check(self.ix.push_build(ctypes.byref(self.ix), r.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(r)),
               self._err)

    def finish_building(self):
        _check(self.ix.finish_building(ctypes.byref(self.ix)), self._err)

    def build_from_refiner(self, refiner):
        rc = self.lib.SjhipIndexBuildFromRefiner(ctypes.byref(self.ix), ctypes.byref(refiner.rf))
        if rc:
            raise GpuSpatialError(self.lib.SjhipLastError().decode())

    def probe(self, rects, callback_rc=0):
        """(build_indices, probe_indices) of all candidate pairs
        (libgpuspatial.rs:239-308 + 532-572)."""
        r = np.ascontiguousarray(rects, dtype=np.float32).reshape(-1, 4)
        outs_b, outs_p = [], []

        def cb(b, p, n, ud):
            if n:
                outs_b.append(np.ctypeslib.as_array(b, shape=(n,)).copy())
                outs_p.append(np.ctypeslib.as_array(p, shape=(n,)).copy())
            return callback_rc

        ccb = L.PROBE_CB(cb)
        ctx = L.SedonaSpatialIndexContext()
        self.ix.create_context(ctypes.byref(ctx))
        try:
            rc = self.ix.probe(ctypes.byref(self.ix), ctypes.byref(ctx),
                               r.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(r), ccb, None)
            if rc:
                msg = self.ix.context_get_last_error(ctypes.byref(ctx)).decode()
                raise GpuSpatialError(f"probe failed (rc={rc}): {msg}")
        finally:
            self.ix.destroy_context(ctypes.byref(ctx))
        if not outs_b:
            return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
        return np.concatenate(outs_b), np.concatenate(outs_p)

    def set_tuning(self, kind=0, refine_order=-1, pip_waves=0):
        """Sets the tuned candidate order for joins of refine `kind` and the
        point-in-polygon kernel's resident waves (0 keeps them)."""
        rc = self.lib.SjhipIndexSetTuning(ctypes.byref(self.ix), int(kind), int(refine_order), int(pip_waves))
        if rc:
            raise GpuSpatialError(self.lib.SjhipLastError().decode())

    def set_filter_mode(self, mode):
        """0: grid for point probes, tree otherwise, and the build-rect pass
        (polygon-major) in fused point-in-polygon joins; 1: tree; 2: grid;
        3: as 0 without the build-rect pass."""
        rc = self.lib.SjhipIndexSetFilterMode(ctypes.byref(self.ix), int(mode))
        if rc:
            raise GpuSpatialError(self.lib.SjhipLastError().decode())

    def release(self):
        if self.ix.private_data:
            self.ix.release(ctypes.byref(self.ix))

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Refiner (lib.rs:137-191)


class GpuSpatialRefiner:
    def __init__(self, options=None, runtime=None):
        self.options = options or GpuSpatialOptions()
        self.lib = L.lib()
        rt = runtime.rt if runtime is not None else global_runtime(self.options)
        self._rt = rt
        self.cfg = L.GpuSpatialRefinerConfig(ctypes.pointer(rt), self.options.concurrency, self.options.compress_bvh,
                                             self.options.pipeline_batches)
        self.rf = L.SedonaSpatialRefiner()
        rc = self.lib.GpuSpatialRefinerCreate(ctypes.byref(self.rf), ctypes.byref(self.cfg))
        if rc:
            raise GpuSpatialError(f"refiner create failed: {self.lib.SjhipLastError().decode()}")

    try_new = classmethod(lambda cls, options=None: cls(options))

    def _err(self):
        return self.rf.get_last_error(ctypes.byref(self.rf)).decode()

    def clear(self):
        _check(self.rf.clear(ctypes.byref(self.rf)), self._err)

    def init_build_schema(self, format_or_array):
        if isinstance(format_or_array, (bytes, str)):
            sch = L.ArrowSchema()
            sch.format = format_or_array.encode() if isinstance(format_or_array, str) else format_or_array
            _check(self.rf.init_build_schema(ctypes.byref(self.rf), ctypes.byref(sch)), self._err)
        else:
            ex = ExportedArrow(format_or_array)
            _check(self.rf.init_build_schema(ctypes.byref(self.rf), ctypes.byref(ex.schema)), self._err)

    def push_build(self, arr):
        ex = ExportedArrow(arr)
        _check(self.rf.push_build(ctypes.byref(self.rf), ctypes.byref(ex.array)), self._err)

    def push_build_device(self, dev):
        rc = self.lib.SjhipRefinerPushBuildDevice(ctypes.byref(self.rf), ctypes.byref(dev.schema),
                                                  ctypes.byref(dev.dev))
        if rc:
            raise GpuSpatialError(self.lib.SjhipLastError().decode())

    def finish_building(self):
        _check(self.rf.finish_building(ctypes.byref(self.rf)), self._err)

    def set_segment_leaf(self, segments_per_leaf):
        """Rebuilds the ring index with 16 or 32 segments per tree leaf."""
        lib = self.lib
        rc = lib.SjhipRefinerSetSegmentLeaf(ctypes.byref(self.rf), int(segments_per_leaf))
        if rc:
            raise GpuSpatialError(lib.SjhipLastError().decode())

    def set_precision(self, precision):
        """0: certified FP32 with FP64 fallback (default); 1: all-FP64 plain
        kernels, for refine and refine_distance."""
        lib = self.lib
        rc = lib.SjhipRefinerSetPrecision(ctypes.byref(self.rf), int(precision))
        if rc:
            raise GpuSpatialError(lib.SjhipLastError().decode())

    def refine(self, probe_array, predicate, build_indices, probe_indices):
        """Filters the candidate pairs in place (libgpuspatial.rs:453-486):
        returns the kept (build, probe) arrays sorted by (probe, build)."""
        ex = ExportedArrow(probe_array)
        b = _u32(build_indices).copy()
        p = _u32(probe_indices).copy()
        new_n = ctypes.c_uint32(0)
        pred = PREDICATE_BY_NAME[predicate] if isinstance(predicate, str) else predicate
        _check(self.rf.refine(ctypes.byref(self.rf), ctypes.byref(ex.schema), ctypes.byref(ex.array), int(pred),
                              _ptr32(b), _ptr32(p), len(b), ctypes.byref(new_n)), self._err)
        return b[:new_n.value], p[:new_n.value]

    def refine_distance(self, probe_array, distance, build_indices, probe_indices):
        ex = ExportedArrow(probe_array)
        b = _u32(build_indices).copy()
        p = _u32(probe_indices).copy()
        new_n = ctypes.c_uint32(0)
        _check(self.lib.SjhipRefinerRefineDistance(ctypes.byref(self.rf), ctypes.byref(ex.schema),
                                                   ctypes.byref(ex.array), float(distance), _ptr32(b), _ptr32(p),
                                                   len(b), ctypes.byref(new_n)), self._err)
        return b[:new_n.value], p[:new_n.value]

    def knn(self, probe_array, k):
        ex = ExportedArrow(probe_array)
        n = ex.array.length
        ids = np.zeros((n, k), np.uint32)
        d = np.zeros((n, k), np.float64)
        _check(self.lib.SjhipRefinerKnn(ctypes.byref(self.rf), ctypes.byref(ex.schema), ctypes.byref(ex.array), k,
                                        _ptr32(ids), d.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), self._err)
        return ids, d

    def pair_distances(self, probe_array, build_indices, probe_indices):
        """FP64 reference distances of (build row, probe point) pairs."""
        ex = ExportedArrow(probe_array)
        b = np.ascontiguousarray(build_indices, np.uint32)
        p = np.ascontiguousarray(probe_indices, np.uint32)
        out = np.zeros(max(len(b), 1), np.float64)
        _check(self.lib.SjhipRefinerPairDistances(ctypes.byref(self.rf), ctypes.byref(ex.schema), ctypes.byref(ex.array),
                                                  b.ctypes.data, p.ctypes.data, len(b), out.ctypes.data), self._err)
        return out[:len(b)]

    def knn_ties(self, probe_array, k, include_tie_breakers=True):
        """KNN with variable-length rows (SjhipRefinerKnnTies): (offsets,
        ids, distances); probe q's neighbours are ids[offsets[q]:offsets[q+1]]."""
        ex = ExportedArrow(probe_array)
        n = ex.array.length
        cap = max(1, n * k + 1024)
        for _ in range(2):
            off = np.zeros(n + 1, np.uint64)
            ids = np.zeros(cap, np.uint32)
            d = np.zeros(cap, np.float64)
            total = ctypes.c_uint64()
            rc = self.lib.SjhipRefinerKnnTies(ctypes.byref(self.rf), ctypes.byref(ex.schema), ctypes.byref(ex.array),
                                              int(k), 1 if include_tie_breakers else 0, off.ctypes.data,
                                              ids.ctypes.data, d.ctypes.data, cap, ctypes.byref(total))
            if rc == 0:
                t = int(total.value)
                return off, ids[:t], d[:t]
            if int(total.value) <= cap:
                break
            cap = int(total.value)
        raise GpuSpatialError(self._err())

    def knn_device(self, probe, k):
        """KNN of a DeviceWkb probe column with device-resident outputs:
        (ids DeviceBuffer, distances DeviceBuffer, kernel ms)."""
        n = probe.dev.array.length
        ids = DeviceBuffer(max(1, n * k * 4))
        d = DeviceBuffer(max(1, n * k * 8))
        ms = ctypes.c_double()
        _check(self.lib.SjhipRefinerKnnDevice(ctypes.byref(self.rf), ctypes.byref(probe.schema),
                                              ctypes.byref(probe.dev), int(k), ids.ptr, d.ptr, ctypes.byref(ms)),
               self._err)
        return ids, d, ms.value

    def release(self):
        if self.rf.private_data:
            self.rf.release(ctypes.byref(self.rf))

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Device-resident fused join


class JoinResult:
    def __init__(self, res):
        self.res = res

    @property
    def size(self):
        return self.res.size

    @property
    def candidates(self):
        return self.res.candidates

    @property
    def segment_tests(self):
        return self.res.segment_tests

    @property
    def refine_candidates(self):
        return self.res.refine_candidates

    @property
    def fallback_tests(self):
        return self.res.fallback_tests

    def to_host(self):
        n = self.res.size
        b = np.zeros(max(n, 1), np.uint32)
        p = np.zeros(max(n, 1), np.uint32)
        rc = L.lib().SjhipJoinResultCopyToHost(ctypes.byref(self.res), _ptr32(b), _ptr32(p))
        if rc:
            raise GpuSpatialError(L.lib().SjhipLastError().decode())
        return b[:n], p[:n]

    def release(self):
        if self.res.private_data:
            self.res.release(ctypes.byref(self.res))

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _join_options(predicate, distance, precision, sort_output, sort_probes, profile, refine_order, stream):
    if isinstance(predicate, str):
        pred = 100 if predicate == "dwithin" else int(PREDICATE_BY_NAME[predicate])
    else:
        pred = int(predicate)
    return L.SjhipJoinOptions(pred, float(distance), int(precision), 1 if sort_output else 0, 1 if sort_probes else 0,
                              1 if profile else 0, int(refine_order), stream)


def join_device(index, refiner, probe, predicate, distance=0.0, precision=0, sort_output=True, sort_probes=True,
                profile=False, refine_order=1, stream=None):
    """One probe batch through the fused device join.  `probe` is a
    DeviceWkb (device memory) or any Arrow-exportable host column."""
    opts = _join_options(predicate, distance, precision, sort_output, sort_probes, profile, refine_order, stream)
    res = L.SjhipJoinResult()
    if isinstance(probe, DeviceWkb):
        sch, dev = probe.schema, probe.dev
        keep = None
    else:
        keep = ExportedArrow(probe)
        sch, dev = keep.schema, host_device_array(keep)
    rc = L.lib().SjhipJoinDevice(ctypes.byref(index.ix), ctypes.byref(refiner.rf), ctypes.byref(sch),
                                 ctypes.byref(dev), ctypes.byref(opts), ctypes.byref(res))
    del keep
    if rc:
        raise GpuSpatialError(L.lib().SjhipLastError().decode())
    return JoinResult(res)


def join_host_stream(index, refiner, probe, predicate, batch_rows=1_000_000, slots=2, distance=0.0, precision=0,
                     sort_output=False, sort_probes=True, refine_order=-1, callback_rc=0):
    """The pipelined join of a host column (SjhipJoinHostStream): batches of
    batch_rows rows, `slots` in flight.  Returns (build, probe) host arrays,
    probe indices global, batches in order."""
    opts = _join_options(predicate, distance, precision, sort_output, sort_probes, False, refine_order, None)
    ex = ExportedArrow(probe)
    outs_b, outs_p = [], []

    def cb(b, p, n, ud):
        if n:
            outs_b.append(np.ctypeslib.as_array(b, shape=(n,)).copy())
            outs_p.append(np.ctypeslib.as_array(p, shape=(n,)).copy())
        return callback_rc

    ccb = L.PROBE_CB(cb)
    rc = L.lib().SjhipJoinHostStream(ctypes.byref(index.ix), ctypes.byref(refiner.rf), ctypes.byref(ex.schema),
                                     ctypes.byref(ex.array), ctypes.byref(opts), int(batch_rows), int(slots), ccb,
                                     None)
    del ex
    if rc:
        raise GpuSpatialError(f"streaming join failed (rc={rc}): {L.lib().SjhipLastError().decode()}")
    if not outs_b:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    return np.concatenate(outs_b), np.concatenate(outs_p)


def partitioned_join(build, probe, predicate, device_budget_bytes, distance=0.0, precision=0, sort_output=False,
                     refine_order=-1, options=None, callback_rc=0):
    """A join whose build side is split into partitions of at most
    device_budget_bytes resident bytes, built and joined one at a time
    (SjhipPartitionedJoin).  `build` is an Arrow-exportable host column,
    `probe` a DeviceWkb or a host column.  Returns (build, probe) host
    arrays in partition order and the call's statistics."""
    opts = _join_options(predicate, distance, precision, sort_output, True, False, refine_order, None)
    rt = global_runtime(options or GpuSpatialOptions())
    bex = ExportedArrow(build)
    if isinstance(probe, DeviceWkb):
        pkeep, psch, pdev = None, probe.schema, probe.dev
    else:
        pkeep = ExportedArrow(probe)
        psch, pdev = pkeep.schema, host_device_array(pkeep)
    outs_b, outs_p = [], []

    def cb(b, p, n, ud):
        if n:
            outs_b.append(np.ctypeslib.as_array(b, shape=(n,)).copy())
            outs_p.append(np.ctypeslib.as_array(p, shape=(n,)).copy())
        return callback_rc

    ccb = L.PROBE_CB(cb)
    st = L.SjhipPartitionStats()
    rc = L.lib().SjhipPartitionedJoin(ctypes.byref(rt), ctypes.byref(bex.schema), ctypes.byref(bex.array),
                                      ctypes.byref(psch), ctypes.byref(pdev), ctypes.byref(opts),
                                      int(device_budget_bytes), ccb, None, ctypes.byref(st))
    del bex, pkeep
    if rc:
        raise GpuSpatialError(f"partitioned join failed (rc={rc}): {L.lib().SjhipLastError().decode()}")
    stats = {k: getattr(st, k) for k, _ in L.SjhipPartitionStats._fields_}
    if not outs_b:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32), stats
    return np.concatenate(outs_b), np.concatenate(outs_p), stats


# SJHIP_NOALLOC_* of include/gpuspatial_hip.h
NOALLOC_MAX_ROWS = 65536
NOALLOC_MAX_CANDIDATES = 1 << 20


def join_device_no_alloc(index, refiner, probe, predicate, out_build, out_probe, capacity, distance=0.0):
    """The fused join without allocation (SjhipJoinDeviceNoAlloc): `probe`
    is a DeviceWkb point column, out_build / out_probe device buffers (for
    example DeviceBuffer) of `capacity` u32 entries.  Returns the number of
    pairs written, in refine order."""
    opts = _join_options(predicate, distance, 0, False, True, False, 1, None)
    n = ctypes.c_uint64(0)
    ob = out_build.ptr if isinstance(out_build, DeviceBuffer) else out_build
    op = out_probe.ptr if isinstance(out_probe, DeviceBuffer) else out_probe
    rc = L.lib().SjhipJoinDeviceNoAlloc(ctypes.byref(index.ix), ctypes.byref(refiner.rf), ctypes.byref(probe.schema),
                                        ctypes.byref(probe.dev), ctypes.byref(opts), ob, op, int(capacity),
                                        ctypes.byref(n))
    if rc:
        err = GpuSpatialError(L.lib().SjhipLastError().decode())
        err.pairs = int(n.value)  # the pair count when only the capacity was short
        raise err
    return int(n.value)


def autotune(index, refiner, probe_sample, predicate, distance=0.0, precision=0):
    """Tunes the index for joins like join_device(index, refiner, probe,
    predicate, refine_order=-1): returns the chosen parameters."""
    opts = _join_options(predicate, distance, precision, False, True, False, -1, None)
    if isinstance(probe_sample, DeviceWkb):
        keep, sch, dev = None, probe_sample.schema, probe_sample.dev
    else:
        keep = ExportedArrow(probe_sample)
        sch, dev = keep.schema, host_device_array(keep)
    out = L.SjhipTuneResult()
    rc = L.lib().SjhipAutotune(ctypes.byref(index.ix), ctypes.byref(refiner.rf), ctypes.byref(sch), ctypes.byref(dev),
                               ctypes.byref(opts), ctypes.byref(out))
    del keep
    if rc:
        raise GpuSpatialError(L.lib().SjhipLastError().decode())
    return {"leaf": out.leaf, "refine_order": out.refine_order, "pip_waves": out.pip_waves,
            "seg_leaf": out.seg_leaf, "evaluations": out.evaluations, "best_ms": out.best_ms}


class GpuSpatialServer:
    """Request loop over a built index + refiner: up to `slots` probe
    batches joined concurrently on the device, admitted in arrival order by
    free slots and a device-memory budget (SjhipServer* entry points)."""

    def __init__(self, index, refiner, slots=4, memory_budget=0, graphs=False):
        self._index, self._refiner = index, refiner
        cfg = L.SjhipServerConfig(int(slots), int(memory_budget), 1 if graphs else 0)
        h = ctypes.c_void_p()
        rc = L.lib().SjhipServerCreate(ctypes.byref(index.ix), ctypes.byref(refiner.rf), ctypes.byref(cfg),
                                       ctypes.byref(h))
        if rc:
            raise GpuSpatialError(L.lib().SjhipLastError().decode())
        self.h = h.value
        self._keep = {}

    def submit(self, probe, predicate, distance=0.0, precision=0, sort_output=False, sort_probes=True,
               refine_order=1):
        opts = _join_options(predicate, distance, precision, sort_output, sort_probes, False, refine_order, None)
        if isinstance(probe, DeviceWkb):
            keep, sch, dev = probe, probe.schema, probe.dev
        else:
            exp = ExportedArrow(probe)
            keep, sch, dev = exp, exp.schema, host_device_array(exp)
            keep = (exp, dev)
        rid = ctypes.c_uint64()
        rc = L.lib().SjhipServerSubmit(self.h, ctypes.byref(sch), ctypes.byref(dev), ctypes.byref(opts),
                                       ctypes.byref(rid))
        if rc:
            raise GpuSpatialError(L.lib().SjhipLastError().decode())
        self._keep[rid.value] = keep
        return rid.value

    def wait(self, rid, timeout_ms=-1):
        """(JoinResult, timing dict), or None while the request is not done."""
        res = L.SjhipJoinResult()
        t = L.SjhipRequestTiming()
        rc = L.lib().SjhipServerWait(self.h, int(rid), int(timeout_ms), ctypes.byref(res), ctypes.byref(t))
        if rc == errno.EAGAIN:
            return None
        self._keep.pop(int(rid), None)
        if rc:
            raise GpuSpatialError(L.lib().SjhipLastError().decode())
        timing = {"submit_ms": t.submit_ms, "admit_ms": t.admit_ms, "done_ms": t.done_ms,
                  "estimate_bytes": int(t.estimate_bytes)}
        return JoinResult(res), timing

    def close(self):
        if getattr(self, "h", None):
            L.lib().SjhipServerDestroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
