"""ctypes wrapper of the CPU oracle (sjoin_oracle.c) and of the reference's
own TG refine library built into oracle/_ref/ (tgref_driver.c + tg.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never imports
this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libtgref.so")

PRED = {"equals": 0, "disjoint": 1, "touches": 2, "contains": 3, "covers": 4,
        "intersects": 5, "within": 6, "coveredby": 7, "dwithin": 100}

_P = ctypes.c_void_p
_I64 = ctypes.c_int64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _load(path):
    lib = ctypes.CDLL(path)
    lib.ora_store_from_wkb.argtypes = [_P, _P, _P, _I64, ctypes.POINTER(_P), ctypes.POINTER(_I64)]
    lib.ora_store_free.argtypes = [_P]
    lib.ora_store_sizes.argtypes = [_P, _P]
    lib.ora_store_export.argtypes = [_P] * 6
    lib.ora_rects.argtypes = [_P, ctypes.c_int, _P]
    lib.ora_expand_rects.argtypes = [_P, _I64, ctypes.c_double]
    lib.ora_filter.argtypes = [_P, _I64, _P, _I64, _P, _P, _I64, ctypes.c_int]
    lib.ora_filter.restype = _I64
    lib.ora_refine.argtypes = [_P, _P, ctypes.c_int, ctypes.c_double, _P, _P, _I64, _P, ctypes.c_int]
    lib.ora_join.argtypes = [_P, _P, ctypes.c_int, ctypes.c_double, _P, _P, _I64, ctypes.c_int]
    lib.ora_join.restype = _I64
    lib.ora_distance.argtypes = [_P, _I64, _P, _I64]
    lib.ora_distance.restype = ctypes.c_double
    lib.ora_knn.argtypes = [_P, _P, ctypes.c_int, _P, _P, ctypes.c_int]
    lib.ora_knn_brute.argtypes = [_P, _P, ctypes.c_int, ctypes.c_double, _P, _P, _P, _I64, ctypes.c_int]
    lib.ora_knn_brute.restype = _I64
    lib.ora_pair_distances.argtypes = [_P, _P, _P, _P, _I64, _P, ctypes.c_int]
    d6 = [ctypes.c_double] * 6
    lib.ora_raycast.argtypes = d6
    lib.ora_collinear.argtypes = d6
    lib.ora_orient2d_sign.argtypes = d6
    lib.ora_segment_intersects.argtypes = [ctypes.c_double] * 8
    lib.ora_f64_interval_to_f32.argtypes = [ctypes.c_double, ctypes.c_double, _P, _P]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        _lib = _load(LIB_PATH)
    return _lib


def wkb_buffers(arr):
    """(int64 offsets[n+1], uint8 data, uint8 valid or None) of a pyarrow
    Binary/LargeBinary array, or passes through an (offsets, data, valid)
    tuple of numpy arrays."""
    if isinstance(arr, tuple):
        off, data, valid = arr
        return np.ascontiguousarray(off, dtype=np.int64), np.ascontiguousarray(data, dtype=np.uint8), valid
    import pyarrow as pa
    if isinstance(arr, pa.ChunkedArray):
        arr = arr.combine_chunks()
    if pa.types.is_binary_view(arr.type):
        arr = arr.cast(pa.large_binary())
    bufs = arr.buffers()
    n = len(arr)
    odt = np.int64 if pa.types.is_large_binary(arr.type) else np.int32
    off = np.frombuffer(bufs[1], dtype=odt)[arr.offset:arr.offset + n + 1].astype(np.int64)
    data = np.frombuffer(bufs[2], dtype=np.uint8) if bufs[2] is not None else np.zeros(1, np.uint8)
    valid = None
    if arr.null_count:
        valid = np.asarray(arr.is_valid(), dtype=np.uint8)
    return off, data, valid


class Store:
    """Decoded geometry column (row -> parts -> rings -> points)."""

    def __init__(self, arr):
        off, data, valid = wkb_buffers(arr)
        self._keep = (off, data, valid)
        self.n = len(off) - 1
        h = _P()
        bad = _I64(-1)
        rc = lib().ora_store_from_wkb(_ptr(off), _ptr(data), _ptr(valid), self.n, ctypes.byref(h), ctypes.byref(bad))
        if rc:
            raise ValueError(f"oracle WKB decode failed (rc={rc}) at row {bad.value}")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_store_free(self.h)
            self.h = None

    def soa(self):
        sz = np.zeros(4, np.int64)
        lib().ora_store_sizes(self.h, _ptr(sz))
        ng, npart, nring, npt = (int(x) for x in sz)
        t = np.zeros(ng, np.uint8)
        po = np.zeros(ng + 1, np.int64)
        ro = np.zeros(npart + 1, np.int64)
        vo = np.zeros(nring + 1, np.int64)
        xy = np.zeros(2 * npt, np.float64)
        lib().ora_store_export(self.h, _ptr(t), _ptr(po), _ptr(ro), _ptr(vo), _ptr(xy))
        return dict(type=t, part_off=po, ring_off=ro, pt_off=vo, xy=xy.reshape(-1, 2))

    def rects(self, point_nearest=True):
        out = np.zeros((self.n, 4), np.float32)
        lib().ora_rects(self.h, 1 if point_nearest else 0, _ptr(out))
        return out


def expand_rects(rects, distance):
    r = np.ascontiguousarray(rects, dtype=np.float32).copy()
    lib().ora_expand_rects(_ptr(r), len(r), float(distance))
    return r


def f64_interval_to_f32(lo, hi):
    a = np.zeros(1, np.float32)
    b = np.zeros(1, np.float32)
    lib().ora_f64_interval_to_f32(float(lo), float(hi), _ptr(a), _ptr(b))
    return float(a[0]), float(b[0])


def filter_pairs(brects, prects, nthreads=0):
    br = np.ascontiguousarray(brects, np.float32)
    pr = np.ascontiguousarray(prects, np.float32)
    n = lib().ora_filter(_ptr(br), len(br), _ptr(pr), len(pr), None, None, 0, nthreads)
    ob = np.zeros(max(n, 1), np.uint32)
    op = np.zeros(max(n, 1), np.uint32)
    lib().ora_filter(_ptr(br), len(br), _ptr(pr), len(pr), _ptr(ob), _ptr(op), n, nthreads)
    return ob[:n], op[:n]


def refine(build, probe, pred, bidx, pidx, distance=0.0, nthreads=0):
    bidx = np.ascontiguousarray(bidx, np.uint32)
    pidx = np.ascontiguousarray(pidx, np.uint32)
    mask = np.zeros(max(len(bidx), 1), np.uint8)
    rc = lib().ora_refine(build.h, probe.h, PRED[pred], float(distance), _ptr(bidx), _ptr(pidx),
                          len(bidx), _ptr(mask), nthreads)
    if rc:
        raise ValueError("oracle refine: unsupported geometry/predicate combination")
    return mask[:len(bidx)].astype(bool)


def join(build, probe, pred, distance=0.0, nthreads=0):
    """Refined (build_id, probe_id) pairs sorted by (probe, build)."""
    L = lib()
    n = L.ora_join(build.h, probe.h, PRED[pred], float(distance), None, None, 0, nthreads)
    if n < 0:
        raise ValueError("oracle join: unsupported geometry/predicate combination")
    ob = np.zeros(max(n, 1), np.uint32)
    op = np.zeros(max(n, 1), np.uint32)
    L.ora_join(build.h, probe.h, PRED[pred], float(distance), _ptr(ob), _ptr(op), n, nthreads)
    return ob[:n], op[:n]


def distance(build, b, probe, p):
    return lib().ora_distance(build.h, int(b), probe.h, int(p))


def knn(build, probe, k, nthreads=0):
    ids = np.zeros((probe.n, k), np.uint32)
    d = np.zeros((probe.n, k), np.float64)
    rc = lib().ora_knn(build.h, probe.h, int(k), _ptr(ids), _ptr(d), nthreads)
    if rc:
        raise ValueError("oracle knn: objects must be points")
    return ids, d


def raycast(ax, ay, bx, by, px, py):
    return lib().ora_raycast(ax, ay, bx, by, px, py)


def collinear(*a):
    return lib().ora_collinear(*[float(x) for x in a])


def orient2d_sign(*a):
    return lib().ora_orient2d_sign(*[float(x) for x in a])


def segment_intersects(*a):
    return lib().ora_segment_intersects(*[float(x) for x in a])


# ---------------------------------------------------------------------------
# The reference's own TG library (oracle/_ref), when built.

_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def reflib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference TG library not built: {REF_PATH} (run make -C oracle ref)")
        r = ctypes.CDLL(REF_PATH)
        r.tgref_parse.argtypes = [_P, _P, _P, _I64, ctypes.c_int, ctypes.c_int]
        r.tgref_parse.restype = _P
        r.tgref_free.argtypes = [_P]
        r.tgref_refine.argtypes = [_P, _P, ctypes.c_int, _P, _P, _I64, _P, ctypes.c_int]
        r.tgref_join.argtypes = [_P, _P, _P, _P, ctypes.c_int, _P, _P, _I64, ctypes.c_int]
        r.tgref_join.restype = _I64
        _ref = r
    return _ref


class TgColumn:
    """A WKB column parsed by the reference TG library (index: 0 none,
    1 natural, 2 ystripes)."""

    def __init__(self, arr, index=1, nthreads=0):
        off, data, valid = wkb_buffers(arr)
        self._keep = (off, data, valid)
        self.n = len(off) - 1
        self.h = reflib().tgref_parse(_ptr(off), _ptr(data), _ptr(valid), self.n, index, nthreads)

    def __del__(self):
        if getattr(self, "h", None):
            reflib().tgref_free(self.h)
            self.h = None


def ref_refine(build, probe, pred, bidx, pidx, nthreads=0):
    bidx = np.ascontiguousarray(bidx, np.uint32)
    pidx = np.ascontiguousarray(pidx, np.uint32)
    mask = np.zeros(max(len(bidx), 1), np.uint8)
    rc = reflib().tgref_refine(build.h, probe.h, PRED[pred], _ptr(bidx), _ptr(pidx), len(bidx),
                               _ptr(mask), nthreads)
    if rc:
        raise ValueError("unsupported predicate for TG")
    return mask[:len(bidx)].astype(bool)


def ref_join(build, probe, brects, prects, pred, nthreads=0, cap_hint=0):
    """Pairs sorted by (probe, build).  With cap_hint >= the pair count the
    join runs once; otherwise a counting run comes first."""
    br = np.ascontiguousarray(brects, np.float32)
    pr = np.ascontiguousarray(prects, np.float32)
    R = reflib()
    cap = int(cap_hint)
    if cap <= 0:
        cap = R.tgref_join(build.h, probe.h, _ptr(br), _ptr(pr), PRED[pred], None, None, 0, nthreads)
    ob = np.zeros(max(cap, 1), np.uint32)
    op = np.zeros(max(cap, 1), np.uint32)
    n = R.tgref_join(build.h, probe.h, _ptr(br), _ptr(pr), PRED[pred], _ptr(ob), _ptr(op), cap, nthreads)
    if n < 0:
        raise ValueError("unsupported predicate for TG")
    if n > cap:
        return ref_join(build, probe, brects, prects, pred, nthreads, cap_hint=n)
    return ob[:n], op[:n]


def knn_brute(build, probe, k, include_tie_breakers=False, nthreads=0):
    """KNN over any geometry (one side points) by brute force, rows of
    variable length: (offsets, ids, distances); with include_tie_breakers
    every object within 1e-9 of the k-th distance is kept as well."""
    tol = 1e-9 if include_tie_breakers else -1.0
    L = lib()
    off = np.zeros(probe.n + 1, np.int64)
    n = L.ora_knn_brute(build.h, probe.h, int(k), tol, _ptr(off), None, None, 0, nthreads)
    if n < 0:
        raise ValueError("knn: one side must be points")
    ids = np.zeros(max(n, 1), np.uint32)
    d = np.zeros(max(n, 1), np.float64)
    L.ora_knn_brute(build.h, probe.h, int(k), tol, _ptr(off), _ptr(ids), _ptr(d), n, nthreads)
    return off, ids[:n], d[:n]


def pair_distances(build, probe, bidx, pidx, nthreads=0):
    """ora_distance over (build, probe) pairs."""
    b = np.ascontiguousarray(bidx, np.uint32)
    p = np.ascontiguousarray(pidx, np.uint32)
    out = np.zeros(max(len(b), 1), np.float64)
    lib().ora_pair_distances(build.h, probe.h, _ptr(b), _ptr(p), len(b), _ptr(out), nthreads)
    return out[:len(b)]
