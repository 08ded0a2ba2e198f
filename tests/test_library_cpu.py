"""CPU-side checks of the HIP library: it loads, exports every symbol the
public headers declare, and its scalar geometry code (the same source the
kernels compile, evaluated on the host) agrees with the oracle and with
the reference host's libm."""
import ctypes
import os
import re

import numpy as np
import pytest

import corpus as C
import oracle as O
from helpers import ROOT
from sjhip import _lib as L


def _declared_symbols():
    names = []
    for h in ("gpuspatial_hip.h", "gpuspatial_hip_testing.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        # top-level function declarations (struct members are indented)
        for m in re.finditer(r"^(?:const\s+)?(?:struct\s+)?\w+\s*\**\s*(\w+)\s*\(", src, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    declared = _declared_symbols()
    assert "GpuSpatialRuntimeCreate" in declared and "GpuSpatialRefinerCreate" in declared
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(L.EXPORTED) == sorted(declared)


def test_abi_struct_layouts_match_reference_header():
    """Sizes / member offsets of the gpuspatial_c.h structs on x86-64."""
    assert ctypes.sizeof(L.GpuSpatialRuntimeConfig) == 24
    assert ctypes.sizeof(L.GpuSpatialRuntime) == 32
    assert ctypes.sizeof(L.SedonaFloatIndex2D) == 80
    assert L.SedonaFloatIndex2D.probe.offset == 40
    assert ctypes.sizeof(L.GpuSpatialRefinerConfig) == 24
    assert ctypes.sizeof(L.SedonaSpatialRefiner) == 64
    assert L.SedonaSpatialRefiner.refine.offset == 32
    assert ctypes.sizeof(L.ArrowDeviceArray) == 80 + 8 + 8 + 8 + 24


def test_runtime_init_fails_on_missing_shader_dir():
    """c_wrapper_test.cc:45-65: the error names 'No such file or directory'."""
    from sjhip.gpuspatial import Runtime
    rt = Runtime(ptx_root="/invalid/path/to/ptx")
    assert rt.init() != 0
    assert "No such file or directory" in rt.last_error()
    rt.release()


def test_host_raycast_matches_oracle_on_corpus():
    cb, cp = C.corpus()
    sb, sp = O.Store(cb).soa(), O.Store(cp).soa()
    pts = sp["xy"][sp["pt_off"][:-1]]
    segs = []
    for r in range(len(sb["pt_off"]) - 1):
        ring = sb["xy"][sb["pt_off"][r]:sb["pt_off"][r + 1]]
        for i in range(len(ring) - 1):
            segs.append(np.concatenate([ring[i], ring[i + 1]]))
    segs = np.array(segs)
    rng = np.random.default_rng(1)
    pick = rng.choice(len(pts), size=min(len(pts), 800), replace=False)
    s6 = np.concatenate([np.repeat(segs, len(pick), axis=0), np.tile(pts[pick], (len(segs), 1))], axis=1)
    s6 = np.ascontiguousarray(s6)
    got = np.zeros(len(s6), np.int32)
    L.lib().SjhipHostRaycastBatch(s6.ctypes.data, len(s6), got.ctypes.data)
    want = np.array([O.raycast(*row) for row in s6[::3]])
    assert np.array_equal(got[::3], want)
    assert set(np.unique(got)) == {0, 1, 2}


def test_host_hypot_is_bit_identical_to_glibc():
    """The reference host computes hypot with glibc (Rust f64::hypot); the
    library's hypot must agree bit for bit, including near-tie inputs."""
    rng = np.random.default_rng(2)
    n = 400_000
    xy = np.empty((n, 2))
    xy[:, 0] = (rng.random(n) - 0.5) * 2.0 ** rng.integers(-30, 30, n)
    xy[:, 1] = (rng.random(n) - 0.5) * 2.0 ** rng.integers(-30, 30, n)
    xy[::3, 1] = xy[::3, 0] * (1 + (rng.random(len(xy[::3])) - 0.5) * 1e-3)
    xy[1::7, 1] = 0.0
    xy = np.ascontiguousarray(xy)
    out = np.zeros(n)
    L.lib().SjhipHostHypotBatch(xy.ctypes.data, n, out.ctypes.data)
    want = np.hypot(xy[:, 0], xy[:, 1])
    assert np.array_equal(out.view(np.uint64), want.view(np.uint64))


def test_host_coord_to_line_matches_oracle():
    rng = np.random.default_rng(4)
    lib = L.lib()
    for _ in range(3000):
        v = rng.uniform(0, 10, 6)
        A = O.Store(C.column([C.linestring_wkb([(v[2], v[3]), (v[4], v[5])])]))
        B = O.Store(C.column([C.point_wkb(v[0], v[1])]))
        want = O.distance(A, 0, B, 0)
        got = lib.SjhipHostCoordToLine(*v)
        if want == 0.0:
            continue  # on-line exact zero handled before the segment formula
        assert got == want


def test_host_orientation_and_segment_tests_match_oracle():
    rng = np.random.default_rng(6)
    lib = L.lib()
    for _ in range(5000):
        v = np.round(rng.uniform(0, 4, 8) * 2) / 2  # grid: many degenerate cases
        assert lib.SjhipHostOrient2d(*v[:6]) == O.orient2d_sign(*v[:6])
        assert lib.SjhipHostCollinear(*v[:6]) == O.collinear(*v[:6])
        assert lib.SjhipHostSegmentIntersects(*v) == O.segment_intersects(*v)
    # near-collinear, where a plain FP64 cross product can get the sign wrong
    for _ in range(2000):
        a = rng.uniform(0, 1, 2)
        b = rng.uniform(0, 1, 2)
        t = rng.random()
        c = a + t * (b - a)
        c = c + rng.integers(-2, 3, 2) * np.spacing(c)
        assert lib.SjhipHostOrient2d(*a, *b, *c) == O.orient2d_sign(*a, *b, *c)


def test_fp32_certified_raycast_never_contradicts_fp64():
    """Whenever the FP32 form decides, it decides what the reference's FP64
    sequence decides (inputs: FP32-rounded local coordinates with the bound
    the library uses)."""
    rng = np.random.default_rng(8)
    lib = L.lib()
    decided = 0
    for _ in range(60000):
        o = rng.uniform(0, 1000, 2)
        H = 10.0 ** rng.uniform(-3, 1)
        a = o + rng.uniform(-H, H, 2)
        b = o + rng.uniform(-H, H, 2)
        if rng.random() < 0.5:  # points near the segment or at vertex heights
            t = rng.random()
            p = a + t * (b - a) + rng.normal(0, 1e-7 * H, 2)
        else:
            p = o + rng.uniform(-H, H, 2)
        if rng.random() < 0.1:
            p[1] = a[1]
        E = np.float32(H * 2.384185791015625e-07) * np.float32(1.0001)
        # the kernels also cast the ray toward -x (x mirrored), +y and -y
        # (x and y swapped, then mirrored): both maps are exact
        sx = 1.0 if rng.random() < 0.5 else -1.0
        swap = rng.random() < 0.5
        pts = [(a[0] - o[0], a[1] - o[1]), (b[0] - o[0], b[1] - o[1]), (p[0] - o[0], p[1] - o[1])]
        if swap:
            pts = [(y, x) for x, y in pts]
        pts = [(sx * x, y) for x, y in pts]
        loc = [np.float32(v) for q in pts for v in q]
        f = lib.SjhipHostRaycastF32(*[float(x) for x in loc], float(E))
        if f < 0:
            continue
        decided += 1
        if swap:
            want = lib.SjhipHostRaycast(sx * a[1], a[0], sx * b[1], b[0], sx * p[1], p[0])
        else:
            want = lib.SjhipHostRaycast(sx * a[0], a[1], sx * b[0], b[1], sx * p[0], p[1])
        assert f == want
    assert decided > 20000


def test_fp32_certified_distance_never_contradicts_fp64():
    """The certified FP32 DWithin term (ring-local FP32 vertices with the
    library's bound E = 2^-22 H and the point's own rounding) never decides
    otherwise than the reference's FP64 term (on the segment, or
    distance_coord_to_line <= d): points placed at distance d plus or minus
    tiny offsets, on the segment, past its ends; short and long segments;
    rings of extent 1e-3 .. 1e3 at coordinates up to 1e4."""
    rng = np.random.default_rng(17)
    lib = L.lib()
    decided = total = 0
    seen = {0: 0, 1: 0}
    for _ in range(40000):
        o = rng.uniform(-1e4, 1e4, 2)
        H = 10.0 ** rng.uniform(-3, 3)
        a = o + rng.uniform(-H, H, 2)
        b = a + rng.normal(0, H * 10.0 ** rng.uniform(-4, 0), 2)
        d = H * 10.0 ** rng.uniform(-3, 0.5) if rng.random() < 0.9 else 0.0
        t = rng.uniform(-0.3, 1.3)
        v = b - a
        nv = np.hypot(*v) or 1.0
        nrm = np.array([-v[1], v[0]]) / nv
        off = d * (1 + rng.choice([0.0, 1e-12, -1e-12, 1e-7, -1e-7, 1e-4, -1e-4, 0.3, -0.3]))
        if rng.random() < 0.1:
            off = 0.0  # on the segment's line
        p = a + t * v + off * nrm
        E = np.float32(H * 2.384185791015625e-07) * np.float32(1.0001)
        loc = [np.float32(x) for x in (a[0] - o[0], a[1] - o[1], b[0] - o[0], b[1] - o[1], p[0] - o[0], p[1] - o[1])]
        Ep = np.float32(max(abs(loc[4]), abs(loc[5]))) * np.float32(5.9604645e-08) * np.float32(1.0001)
        f = lib.SjhipHostSegmentWithinF32(*[float(x) for x in loc], float(E), float(Ep), float(d))
        total += 1
        if f < 0:
            continue
        decided += 1
        seen[f] += 1
        on = (lib.SjhipHostOrient2d(a[0], a[1], b[0], b[1], p[0], p[1]) == 0 and
              min(a[0], b[0]) <= p[0] <= max(a[0], b[0]) and min(a[1], b[1]) <= p[1] <= max(a[1], b[1]))
        want = on or lib.SjhipHostCoordToLine(p[0], p[1], a[0], a[1], b[0], b[1]) <= d
        assert bool(f) == want, (a, b, p, d, f)
    # most samples sit within 1e-7 of the threshold on purpose; those far
    # from it are decided both ways
    assert decided > 0.4 * total and min(seen.values()) > 5000, (decided, total, seen)


def test_fp32_certified_segment_crossing_never_contradicts_fp64():
    """The certified FP32 segment x segment test never decides otherwise
    than the reference's FP64 segment_intersects (tg.c:1392-1443), with
    ring-local FP32 coordinates and the bound the polygon x polygon kernels
    use (the ring's E = 2^-22 H plus the other segment's own rounding):
    crossing, touching (an endpoint on the other segment or one ulp off it),
    shared endpoints, collinear overlapping and disjoint, near-parallel and
    disjoint pairs, rings of extent 1e-3 .. 1e3 at coordinates up to 1e4."""
    rng = np.random.default_rng(23)
    lib = L.lib()
    seen = {0: 0, 1: 0, -1: 0}
    for _ in range(60000):
        o = rng.uniform(-1e4, 1e4, 2)
        H = 10.0 ** rng.uniform(-3, 3)
        a = o + rng.uniform(-H, H, 2)
        b = a + rng.normal(0, H * 10.0 ** rng.uniform(-3, 0), 2)
        kind = rng.integers(0, 6)
        if kind == 0:  # general position
            c = o + rng.uniform(-H, H, 2)
            d = c + rng.normal(0, H * 10.0 ** rng.uniform(-3, 0), 2)
        elif kind == 1:  # an endpoint on a-b, or a few ulps off it
            t = rng.random()
            c = a + t * (b - a) + rng.integers(-2, 3, 2) * np.spacing(np.abs(a) + 1.0)
            d = c + rng.normal(0, H * 0.3, 2)
        elif kind == 2:  # shared endpoint
            c = b.copy()
            d = c + rng.normal(0, H * 0.3, 2)
        elif kind == 3:  # collinear, overlapping or not
            t0, t1 = rng.uniform(-0.5, 1.5, 2)
            c, d = a + t0 * (b - a), a + t1 * (b - a)
        elif kind == 4:  # near-parallel, offset by a tiny distance
            v = b - a
            nrm = np.array([-v[1], v[0]]) / (np.hypot(*v) or 1.0)
            off = H * 10.0 ** rng.uniform(-9, -3) * rng.choice([-1.0, 1.0])
            c = a + rng.uniform(-0.5, 0.5) * v + off * nrm
            d = b + rng.uniform(-0.5, 0.5) * v + off * nrm * rng.uniform(0.5, 1.5)
        else:  # crossing near an endpoint of c-d
            t = rng.random()
            m = a + t * (b - a)
            dirn = rng.normal(0, 1, 2)
            c = m + dirn * H * 10.0 ** rng.uniform(-8, -1)
            d = m - dirn * H * rng.uniform(0.1, 1.0)
        E = np.float32(H * 2.384185791015625e-07) * np.float32(1.0001)
        loc = [np.float32(x) for x in (a[0] - o[0], a[1] - o[1], b[0] - o[0], b[1] - o[1])]
        q = [np.float32(x) for x in (c[0] - o[0], c[1] - o[1], d[0] - o[0], d[1] - o[1])]
        Ep = np.float32(max(abs(v) for v in q)) * np.float32(5.9604645e-08) * np.float32(1.0001)
        # the kernels pass the other ring's segment first, the ring's second
        f = lib.SjhipHostSegmentsCrossF32(*[float(x) for x in q], *[float(x) for x in loc], float(E + Ep))
        seen[f] += 1
        if f < 0:
            continue
        want = lib.SjhipHostSegmentIntersects(c[0], c[1], d[0], d[1], a[0], a[1], b[0], b[1])
        assert bool(f) == bool(want), (a, b, c, d, f, kind)
    assert seen[0] > 10000 and seen[1] > 5000 and seen[-1] > 1000, seen

