"""The build-rect (polygon-major) pass of the fused point-in-polygon join.

Filter mode 0 (the default) lets every build rect test the points of the
Morton cells it covers, with the cell-raster split decided in the same
pass; mode 3 takes the probe-major filter, the raster pass and the
regrouping sort.  Both must give the oracle's pair set and the same
candidate count, for every predicate of the point-in-polygon refine and
both precisions, on inputs that reach the pass's edge cases: nulls and
empty points, points on the FP32 rect bounds (cell boundaries), a probe
set of zero width (a degenerate Morton frame), a rect over more cells than
the pass takes (the probe-major fallback), and rows with holes and
multipolygons (the general kernel's rows)."""
import sys

import numpy as np
import pytest

import corpus as C
import oracle as O
from sjhip import datagen as D
from sjhip import gpuspatial as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def opts():
    return G.GpuSpatialOptions(concurrency=1, device_id=0)


def _pair(build, opts):
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(build)
    rf.push_build(build)
    rf.finish_building()
    ix = G.GpuSpatialIndex(opts)
    ix.build_from_refiner(rf)
    return ix, rf


def _points(x, y, valid=None):
    c = D.points_wkb(np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64))
    return D.WkbColumn(c.offsets, c.data, valid)


def _zipf_nulls():
    b, p = D.workload_pip_zipf(200_000, 2_000, seed=21)
    xy = p.data.reshape(-1, 21)[:, 5:].copy().view(np.float64).reshape(-1, 2)
    x, y = xy[:, 0].copy(), xy[:, 1].copy()
    x[::97] = np.nan
    y[::97] = np.nan
    valid = np.ones(len(x), bool)
    valid[5::101] = False
    return b.as_tuple(), _points(x, y, valid)


def _rect_edges():
    """Points on the FP32 rects of stars: their bounds, one ULP inside and
    outside, and midpoints, in every combination, plus random points."""
    b, _ = D.workload_pip_zipf(10, 400, seed=22)
    r = G.compute_rects(b.to_arrow())

    def around(v):
        return [v, np.nextafter(v, np.float32(-np.inf)), np.nextafter(v, np.float32(np.inf))]

    xv = around(r[:, 0]) + around(r[:, 2]) + [0.5 * (r[:, 0] + r[:, 2])]
    yv = around(r[:, 1]) + around(r[:, 3]) + [0.5 * (r[:, 1] + r[:, 3])]
    xs = [np.asarray(xa, np.float64) for xa in xv for _ in yv]
    ys = [np.asarray(ya, np.float64) for _ in xv for ya in yv]
    rng = np.random.default_rng(23)
    lo, hi = r[:, :2].min(axis=0), r[:, 2:].max(axis=0)
    xs.append(rng.uniform(lo[0], hi[0], 50_000))
    ys.append(rng.uniform(lo[1], hi[1], 50_000))
    return b.as_tuple(), _points(np.concatenate(xs), np.concatenate(ys))


def _zero_width():
    """Every probe on one vertical line (x extent 0: the frame's x scale is 0)."""
    b, _ = D.workload_pip_zipf(10, 300, seed=24)
    r = G.compute_rects(b.to_arrow()).astype(np.float64)
    x0 = float(np.median(0.5 * (r[:, 0] + r[:, 2])))
    y = np.random.default_rng(25).uniform(r[:, 1].min(), r[:, 3].max(), 40_000)
    return b.as_tuple(), _points(np.full_like(y, x0), y)


def _huge_rect():
    """Small stars plus one star over the whole extent: its rect covers more
    cells than the pass takes, so the join falls back to the probe-major
    filter."""
    b, p = D.workload_pip_zipf(60_000, 600, seed=26)
    rng = np.random.default_rng(27)
    ext = 1000.0 * np.sqrt(600 / 1_000_000)
    ang = np.sort(rng.uniform(0, 2 * np.pi, 64))
    rad = ext * (0.45 + 0.05 * rng.random(64))
    ring = np.stack([ext / 2 + rad * np.cos(ang), ext / 2 + rad * np.sin(ang)], axis=1)
    ring = np.vstack([ring, ring[:1]])
    big = D.rings_wkb([len(ring)], ring, 3)
    off = np.concatenate([b.offsets, b.offsets[-1] + big.offsets[1:]])
    return (off, np.concatenate([b.data, big.data]), None), p


def _corpus():
    cb, cp = C.corpus()
    hb = C.column(C.hole_polygons())
    off = np.concatenate([cb[0], cb[0][-1] + hb[0][1:]])
    return (off, np.concatenate([cb[1], hb[1]]), None), D.WkbColumn(cp[0], cp[1])


CASES = {"zipf_nulls": _zipf_nulls, "rect_edges": _rect_edges, "zero_width": _zero_width, "huge_rect": _huge_rect,
         "corpus_holes": _corpus}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", [0, 1])
def test_build_rect_pass_equals_probe_major_and_oracle(case, precision, opts):
    bt, p = CASES[case]()
    build = D.WkbColumn(*bt)
    ix, rf = _pair(build.to_arrow(), opts)
    dev = p.to_arrow()  # (an Arrow column keeps the nulls)
    for pred in ["contains", "intersects", "covers"]:
        say("build-rect pass", case, "precision", precision, pred, "rows", len(build), "probes", len(p))
        ix.set_filter_mode(0)
        r0 = G.join_device(ix, rf, dev, pred, precision=precision)
        ix.set_filter_mode(3)
        r3 = G.join_device(ix, rf, dev, pred, precision=precision)
        ix.set_filter_mode(0)
        b0, p0 = r0.to_host()
        b3, p3 = r3.to_host()
        ob, op = O.join(O.Store(bt), O.Store(p.as_tuple()), pred, nthreads=8)
        assert r0.candidates == r3.candidates, (r0.candidates, r3.candidates)
        assert np.array_equal(b0, ob) and np.array_equal(p0, op), (len(b0), len(ob))
        assert np.array_equal(b3, ob) and np.array_equal(p3, op), (len(b3), len(ob))
        if case != "zero_width":
            assert len(ob) > 0


@pytest.mark.parametrize("mode", [0, 3])
def test_every_candidate_decided_by_the_rasters(mode, opts):
    """Probes only in clean inside cells of a 200-vertex star (it has a
    segment tree and a cell raster): the raster split decides every
    candidate and leaves the walk none, and the kept pairs are the join's."""
    ang = np.linspace(0, 2 * np.pi, 200, endpoint=False)
    rad = 10.0 * (0.9 + 0.1 * np.cos(7 * ang))
    ring = np.stack([50 + rad * np.cos(ang), 50 + rad * np.sin(ang)], axis=1)
    ring = np.vstack([ring, ring[:1]])
    build = D.rings_wkb([len(ring)], ring, 3)
    rng = np.random.default_rng(29)
    p = _points(50 + rng.uniform(-2, 2, 500), 50 + rng.uniform(-2, 2, 500))
    ix, rf = _pair(build.to_arrow(), opts)
    ix.set_filter_mode(mode)
    for pred in ["contains", "intersects"]:
        r = G.join_device(ix, rf, p.to_arrow(), pred, precision=0)
        gb, gp = r.to_host()
        ob, op = O.join(O.Store(build.as_tuple()), O.Store(p.as_tuple()), pred, nthreads=8)
        say("raster-decided", mode, pred, len(gb), "walked", r.res.refine_candidates)
        assert len(ob) == 500
        assert r.res.refine_candidates == 0
        assert np.array_equal(gb, ob) and np.array_equal(gp, op)


def _walk_probes(d):
    """Random-walk lines; probes: points at the lines' FP32 rect bounds
    widened by d (and one FP64 ulp either side), nulls, empty points and
    uniform points."""
    b, p = D.workload_dwithin_walks(10_000, 1_000, seed=33)
    assert b.data[0] == 1 and int.from_bytes(bytes(b.data[1:5]), "little") == 2  # little-endian linestrings
    r = G.compute_rects(b.to_arrow()).astype(np.float64)
    rng = np.random.default_rng(34)
    xs, ys = [], []
    for dx in (-np.inf, 0.0, np.inf):
        for side_x, side_y in ((r[:, 0] - d, 0.5 * (r[:, 1] + r[:, 3])), (r[:, 2] + d, 0.5 * (r[:, 1] + r[:, 3])),
                               (0.5 * (r[:, 0] + r[:, 2]), r[:, 1] - d), (0.5 * (r[:, 0] + r[:, 2]), r[:, 3] + d)):
            xs.append(side_x if dx == 0.0 else np.nextafter(side_x, dx))
            ys.append(side_y)
    xy = p.data.reshape(-1, 21)[:, 5:].copy().view(np.float64).reshape(-1, 2)
    xs.append(xy[:, 0])
    ys.append(xy[:, 1])
    # the vertices of the first lines: at distance 0 from them
    for i in range(200):
        v = b.data[b.offsets[i] + 9:b.offsets[i + 1]].copy().view(np.float64).reshape(-1, 2)
        xs.append(v[:, 0])
        ys.append(v[:, 1])
    x, y = np.concatenate(xs), np.concatenate(ys)
    x[::89] = np.nan
    y[::89] = np.nan
    valid = np.ones(len(x), bool)
    valid[3::97] = False
    perm = rng.permutation(len(x))
    return b, _points(x[perm], y[perm], valid)


@pytest.mark.parametrize("d", [0.0, 0.05, 0.7])
def test_dwithin_build_rect_pass_equals_probe_major_and_oracle(d, opts):
    """ST_DWithin over single lines with line rasters: the build-rect pass
    (each line's rect widened by d visits the Morton cells of the points,
    the line rasters' drops decided in the pass) against the tree filter,
    the raster split and the grouping sort, and against the oracle."""
    b, p = _walk_probes(d)
    ix, rf = _pair(b.to_arrow(), opts)
    say("dwithin build-rect pass", d, "lines", len(b), "probes", len(p))
    ix.set_filter_mode(0)
    r0 = G.join_device(ix, rf, p.to_arrow(), "dwithin", distance=d)
    ix.set_filter_mode(3)
    r3 = G.join_device(ix, rf, p.to_arrow(), "dwithin", distance=d)
    ix.set_filter_mode(0)
    b0, p0 = r0.to_host()
    b3, p3 = r3.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "dwithin", distance=d, nthreads=8)
    assert r0.candidates == r3.candidates, (r0.candidates, r3.candidates)
    assert np.array_equal(b0, ob) and np.array_equal(p0, op), (len(b0), len(ob))
    assert np.array_equal(b3, ob) and np.array_equal(p3, op), (len(b3), len(ob))
    assert len(ob) > 1000


@pytest.mark.parametrize("d", [0.0, 0.05, 0.7])
def test_dwithin_polygons_build_rect_pass_equals_probe_major_and_oracle(d, opts):
    """ST_DWithin over polygons (Zipf stars with cell rasters, plus rows with
    holes and multipolygons): the build-rect pass keeps the points in clean
    inside cells (distance 0) and walks the rest; against the tree filter in
    filter order and against the oracle."""
    b, p = D.workload_pip_zipf(60_000, 600, seed=35)
    hb = C.column(C.hole_polygons())
    off = np.concatenate([b.offsets, b.offsets[-1] + hb[0][1:]])
    build = D.WkbColumn(off, np.concatenate([b.data, hb[1]]))
    rng = np.random.default_rng(36)
    xy = p.data.reshape(-1, 21)[:, 5:].copy().view(np.float64).reshape(-1, 2)
    x = np.concatenate([xy[:, 0], rng.uniform(-1, 17, 20_000)])
    y = np.concatenate([xy[:, 1], rng.uniform(-1, 17, 20_000)])
    x[::83] = np.nan
    y[::83] = np.nan
    valid = np.ones(len(x), bool)
    valid[7::101] = False
    pts = _points(x, y, valid)
    ix, rf = _pair(build.to_arrow(), opts)
    say("dwithin polygons build-rect pass", d, "rows", len(build), "probes", len(pts))
    ix.set_filter_mode(0)
    r0 = G.join_device(ix, rf, pts.to_arrow(), "dwithin", distance=d)
    ix.set_filter_mode(3)
    r3 = G.join_device(ix, rf, pts.to_arrow(), "dwithin", distance=d, refine_order=0)
    ix.set_filter_mode(0)
    b0, p0 = r0.to_host()
    b3, p3 = r3.to_host()
    ob, op = O.join(O.Store(build.as_tuple()), O.Store(pts.as_tuple()), "dwithin", distance=d, nthreads=8)
    assert r0.candidates == r3.candidates, (r0.candidates, r3.candidates)
    assert np.array_equal(b0, ob) and np.array_equal(p0, op), (len(b0), len(ob))
    assert np.array_equal(b3, ob) and np.array_equal(p3, op), (len(b3), len(ob))
    assert len(ob) > 1000
