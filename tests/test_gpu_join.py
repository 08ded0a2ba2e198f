"""Parity tests of the HIP library on the GPU, through the C ABI.

Each test compares the library's output with the golden vectors minted by
the reference's TG library, or with the oracle on the same seeded inputs,
bit for bit (pair sets, neighbour ids, distances)."""
import sys

import numpy as np
import pytest

import corpus as C
import oracle as O
from touch_cases import touch_cases
from helpers import expected, pairs_equal, read_fixture
from sjhip import datagen as D
from sjhip import gpuspatial as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def opts():
    return G.GpuSpatialOptions(concurrency=1, device_id=0)


def build_pair(build_col, opts, from_refiner=True, point_nearest=True):
    """Index + refiner built from the same build rows (the operator's
    finish(), rust/sedona-spatial-join-gpu/src/index/gpu_spatial_index_builder.rs:150-259)."""
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(build_col)
    rf.push_build(build_col)
    rf.finish_building()
    ix = G.GpuSpatialIndex(opts)
    if from_refiner:
        ix.build_from_refiner(rf)
    else:
        ix.push_build(G.compute_rects(build_col, point_nearest=point_nearest))
        ix.finish_building()
    return ix, rf


def arrow(col):
    return col.to_arrow() if isinstance(col, D.WkbColumn) else col


def corpus_arrow(cols):
    import pyarrow as pa
    off, data, _ = cols
    return pa.Array.from_buffers(pa.binary(), len(off) - 1,
                                 [None, pa.py_buffer(off.astype(np.int32)), pa.py_buffer(data)])


FIXTURES = [
    ("pip", "synthetic_pip_polygons.parquet", "synthetic_pip_points.parquet",
     ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]),
    ("countries", "countries.parquet", "countries_points.parquet",
     ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]),
    ("points_in_countries_rev", "countries_points.parquet", "countries.parquet",
     ["within", "intersects", "coveredby", "touches"]),
    ("polypoly", "synthetic_pip_polygons.parquet", "synthetic_poly_polygons2.parquet",
     ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]),
    ("countries_countries", "countries.parquet", "countries.parquet",
     ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]),
]


@pytest.mark.parametrize("name,bf,pf,preds", FIXTURES)
def test_c_abi_join_matches_reference_goldens(name, bf, pf, preds, opts):
    """The flow of the reference's c_wrapper_test.cc:151-329 (index push /
    finish / probe, refiner push / finish / refine), checked against the
    pair sets of the reference's TG refine."""
    say("case", name)
    b, p = read_fixture(bf), read_fixture(pf)
    ix, rf = build_pair(b, opts, from_refiner=False)
    cb, cp = ix.probe(G.compute_rects(p, point_nearest=True))
    E = expected()
    for pred in preds:
        say("  predicate", pred, "candidates", len(cb))
        ob, op = rf.refine(p, pred, cb, cp)
        want = (E[f"{name}/{pred}/build"], E[f"{name}/{pred}/probe"])
        eq, miss, extra = pairs_equal((ob, op), want)
        assert eq, (name, pred, miss, extra)
        # sorted by (probe, build) like the reference's refiner output
        key = (op.astype(np.uint64) << 32) | ob
        assert np.all(np.diff(key.astype(np.int64)) > 0) or len(key) < 2


@pytest.mark.parametrize("pred", ["touches", "equals", "intersects", "covers"])
def test_touches_polygons_multipolygons_both_roles_vs_oracle(pred, opts):
    """tg_geom_touches: a polygon's boundary (holes included); a multipolygon
    (as the first argument) only when no other part holds the point inside;
    a point (as the first argument) touching any part."""
    bcol, pcol = touch_cases()
    B, P = O.Store(bcol), O.Store(pcol)
    bi = np.repeat(np.arange(B.n), P.n).astype(np.uint32)
    pi = np.tile(np.arange(P.n), B.n).astype(np.uint32)
    for roles in ("poly_build", "point_build"):
        if roles == "poly_build":
            rf = G.GpuSpatialRefiner(opts)
            rf.init_build_schema(corpus_arrow(bcol))
            rf.push_build(corpus_arrow(bcol))
            rf.finish_building()
            ob, op = rf.refine(corpus_arrow(pcol), pred, bi, pi)
            want = O.refine(B, P, pred, bi, pi)
            exp = (bi[want], pi[want])
        else:
            rf = G.GpuSpatialRefiner(opts)
            rf.init_build_schema(corpus_arrow(pcol))
            rf.push_build(corpus_arrow(pcol))
            rf.finish_building()
            ob, op = rf.refine(corpus_arrow(bcol), pred, pi, bi)
            want = O.refine(P, B, pred, pi, bi)
            exp = (pi[want], bi[want])
        say(pred, roles, int(want.sum()))
        eq, miss, extra = pairs_equal((ob, op), exp)
        assert eq, (pred, roles, miss, extra)
        if pred == "touches":
            assert want.sum() >= 8


def test_filter_matches_oracle_candidates(opts):
    b = read_fixture("synthetic_pip_polygons.parquet")
    p = read_fixture("synthetic_pip_points.parquet")
    br, pr = G.compute_rects(b), G.compute_rects(p)
    ob, op = O.filter_pairs(br, pr)
    ix = G.GpuSpatialIndex(opts)
    ix.push_build(br)
    ix.finish_building()
    gb, gp = ix.probe(pr)
    eq, miss, extra = pairs_equal((gb, gp), (ob, op))
    assert eq, (miss, extra)


def brute_force_candidates(br, qr):
    """(build, query) pairs whose FP32 rects meet (inclusive), valid rects only."""
    okb = (br[:, 0] <= br[:, 2]) & (br[:, 1] <= br[:, 3])
    okq = (qr[:, 0] <= qr[:, 2]) & (qr[:, 1] <= qr[:, 3])
    bs, ps = [], []
    for a in range(0, len(qr), 2048):
        q = qr[a:a + 2048]
        h = ((br[None, :, 0] <= q[:, None, 2]) & (br[None, :, 2] >= q[:, None, 0]) &
             (br[None, :, 1] <= q[:, None, 3]) & (br[None, :, 3] >= q[:, None, 1]) &
             okb[None, :] & okq[a:a + 2048, None])
        qi, bi = np.nonzero(h)
        bs.append(bi.astype(np.uint32))
        ps.append((qi + a).astype(np.uint32))
    return np.concatenate(bs), np.concatenate(ps)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("layout", ["dense", "sparse", "mixed", "boxes"])
def test_filter_both_walk_modes_vs_brute_force(layout, mode, opts):
    """The tree filter walks a wavefront of Morton-ordered queries once with
    the union of their rects when they lie close together (dense), and one
    lane per query when they lie far apart (sparse); the grid filter visits
    the cells of each query (mode 0: grid for point queries, 1: tree only,
    2: grid for every query, boxes spanning cells included); every form must
    give exactly the rect hits, edges included, each pair once."""
    rng = np.random.default_rng(17)
    nb = 3000
    c = rng.uniform(0, 100, size=(nb, 2)).astype(np.float32)
    h = rng.uniform(0.05, 1.5, size=(nb, 2)).astype(np.float32)
    br = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    if layout == "dense":
        xy = rng.uniform(40, 42, size=(30000, 2)).astype(np.float32)
    elif layout == "sparse":
        xy = rng.uniform(0, 100, size=(3000, 2)).astype(np.float32)
    else:
        xy = np.concatenate([rng.uniform(40, 42, size=(20000, 2)), rng.uniform(0, 100, size=(3000, 2))]).astype(
            np.float32)
    # points on build rect edges and corners (the compare is inclusive)
    e = rng.integers(0, nb, size=500)
    xy = np.concatenate([xy, np.stack([br[e, 0], br[e, 1]], 1), np.stack([br[e, 2], rng.uniform(br[e, 1], br[e, 3])], 1)
                         ]).astype(np.float32)
    qr = np.concatenate([xy, xy], axis=1)
    if layout == "boxes":
        qr[:, 2:] += rng.uniform(0, 0.3, size=(len(qr), 2)).astype(np.float32)
    qr[::97] = np.nan  # null rows
    qr[5::101, 0] = qr[5::101, 2] + 1  # inverted (empty) rects
    ix = G.GpuSpatialIndex(opts)
    ix.push_build(br)
    ix.finish_building()
    ix.set_filter_mode(mode)
    gb, gp = ix.probe(qr)
    ob, op = brute_force_candidates(br, qr)
    say(layout, mode, len(qr), len(ob))
    eq, miss, extra = pairs_equal((gb, gp), (ob, op))
    assert eq, (layout, miss, extra)


def test_device_rects_match_reference_rounding():
    """bounds.rs:147-162 outward rounding and the GPU operator's nearest
    rounding for points, as restated by the oracle."""
    for f in ["countries.parquet", "synthetic_pip_polygons.parquet", "countries_points.parquet"]:
        col = read_fixture(f)
        st = O.Store(col)
        for nearest in (True, False):
            got = G.compute_rects(col, point_nearest=nearest)
            want = st.rects(point_nearest=nearest)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (f, nearest)
    pts = read_fixture("countries_points.parquet")
    got = G.compute_rects(pts, point_nearest=False, expand_distance=0.05)
    want = O.expand_rects(O.Store(pts).rects(point_nearest=False), 0.05)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("precision", [0, 1])
def test_corpus_all_pairs_match_reference(precision, opts):
    """Every (polygon, point) pair of the adversarial corpus, both roles."""
    cb, cp = C.corpus()
    b, p = corpus_arrow(cb), corpus_arrow(cp)
    nb, np_ = len(b), len(p)
    bi = np.repeat(np.arange(nb), np_).astype(np.uint32)
    pi = np.tile(np.arange(np_), nb).astype(np.uint32)
    E = expected()
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(b)
    rf.push_build(b)
    rf.finish_building()
    rf.set_precision(precision)
    for pred in ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]:
        say("corpus", pred, "precision", precision)
        want = np.unpackbits(E[f"corpus/{pred}/mask"])[: len(bi)].astype(bool)
        ob, op = rf.refine(p, pred, bi, pi)
        eq, miss, extra = pairs_equal((ob, op), (bi[want], pi[want]))
        assert eq, (pred, miss, extra)
    rrf = G.GpuSpatialRefiner(opts)
    rrf.init_build_schema(p)
    rrf.push_build(p)
    rrf.finish_building()
    rrf.set_precision(precision)
    for pred in ["intersects", "within", "coveredby", "touches", "equals"]:
        want = np.unpackbits(E[f"corpus_rev/{pred}/mask"])[: len(bi)].astype(bool)
        ob, op = rrf.refine(b, pred, pi, bi)
        eq, miss, extra = pairs_equal((ob, op), (pi[want], bi[want]))
        assert eq, ("rev", pred, miss, extra)
    bb = np.repeat(np.arange(nb), nb).astype(np.uint32)
    pp = np.tile(np.arange(nb), nb).astype(np.uint32)
    want = np.unpackbits(E["corpus_polypoly/intersects/mask"])[: len(bb)].astype(bool)
    ob, op = rf.refine(b, "intersects", bb, pp)
    assert pairs_equal((ob, op), (bb[want], pp[want]))[0]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("sort_probes", [True, False])
def test_device_join_uniform_stars_golden(precision, sort_probes, opts):
    b, p = D.workload_pip_uniform(20_000, 200, seed=7, extent=100.0 * np.sqrt(200 / 10_000))
    ix, rf = build_pair(b.to_arrow(), opts)
    r = G.join_device(ix, rf, p.to_arrow(), "intersects", precision=precision, sort_probes=sort_probes)
    gb, gp = r.to_host()
    E = expected()
    assert np.array_equal(gb, E["uniform_small/intersects/build"])
    assert np.array_equal(gp, E["uniform_small/intersects/probe"])


@pytest.mark.parametrize("precision", [0, 1])
def test_c1_full_size_vs_oracle(precision, opts):
    """The oracle's own configuration at its full size: 1M uniform points x
    10k stars (8..32 vertices, radius lognormal with mean 0.5), ST_Intersects,
    through the fused device join and through the reference ABI flow
    (index probe, then refine of the host candidates)."""
    b, p = D.workload_pip_uniform(1_000_000, 10_000, seed=1)
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "intersects", nthreads=8)
    say("C1 full size, precision", precision, "oracle pairs", len(ob))
    assert 600_000 < len(ob) < 700_000
    ix, rf = build_pair(b.to_arrow(), opts)
    rf.set_precision(precision)
    r = G.join_device(ix, rf, G.DeviceWkb.from_host(p), "intersects", precision=precision)
    gb, gp = r.to_host()
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (len(gb), len(ob))
    cb, cp = ix.probe(G.compute_rects(p.to_arrow(), point_nearest=True))
    hb, hp = rf.refine(p.to_arrow(), "intersects", cb, cp)
    assert np.array_equal(hb, ob) and np.array_equal(hp, op), (len(hb), len(ob))


@pytest.mark.parametrize("pred", ["contains", "intersects", "covers"])
@pytest.mark.parametrize("precision", [0, 1])
def test_device_join_zipf_stars_vs_oracle(pred, precision, opts):
    """The 100M x 1M workload's distributions (Zipf vertex counts up to 4096,
    clustered points) at 1/1000 scale, device-resident probe batch."""
    b, p = D.workload_pip_zipf(100_000, 1_000, seed=11)
    say("zipf", pred, precision, "vertices", int(b.offsets[-1]))
    ix, rf = build_pair(b.to_arrow(), opts)
    dev = G.DeviceWkb.from_host(p)
    r = G.join_device(ix, rf, dev, pred, precision=precision)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), pred, nthreads=8)
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (len(gb), len(ob))
    if precision == 0:
        assert r.segment_tests > 0
        say("  fallback rate", r.fallback_tests / max(1, r.segment_tests))


def _raster_probe_build():
    """Single-ring polygons with trees (and so cell rasters): stars of 40 to
    3000 vertices and a staircase whose edges lie on the raster's cell lines
    (rect [0, 32]^2, cell width 1).  Probes: the lattice of cell corners and
    cell centres (one cell beyond the rect on every side), every vertex and
    edge midpoint, and random points in each rect."""
    rng = np.random.default_rng(31)
    rings = []
    for k, nv in enumerate([40, 200, 1000, 3000]):
        ang = np.sort(rng.random(nv)) * 2 * np.pi
        r = 10.0 * (0.6 + 0.4 * rng.random(nv))
        cx, cy = 100.0 + 40.0 * k, 50.0
        ring = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], axis=1)
        rings.append(np.vstack([ring, ring[:1]]))
    st = [(32.0, 0.0)]
    for j in range(32):
        st += [(32.0 - j, j + 1.0), (31.0 - j, j + 1.0)]
    st += [(0.0, 0.0), (32.0, 0.0)]
    rings.append(np.array(st))
    counts = [len(r) for r in rings]
    b = D.rings_wkb(counts, np.vstack(rings), 3)
    pts = []
    for ring in rings:
        x0, y0 = ring.min(axis=0)
        x1, y1 = ring.max(axis=0)
        cw, ch = (x1 - x0) / 32, (y1 - y0) / 32
        for t in (np.arange(-1, 34, dtype=np.float64), np.arange(-1, 33, dtype=np.float64) + 0.5):
            gx, gy = np.meshgrid(x0 + t * cw, y0 + t * ch)
            pts.append(np.stack([gx.ravel(), gy.ravel()], axis=1))
        pts.append(ring[:-1])
        pts.append(0.5 * (ring[:-1] + ring[1:]))
        pts.append(np.stack([x0 + (x1 - x0) * rng.random(2000), y0 + (y1 - y0) * rng.random(2000)], axis=1))
    xy = np.vstack(pts)
    return b, D.points_wkb(xy[:, 0].copy(), xy[:, 1].copy())


@pytest.mark.parametrize("pred", ["contains", "intersects"])
@pytest.mark.parametrize("sort_probes", [True, False])
def test_cell_raster_boundaries_vs_oracle(pred, sort_probes, opts):
    """Points on the cell lines, in the cell centres, on vertices and edges
    of polygons with cell rasters: the certified path (raster pass, walk,
    FP64 redo) gives the oracle's pairs."""
    b, p = _raster_probe_build()
    say("raster boundaries", pred, sort_probes, "probes", len(p.offsets) - 1)
    ix, rf = build_pair(b.to_arrow(), opts)
    r = G.join_device(ix, rf, p.to_arrow(), pred, precision=0, sort_probes=sort_probes)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), pred, nthreads=8)
    assert len(ob) > 10_000
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (len(gb), len(ob))


@pytest.mark.parametrize("refine_order,d", [(0, 0.05), (0, 0.0), (0, 3.0), (1, 0.05), (2, 0.05)])
def test_dwithin_random_walks_vs_oracle(refine_order, d, opts):
    # ~160 candidate boxes per point at this line density (as at full size);
    # d = 3 spans more raster rows than the line raster pass reads (walked)
    b, p = D.workload_dwithin_walks(10_000, 2_000, seed=13)
    say("dwithin walks", len(b), len(p), refine_order, d)
    ix, rf = build_pair(b.to_arrow(), opts)
    r = G.join_device(ix, rf, p.to_arrow(), "dwithin", distance=d, refine_order=refine_order)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "dwithin", distance=d, nthreads=8)
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (len(gb), len(ob))
    if d > 0:
        assert len(gb) > 1000


def _line_raster_probe_build():
    """Lines with cell rasters and edge cases of the raster: a staircase on
    the cell lines of its rect ([0, 64]^2, cell width 1), a line with
    repeated vertices (zero-length segments), a two-vertex diagonal, a
    3000-vertex random walk, and a horizontal line (zero height: no raster).
    Probes: cell corners and centres (one cell beyond the rect), vertices,
    edge midpoints, points at distance 0.01 from the vertices, random points."""
    rng = np.random.default_rng(41)
    lines = []
    st = [(0.0, 0.0)]
    for j in range(32):
        st += [(2.0 * j + 2.0, 2.0 * j), (2.0 * j + 2.0, 2.0 * j + 2.0)]
    lines.append(np.array(st) + [100.0, 0.0])
    rep = np.cumsum(rng.normal(0, 1.0, size=(40, 2)), axis=0) + [200.0, 50.0]
    lines.append(np.repeat(rep, 2, axis=0))
    lines.append(np.array([[300.0, 10.0], [330.0, 47.0]]))
    lines.append(np.cumsum(rng.normal(0, 0.5, size=(3000, 2)), axis=0) + [400.0, 50.0])
    lines.append(np.array([[500.0, 20.0], [510.0, 20.0], [520.0, 20.0]]))
    counts = [len(l) for l in lines]
    b = D.rings_wkb(counts, np.vstack(lines), 2)
    pts = []
    for ln in lines:
        x0, y0 = ln.min(axis=0)
        x1, y1 = ln.max(axis=0)
        cw, ch = max(x1 - x0, 1.0) / 64, max(y1 - y0, 1.0) / 64
        for t in (np.arange(-1, 66, dtype=np.float64), np.arange(-1, 65, dtype=np.float64) + 0.5):
            gx, gy = np.meshgrid(x0 + t * cw, y0 + t * ch)
            pts.append(np.stack([gx.ravel(), gy.ravel()], axis=1))
        pts.append(ln)
        pts.append(0.5 * (ln[:-1] + ln[1:]))
        pts.append(ln + [0.01, 0.0])
        pts.append(np.stack([x0 + (x1 - x0 + 2) * rng.random(3000) - 1, y0 + (y1 - y0 + 2) * rng.random(3000) - 1], axis=1))
    xy = np.vstack(pts)
    return b, D.points_wkb(xy[:, 0].copy(), xy[:, 1].copy())


@pytest.mark.parametrize("d", [0.0, 0.01, 0.3])
def test_line_raster_edges_vs_oracle(d, opts):
    b, p = _line_raster_probe_build()
    say("line raster edges", d, "probes", len(p.offsets) - 1)
    ix, rf = build_pair(b.to_arrow(), opts)
    r = G.join_device(ix, rf, p.to_arrow(), "dwithin", distance=d, refine_order=0)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "dwithin", distance=d, nthreads=8)
    assert len(ob) > 1000
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (len(gb), len(ob))


def test_corpus_dwithin_all_pairs_vs_oracle(opts):
    """Every (polygon, point) pair of the adversarial corpus (points on
    vertices and edges, one ulp off them, in holes, rings of up to 39
    segments with segment trees) under the distance refine."""
    cb, cp = C.corpus()
    b, p = corpus_arrow(cb), corpus_arrow(cp)
    nb, np_ = len(b), len(p)
    bi = np.repeat(np.arange(nb), np_).astype(np.uint32)
    pi = np.tile(np.arange(np_), nb).astype(np.uint32)
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(b)
    rf.push_build(b)
    rf.finish_building()
    sb, sp = O.Store(cb), O.Store(cp)
    for d in [0.0, 1e-9, 0.25, 1.0]:
        say("corpus dwithin", d)
        want = O.refine(sb, sp, "dwithin", bi, pi, distance=d, nthreads=8)
        ob, op = rf.refine_distance(p, d, bi, pi)
        eq, miss, extra = pairs_equal((ob, op), (bi[want], pi[want]))
        assert eq, (d, miss, extra)


@pytest.mark.parametrize("refine_order", [0, 1, 2])
@pytest.mark.parametrize("d", [0.0, 0.05, 0.7])
def test_dwithin_zipf_polygons_vs_oracle(d, refine_order, opts):
    """Large rings (up to 4096 vertices, segment trees) under the distance
    refine: inside, on-boundary-free outside within d, and beyond d; in
    filter order, grouped by polygon, and fused into the filter walk."""
    b, p = D.workload_pip_zipf(60_000, 600, seed=19)
    say("dwithin zipf", d, refine_order)
    ix, rf = build_pair(b.to_arrow(), opts)
    r = G.join_device(ix, rf, G.DeviceWkb.from_host(p), "dwithin", distance=d, refine_order=refine_order)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "dwithin", distance=d, nthreads=8)
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (len(gb), len(ob))
    assert len(gb) > 100


@pytest.mark.parametrize("pred,d", [("contains", 0.0), ("dwithin", 0.05)])
def test_autotuned_join_gives_the_same_pairs(pred, d, opts):
    """The autotuner rebuilds the tree with its chosen leaf size and picks
    the refine order and kernel variant; the tuned join (refine_order -1)
    returns the oracle's pairs."""
    b, p = D.workload_pip_zipf(80_000, 800, seed=31)
    ix, rf = build_pair(b.to_arrow(), opts)
    dev = G.DeviceWkb.from_host(p)
    t = G.autotune(ix, rf, dev.slice(0, 40_000), pred, distance=d)
    say("autotune", pred, t)
    assert t["leaf"] in (4, 8, 16) and t["evaluations"] >= 5
    r = G.join_device(ix, rf, dev, pred, distance=d, refine_order=-1)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), pred, distance=d, nthreads=8)
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (len(gb), len(ob))


def test_set_tuning_orders_and_waves_give_the_oracle_pairs(opts):
    """SjhipIndexSetTuning: every candidate order and point-in-polygon wave
    count it accepts gives the oracle's pairs through refine_order=-1;
    values outside its range are rejected."""
    b, p = D.workload_pip_zipf(100_000, 1_000, seed=13)
    ix, rf = build_pair(b.to_arrow(), opts)
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "contains", nthreads=8)
    for order, waves in [(0, 4), (1, 6), (1, 8), (-1, 0)]:
        ix.set_tuning(kind=0, refine_order=order, pip_waves=waves)
        r = G.join_device(ix, rf, p.to_arrow(), "contains", refine_order=-1)
        gb, gp = r.to_host()
        say("tuning", order, waves, len(gb))
        assert np.array_equal(gb, ob) and np.array_equal(gp, op), (order, waves, len(gb), len(ob))
    for kind, order, waves in [(3, 0, 0), (0, 3, 0), (0, 1, 5)]:
        with pytest.raises(G.GpuSpatialError):
            ix.set_tuning(kind=kind, refine_order=order, pip_waves=waves)


def test_segment_leaf_32_gives_the_oracle_pairs(opts):
    """Every walk over the build rings' segment trees with 32 segments per
    leaf (the autotuner's other choice): point in polygon (contains,
    intersects, the certified and the FP64 path), distance to polygons,
    polygon x polygon, and the adversarial corpus."""
    b, p = D.workload_pip_zipf(100_000, 1_000, seed=11)
    ix, rf = build_pair(b.to_arrow(), opts)
    rf.set_segment_leaf(32)
    OB, OP = O.Store(b.as_tuple()), O.Store(p.as_tuple())
    for pred, d, precision in [("contains", 0.0, 0), ("intersects", 0.0, 1), ("dwithin", 0.05, 0)]:
        r = G.join_device(ix, rf, p.to_arrow(), pred, distance=d, precision=precision)
        gb, gp = r.to_host()
        ob, op = O.join(OB, OP, pred, distance=d, nthreads=8)
        say("leaf 32", pred, precision, len(gb))
        assert np.array_equal(gb, ob) and np.array_equal(gp, op), (pred, len(gb), len(ob))
    a, q = D.workload_poly_poly(3_000, 3_000, seed=29)
    ix2, rf2 = build_pair(a.to_arrow(), opts)
    rf2.set_segment_leaf(32)
    r = G.join_device(ix2, rf2, G.DeviceWkb.from_host(q), "intersects")
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(a.as_tuple()), O.Store(q.as_tuple()), "intersects", nthreads=8)
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), ("polypoly", len(gb), len(ob))
    cb, cp = C.corpus()
    bcol, pcol = corpus_arrow(cb), corpus_arrow(cp)
    rf3 = G.GpuSpatialRefiner(opts)
    rf3.init_build_schema(bcol)
    rf3.push_build(bcol)
    rf3.finish_building()
    rf3.set_segment_leaf(32)
    nb, np_ = len(bcol), len(pcol)
    bi = np.repeat(np.arange(nb), np_).astype(np.uint32)
    pi = np.tile(np.arange(np_), nb).astype(np.uint32)
    E = expected()
    for pred in ["intersects", "contains", "covers", "touches"]:
        want = np.unpackbits(E[f"corpus/{pred}/mask"])[: len(bi)].astype(bool)
        ob, op = rf3.refine(pcol, pred, bi, pi)
        eq, miss, extra = pairs_equal((ob, op), (bi[want], pi[want]))
        assert eq, ("corpus", pred, miss, extra)
    with pytest.raises(G.GpuSpatialError, match="16 or 32"):
        rf3.set_segment_leaf(8)


def test_poly_poly_large_rings_vs_oracle(opts):
    """Polygon x polygon intersects with rings of up to 1024 vertices (both
    sides get segment trees): the indexed crossing counts and point tests
    decide every pair as the oracle does."""
    a, b = D.workload_poly_poly(3_000, 3_000, seed=29)
    say("polypoly large", int(a.offsets[-1]), int(b.offsets[-1]))
    ix, rf = build_pair(a.to_arrow(), opts)
    r = G.join_device(ix, rf, G.DeviceWkb.from_host(b), "intersects")
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(a.as_tuple()), O.Store(b.as_tuple()), "intersects", nthreads=8)
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (len(gb), len(ob))
    assert len(gb) > 1000


def _star(cx, cy, r_in, r_out, n, phase=0.0):
    ang = phase + np.arange(n) * (2 * np.pi / n)
    r = np.where(np.arange(n) % 2 == 0, r_out, r_in)
    ring = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], axis=1)
    return np.vstack([ring, ring[:1]])


def test_poly_poly_raster_cases_vs_oracle(opts):
    """Polygon pairs that the raster pass decides or leaves to the exact
    kernel: nested stars (a common point at the centre of the rects'
    overlap), stars apart whose rects overlap (a raster with no point in
    the overlap), stars touching at a vertex, squares sharing an edge and a
    corner (many vertices on their sides), and crossing stars."""
    side = np.linspace(0.0, 10.0, 41)[:-1]
    sq = lambda x0, y0: np.vstack([np.stack([x0 + side, y0 + 0 * side], 1), np.stack([x0 + 10 + 0 * side, y0 + side], 1),
                                   np.stack([x0 + 10 - side, y0 + 10 + 0 * side], 1),
                                   np.stack([x0 + 0 * side, y0 + 10 - side], 1), [[x0, y0]]])
    build = [_star(0, 0, 6, 10, 200), _star(40, 0, 2, 10, 64), sq(100, 0), _star(200, 0, 7, 10, 300)]
    touch = _star(60, 0, 7, 10, 128)
    touch[64] = (50.0, 0.0)  # build 1's first vertex: the two stars touch there
    probe = [_star(0, 0, 1, 2, 50),  # nested in build 0
             _star(47.5, 7.5, 0.5, 2, 40),  # in build 1's rect corner, outside the star
             touch,
             sq(110, 0), sq(110, 10), sq(95, 5),  # shares an edge / a corner / overlaps build 2
             _star(212, 3, 6, 9, 256, phase=0.3),  # crosses build 3
             _star(0, 30, 6, 10, 200)]  # disjoint rects
    b = D.rings_wkb([len(r) for r in build], np.vstack(build), 3)
    p = D.rings_wkb([len(r) for r in probe], np.vstack(probe), 3)
    ix, rf = build_pair(b.to_arrow(), opts)
    r = G.join_device(ix, rf, p.to_arrow(), "intersects", sort_output=True)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "intersects", nthreads=8)
    say("poly raster cases", list(zip(ob.tolist(), op.tolist())))
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (list(zip(gb, gp)), list(zip(ob, op)))
    assert (1, 2) in set(zip(ob.tolist(), op.tolist())) and len(ob) >= 6


def _near_contact_pairs(rng, n_base=40):
    """Star polygons and polygons in exact or near contact with them."""
    def close(r):
        r = np.asarray(r, np.float64).copy()
        r[-1] = r[0]
        return r

    build, probe = [], []
    for _ in range(n_base):
        c = rng.uniform(0, 1000, 2)
        n = int(rng.integers(4, 300))
        R = float(rng.uniform(0.5, 3.0))
        A = _star(c[0], c[1], 0.7 * R, R, n, phase=float(rng.uniform(0, 1)))
        build.append(A)
        m = len(A) - 1
        j = int(rng.integers(0, m))
        # a copy moved so that its first vertex lands exactly on vertex j
        B = _star(c[0], c[1], 0.6 * R, 0.9 * R, int(rng.integers(4, 64)))
        B = close(B - B[0] + A[j])
        B[0] = A[j]
        B[-1] = A[j]
        probe.append(B)
        # the mirror image across the line of edge (j, j + 1): the edge is shared
        a, b = A[j], A[j + 1]
        d = (b - a) / np.hypot(*(b - a))
        v = A - a
        M = a + 2 * (v @ d)[:, None] * d[None, :] - v
        M[j], M[j + 1] = a, b
        probe.append(close(M) if j != 0 else np.vstack([M[:-1], M[:1]]))
        # every vertex moved by a few ulps
        P = A[:-1] + np.spacing(A[:-1]) * rng.integers(-3, 4, A[:-1].shape)
        probe.append(close(np.vstack([P, P[:1]])))
        # the mirror image across the vertical line through the rightmost vertex
        k = int(np.argmax(A[:-1, 0]))
        X = A.copy()
        X[:, 0] = 2 * A[k, 0] - A[:, 0]
        X[k] = A[k]
        probe.append(close(X) if k != 0 else np.vstack([X[:-1], X[:1]]))
        # a copy scaled by 1 + 1e-9 about a far point
        far = c + 1e4
        probe.append(close(far + (A - far) * (1 + 1e-9)))
    return build, probe


@pytest.mark.parametrize("precision", [0, 1])
def test_poly_poly_near_contact_vs_oracle(precision, opts):
    """Polygon pairs in exact or near contact, where the certified FP32
    point and crossing tests of the two-phase kernels are uncertain and hand
    pairs to the inline-precision kernel: a shared vertex, a shared edge,
    every vertex a few ulps off, a touching mirror image and a copy scaled
    by 1 + 1e-9; every (build, probe) pair through the refiner, certified
    (precision 0) and inline (precision 1), against the oracle."""
    build, probe = _near_contact_pairs(np.random.default_rng(31))
    b = D.rings_wkb([len(r) for r in build], np.vstack(build), 3)
    p = D.rings_wkb([len(r) for r in probe], np.vstack(probe), 3)
    B, P = O.Store(b.as_tuple()), O.Store(p.as_tuple())
    bi = np.repeat(np.arange(B.n), P.n).astype(np.uint32)
    pi = np.tile(np.arange(P.n), B.n).astype(np.uint32)
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(b.to_arrow())
    rf.push_build(b.to_arrow())
    rf.finish_building()
    rf.set_precision(precision)
    ob, op = rf.refine(p.to_arrow(), "intersects", bi, pi)
    want = O.refine(B, P, "intersects", bi, pi)
    say("near contact", precision, int(want.sum()))
    eq, miss, extra = pairs_equal((ob, op), (bi[want], pi[want]))
    assert eq, (precision, miss, extra)
    # every base polygon meets the five polygons made from it
    assert int(want.sum()) >= 5 * B.n


@pytest.mark.parametrize("rev", [False, True])
def test_hole_corpus_polypoly_vs_reference_goldens(rev, opts):
    """Polygons with convex, concave and pentagram holes x small polygons,
    every pair through the refiner ABI and the fused join, against the
    reference TG goldens (the holes decided as tg_ring_contains_ring does)."""
    hb, hp = C.hole_corpus()
    b, p = corpus_arrow(hb), corpus_arrow(hp)
    if rev:
        b, p = p, b
    nb, np_ = len(b), len(p)
    bi = np.repeat(np.arange(nb), np_).astype(np.uint32)
    pi = np.tile(np.arange(np_), nb).astype(np.uint32)
    E = expected()
    key = "holes_polypoly_rev" if rev else "holes_polypoly"
    # the goldens index (hole polygon, small polygon) pairs in build-major order
    want = np.unpackbits(E[f"{key}/intersects/mask"])[: len(bi)].astype(bool)
    if rev:
        want = want.reshape(np_, nb).T.reshape(-1)
    say("hole corpus", rev, int(want.sum()))
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(b)
    rf.push_build(b)
    rf.finish_building()
    ob, op = rf.refine(p, "intersects", bi, pi)
    eq, miss, extra = pairs_equal((ob, op), (bi[want], pi[want]))
    assert eq, (rev, miss, extra)
    ix = G.GpuSpatialIndex(opts)
    ix.build_from_refiner(rf)
    r = G.join_device(ix, rf, p, "intersects", sort_output=True)
    gb, gp = r.to_host()
    assert pairs_equal((gb, gp), (bi[want], pi[want]))[0]


RELATE = [("pp", "p", "p", ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"])] + \
    [(a + b, a, b, ["intersects"]) for a, b in (("l", "p"), ("p", "l"), ("l", "l"), ("q", "l"), ("l", "q"))]


@pytest.mark.parametrize("name,a,b,preds", RELATE)
def test_relate_corpus_vs_reference_goldens(name, a, b, preds, opts):
    """(Multi)polygon x (multi)polygon under every predicate and intersects
    with (multi)linestrings, every pair of the relate corpus through the
    refiner ABI, and the rect-filtered pairs through the fused join, against
    the reference TG goldens."""
    cols = {"p": C.relate_polygons(), "l": C.relate_lines(), "q": C.relate_points()}
    A, B = corpus_arrow(C.column(cols[a])), corpus_arrow(C.column(cols[b]))
    na, nb = len(A), len(B)
    ai = np.repeat(np.arange(na), nb).astype(np.uint32)
    bi = np.tile(np.arange(nb), na).astype(np.uint32)
    E = expected()
    ix, rf = build_pair(A, opts)
    for pred in preds:
        say("relate", name, pred)
        want = np.unpackbits(E[f"relate_{name}/{pred}/mask"])[: len(ai)].astype(bool)
        ob, op = rf.refine(B, pred, ai, bi)
        eq, miss, extra = pairs_equal((ob, op), (ai[want], bi[want]))
        assert eq, (name, pred, miss, extra)
        r = G.join_device(ix, rf, B, pred, sort_output=True)
        gb, gp = r.to_host()
        r.release()
        assert pairs_equal((gb, gp), (ai[want], bi[want]))[0], (name, pred, "fused")


@pytest.mark.parametrize("d", [0.0, 0.3, 1.5])
def test_dwithin_hole_polygons_vs_oracle(d, opts):
    """ST_DWithin of points against polygons with holes and multipolygons:
    points inside holes, on hole edges and deep inside solid parts."""
    hb, _ = C.hole_corpus()
    b = corpus_arrow(hb)
    g = np.arange(-1.0, 17.01, 0.25)
    xs, ys = np.meshgrid(g, g)
    rng = np.random.default_rng(41)
    px = np.concatenate([xs.ravel(), rng.uniform(-1, 17, 4000)])
    py = np.concatenate([ys.ravel(), rng.uniform(-1, 17, 4000)])
    p = D.points_wkb(px, py)
    ix, rf = build_pair(b, opts)
    r = G.join_device(ix, rf, p.to_arrow(), "dwithin", distance=d, sort_output=True)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(hb), O.Store(p.as_tuple()), "dwithin", distance=d, nthreads=8)
    say("hole dwithin", d, len(ob))
    assert np.array_equal(gb, ob) and np.array_equal(gp, op), (d, len(gb), len(ob))


def test_pair_distances_match_oracle(opts):
    """SjhipRefinerPairDistances: the FP64 distance of every (polygon or
    line, point) pair of the hole and relate corpora equals the oracle's bit
    for bit (0 inside polygons and on lines, holes included)."""
    hb, _ = C.hole_corpus()
    rng = np.random.default_rng(59)
    q = np.concatenate([np.round(rng.uniform(-2, 18, size=(600, 2)) * 4) / 4, rng.uniform(-2, 18, size=(600, 2))])
    p = D.points_wkb(q[:, 0], q[:, 1])
    for col in (hb, C.column(C.relate_lines()), C.column(C.relate_polygons())):
        b = corpus_arrow(col)
        rf = G.GpuSpatialRefiner(opts)
        rf.init_build_schema(b)
        rf.push_build(b)
        rf.finish_building()
        nb = len(b)
        bi = np.repeat(np.arange(nb), len(q)).astype(np.uint32)
        pi = np.tile(np.arange(len(q)), nb).astype(np.uint32)
        got = rf.pair_distances(p.to_arrow(), bi, pi)
        want = O.pair_distances(O.Store(col), O.Store(p.as_tuple()), bi, pi, nthreads=8)
        assert np.array_equal(got, want, equal_nan=True), int(np.sum(~((got == want) | (np.isnan(got) & np.isnan(want)))))
        assert np.sum(got == 0) > 100


@pytest.mark.parametrize("d", [0.0, 0.05, 0.4])
def test_dwithin_certified_equals_plain_fp64(d, opts):
    """The certified FP32 distance terms with FP64 re-evaluation give the
    pair set of the all-FP64 plain kernel, on random walks (line rasters and
    segment trees) and on Zipf polygons."""
    for b, p in (D.workload_dwithin_walks(400_000, 20_000, seed=61), D.workload_pip_zipf(200_000, 2_000, seed=62)):
        ix, rf = build_pair(b.to_arrow(), opts)
        dev = G.DeviceWkb.from_host(p)
        r0 = G.join_device(ix, rf, dev, "dwithin", distance=d, precision=0, sort_output=True)
        r1 = G.join_device(ix, rf, dev, "dwithin", distance=d, precision=1, sort_output=True)
        a, b_ = r0.to_host(), r1.to_host()
        say("certified dwithin", d, len(a[0]), r0.segment_tests, r0.fallback_tests)
        assert np.array_equal(a[0], b_[0]) and np.array_equal(a[1], b_[1]), (d, len(a[0]), len(b_[0]))
        assert len(a[0]) > 0 or d == 0.0


def test_dwithin_polygons_and_refine_distance_abi(opts):
    b, p = D.workload_pip_uniform(20_000, 200, seed=17, extent=100.0 * np.sqrt(200 / 10_000))
    ix, rf = build_pair(b.to_arrow(), opts)
    r = G.join_device(ix, rf, p.to_arrow(), "dwithin", distance=0.05)
    gb, gp = r.to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "dwithin", distance=0.05, nthreads=8)
    assert np.array_equal(gb, ob) and np.array_equal(gp, op)
    # the ABI extension over oracle candidates of the expanded filter
    pr = O.expand_rects(O.Store(p.as_tuple()).rects(False), 0.05)
    cb, cp = O.filter_pairs(O.Store(b.as_tuple()).rects(False), pr)
    kb, kp = rf.refine_distance(p.to_arrow(), 0.05, cb, cp)
    assert pairs_equal((kb, kp), (ob, op))[0]


@pytest.mark.parametrize("k", [1, 8])
def test_knn_exact_with_id_ties(k, opts):
    rng = np.random.default_rng(19)
    obj = np.round(rng.uniform(0, 50, size=(200_000, 2)) * 8) / 8  # grid: exact ties
    q = D.clustered_xy(rng, 20_000, 50.0, n_clusters=8)
    q[::5] = np.round(q[::5] * 8) / 8
    bcol = D.points_wkb(obj[:, 0], obj[:, 1])
    qcol = D.points_wkb(q[:, 0], q[:, 1])
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(bcol.to_arrow())
    rf.push_build(bcol.to_arrow())
    rf.finish_building()
    ids, d = rf.knn(qcol.to_arrow(), k)
    oid, od = O.knn(O.Store(bcol.as_tuple()), O.Store(qcol.as_tuple()), k, nthreads=8)
    assert np.array_equal(ids, oid)
    assert np.array_equal(d.view(np.uint64), od.view(np.uint64))


def _knn_rows(ids, d):
    """Fixed rows (id 0xffffffff / inf padding) as variable-length rows."""
    keep = ids != 0xFFFFFFFF
    off = np.concatenate([[0], np.cumsum(keep.sum(axis=1))]).astype(np.uint64)
    return off, ids[keep], d[keep]


@pytest.mark.parametrize("k", [1, 6])
def test_knn_point_queries_against_polygons_and_lines(k, opts):
    """Point queries against (multi)polygons with holes and linestrings:
    geo distances (0 inside a polygon or on a line), ids and FP64 distances
    identical to the brute-force oracle."""
    hb, _ = C.hole_corpus()
    lines = C.column(C.relate_lines())
    for name, col in (("polygons", hb), ("lines", lines)):
        b = corpus_arrow(col)
        rng = np.random.default_rng(43 + k)
        q = np.round(rng.uniform(-2, 18, size=(3000, 2)) * 8) / 8
        p = D.points_wkb(q[:, 0], q[:, 1])
        rf = G.GpuSpatialRefiner(opts)
        rf.init_build_schema(b)
        rf.push_build(b)
        rf.finish_building()
        ids, d = rf.knn(p.to_arrow(), k)
        off, oi, od = O.knn_brute(O.Store(col), O.Store(p.as_tuple()), k, nthreads=8)
        go, gi, gd = _knn_rows(ids, d)
        say("knn", name, k, len(oi))
        assert np.array_equal(go, off.astype(np.uint64)) and np.array_equal(gi, oi) and np.array_equal(gd, od), name


def test_knn_polygon_queries_against_points(opts):
    """Queries of any geometry against point objects (the distance of a
    pair is symmetric)."""
    hb, hp = C.hole_corpus()
    rng = np.random.default_rng(47)
    obj = np.round(rng.uniform(-1, 17, size=(5000, 2)) * 4) / 4
    b = D.points_wkb(obj[:, 0], obj[:, 1])
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(b.to_arrow())
    rf.push_build(b.to_arrow())
    rf.finish_building()
    for col in (hb, hp):
        ids, d = rf.knn(corpus_arrow(col), 5)
        off, oi, od = O.knn_brute(O.Store(b.as_tuple()), O.Store(col), 5, nthreads=8)
        go, gi, gd = _knn_rows(ids, d)
        assert np.array_equal(go, off.astype(np.uint64)) and np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("ties", [False, True])
def test_knn_tie_breakers_vs_oracle(ties, opts):
    """include_tie_breakers: objects on a grid (exact ties at the k-th
    distance, some a few ulps apart) and line objects, rows of variable
    length identical to the brute-force oracle's."""
    rng = np.random.default_rng(53)
    obj = np.round(rng.uniform(0, 30, size=(40_000, 2)) * 4) / 4
    obj[:500] += rng.choice([-1e-12, 0.0, 1e-12], size=(500, 2))
    q = np.round(rng.uniform(0, 30, size=(4000, 2)) * 4) / 4
    b, p = D.points_wkb(obj[:, 0], obj[:, 1]), D.points_wkb(q[:, 0], q[:, 1])
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(b.to_arrow())
    rf.push_build(b.to_arrow())
    rf.finish_building()
    for k in (1, 8):
        goff, gi, gd = rf.knn_ties(p.to_arrow(), k, include_tie_breakers=ties)
        off, oi, od = O.knn_brute(O.Store(b.as_tuple()), O.Store(p.as_tuple()), k, include_tie_breakers=ties,
                                  nthreads=8)
        say("knn ties", ties, k, len(oi))
        assert np.array_equal(goff, off.astype(np.uint64)) and np.array_equal(gi, oi) and np.array_equal(gd, od)
        if ties:
            assert np.any(np.diff(off) > k)
    # lines as objects, point queries
    lines = C.column(C.relate_lines())
    lb = corpus_arrow(lines)
    rl = G.GpuSpatialRefiner(opts)
    rl.init_build_schema(lb)
    rl.push_build(lb)
    rl.finish_building()
    q2 = np.round(rng.uniform(-1, 13, size=(2000, 2)) * 2) / 2
    p2 = D.points_wkb(q2[:, 0], q2[:, 1])
    goff, gi, gd = rl.knn_ties(p2.to_arrow(), 3, include_tie_breakers=ties)
    off, oi, od = O.knn_brute(O.Store(lines), O.Store(p2.as_tuple()), 3, include_tie_breakers=ties, nthreads=8)
    assert np.array_equal(goff, off.astype(np.uint64)) and np.array_equal(gi, oi) and np.array_equal(gd, od)


def test_knn_device_outputs_and_resident_tree(opts):
    """The device-output KNN entry point, and repeated calls over the
    object tree kept by the first call, give the oracle's exact lists."""
    rng = np.random.default_rng(23)
    obj = D.clustered_xy(rng, 300_000, 100.0, n_clusters=16)
    q = D.clustered_xy(rng, 30_000, 100.0, n_clusters=16)
    bcol = D.points_wkb(obj[:, 0], obj[:, 1])
    qcol = D.points_wkb(q[:, 0], q[:, 1])
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(bcol.to_arrow())
    rf.push_build(bcol.to_arrow())
    rf.finish_building()
    oid, od = O.knn(O.Store(bcol.as_tuple()), O.Store(qcol.as_tuple()), 8, nthreads=8)
    for _ in range(2):
        ids, db, ms = rf.knn_device(G.DeviceWkb.from_host(qcol), 8)
        gi = ids.download(np.uint32, len(q) * 8).reshape(len(q), 8)
        gd = db.download(np.float64, len(q) * 8).reshape(len(q), 8)
        assert np.array_equal(gi, oid)
        assert np.array_equal(gd.view(np.uint64), od.view(np.uint64))
        assert ms > 0
    hi, hd = rf.knn(qcol.to_arrow(), 8)
    assert np.array_equal(hi, oid)


def test_workspace_reuse_across_joins_of_growing_size(opts):
    """Joins on one refiner reuse its device workspace.  A polygon-major
    join (whose regrouping sort leaves the candidate arrays with different
    sizes) is followed by joins with more candidates and other predicates;
    each gives the oracle's pairs."""
    b, p = D.workload_pip_zipf(120_000, 1_000, seed=17)
    ix, rf = build_pair(b.to_arrow(), opts)
    store_b = O.Store(b.as_tuple())
    plan = [(2_000, "contains", 0.0, 1), (40_000, "intersects", 0.0, 0), (10_000, "contains", 0.0, 1),
            (120_000, "dwithin", 0.2, 0), (60_000, "contains", 0.0, 1), (120_000, "intersects", 0.0, 1)]
    for n, pred, d, order in plan:
        sub = D.take_rows(p, np.arange(n))
        r = G.join_device(ix, rf, sub.to_arrow(), pred, distance=d, refine_order=order, sort_output=True)
        gb, gp = r.to_host()
        r.release()
        ob, op = O.join(store_b, O.Store(sub.as_tuple()), pred, distance=d, nthreads=8)
        say("workspace reuse", n, pred, order, len(gb))
        assert np.array_equal(gb, ob) and np.array_equal(gp, op), (n, pred, len(gb), len(ob))


def test_decode_errors_name_the_row(opts):
    rf = G.GpuSpatialRefiner(opts)
    for bad in C.invalid_wkbs():
        col = corpus_arrow(C.column([C.point_wkb(0, 0), C.point_wkb(1, 1), bad]))
        rf.clear()
        rf.init_build_schema(col)
        with pytest.raises(G.GpuSpatialError, match="row 2"):
            rf.push_build(col)


def test_nulls_empties_big_endian_large_and_view(opts):
    import pyarrow as pa
    polys = [C.polygon_wkb([[(0, 0), (4, 0), (4, 4), (0, 4), (0, 0)]], big_endian=True), None,
             C.polygon_wkb([[(5, 5), (9, 5), (9, 9), (5, 9), (5, 5)]]),
             b"\x01" + (3).to_bytes(4, "little") + (0).to_bytes(4, "little")]
    pts = [C.point_wkb(1, 1, big_endian=True), C.point_wkb(6, 6), None, C.point_wkb(float("nan"), float("nan")),
           C.point_wkb(4, 2), C.point_wkb(7, 7, big_endian=True)]
    want = {(0, 0), (2, 1), (0, 4), (2, 5)}
    for typ in (pa.binary(), pa.large_binary(), pa.binary_view()):
        b = pa.array(polys, type=typ)
        p = pa.array(pts, type=typ)
        ix, rf = build_pair(b, opts)
        cb, cp = ix.probe(G.compute_rects(p))
        ob, op = rf.refine(p, "intersects", cb, cp)
        assert set(zip(ob.tolist(), op.tolist())) == want, typ
        ob, op = rf.refine(p, "contains", cb, cp)
        assert set(zip(ob.tolist(), op.tolist())) == {(0, 0), (2, 1), (2, 5)}, typ


@pytest.mark.parametrize("pred,d", [("contains", 0.0), ("intersects", 0.0), ("dwithin", 0.05)])
def test_invalid_probe_rows_sort_after_corner_points(pred, d, opts):
    """The fused join orders the valid probe points by Morton key and drops
    the invalid rows (null, empty, NaN), whose key sorts after every valid
    one.  Valid points at the far corner of the points' frame get keys whose
    top bits equal the invalid key's; with null, empty and NaN rows mixed
    among them, every corner point must still be joined, in every candidate
    order."""
    import pyarrow as pa
    rng = np.random.default_rng(43)
    n = 50_000
    x, y = rng.uniform(0, 100, n), rng.uniform(0, 100, n)
    # the frame's far corner: the largest x and y, and points within a few
    # key units of it
    x[:200] = 100.0 - rng.uniform(0, 1e-3, 200)
    y[:200] = 100.0 - rng.uniform(0, 1e-3, 200)
    x[0], y[0] = 100.0, 100.0
    col = D.points_wkb(x, y)
    rows = [bytes(col.data[col.offsets[i]:col.offsets[i + 1]]) for i in range(n)]
    # null, empty and NaN rows spread among the valid ones
    for i in rng.choice(np.arange(200, n), 3000, replace=False):
        k = int(rng.integers(0, 3))
        rows[i] = None if k == 0 else (C.point_wkb(float("nan"), float("nan")) if k == 1 else
                                       b"\x01" + (1).to_bytes(4, "little") + np.array([np.nan, np.nan]).tobytes())
    p = pa.array(rows, type=pa.binary())
    # polygons over the corner and over the rest
    b, _ = D.workload_pip_uniform(10, 300, seed=44, extent=100.0)
    corner = [np.array([[99.0, 99.0], [101.0, 99.0], [101.0, 101.0], [99.0, 101.0], [99.0, 99.0]])]
    bc = D.rings_wkb([5], np.vstack(corner), 3)
    bcol = pa.array([bytes(bc.data[bc.offsets[0]:bc.offsets[1]])] +
                    [bytes(b.data[b.offsets[i]:b.offsets[i + 1]]) for i in range(len(b.offsets) - 1)], type=pa.binary())
    ix, rf = build_pair(bcol, opts)
    ob, op = O.join(O.Store(bcol), O.Store(p), pred, distance=d, nthreads=8)
    got = set()
    for order in ((0, 1, 2) if pred == "dwithin" else (0, 1)):
        say("corner points", pred, d, order)
        r = G.join_device(ix, rf, p, pred, distance=d, refine_order=order)
        gb, gp = r.to_host()
        assert np.array_equal(gb, ob) and np.array_equal(gp, op), (pred, order, len(gb), len(ob))
        got = set(gp[gb == 0].tolist())
    # every corner point lies inside the corner square (build row 0)
    assert set(range(200)) <= got, sorted(set(range(200)) - got)[:10]


def test_empty_build_and_probe(opts):
    import pyarrow as pa
    empty = pa.array([], type=pa.binary())
    ix, rf = build_pair(empty, opts)
    p = D.points_wkb(np.array([1.0]), np.array([2.0])).to_arrow()
    cb, cp = ix.probe(G.compute_rects(p))
    assert len(cb) == 0
    ob, op = rf.refine(p, "intersects", cb, cp)
    assert len(ob) == 0
    b = D.points_wkb(np.array([1.0]), np.array([2.0])).to_arrow()
    ix2, rf2 = build_pair(b, opts)
    cb, cp = ix2.probe(np.zeros((0, 4), np.float32))
    assert len(cb) == 0


def test_multi_batch_build_ids_are_global(opts):
    """Build ids are global in push order across push_build calls
    (gpuspatial_c.h:186-192)."""
    b, p = D.workload_pip_uniform(5_000, 60, seed=23, extent=100.0 * np.sqrt(60 / 10_000))
    parts = [b.slice(0, 20), b.slice(20, 45), b.slice(45, 60)]
    rf = G.GpuSpatialRefiner(G.GpuSpatialOptions())
    rf.init_build_schema(parts[0].to_arrow())
    for q in parts:
        rf.push_build(q.to_arrow())
    rf.finish_building()
    ix = G.GpuSpatialIndex(G.GpuSpatialOptions())
    ix.build_from_refiner(rf)
    gb, gp = G.join_device(ix, rf, p.to_arrow(), "intersects").to_host()
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "intersects")
    assert np.array_equal(gb, ob) and np.array_equal(gp, op)


def test_probe_callback_stop_propagates(opts):
    b = read_fixture("synthetic_pip_polygons.parquet")
    p = read_fixture("synthetic_pip_points.parquet")
    ix, _ = build_pair(b, opts, from_refiner=False)
    with pytest.raises(G.GpuSpatialError):
        ix.probe(G.compute_rects(p), callback_rc=7)
