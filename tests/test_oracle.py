"""Pins the CPU oracle (oracle/sjoin_oracle.c) before anything is checked
against it: golden vectors minted with the reference's own TG library, the
reference's hand-written known answers, closed forms and brute force."""
import struct
from fractions import Fraction

import numpy as np
import pytest

import corpus as C
import oracle as O
from helpers import expected, read_fixture

FIXTURES = [
    ("pip", "synthetic_pip_polygons.parquet", "synthetic_pip_points.parquet",
     ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]),
    ("countries", "countries.parquet", "countries_points.parquet",
     ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]),
    ("polypoly", "synthetic_pip_polygons.parquet", "synthetic_poly_polygons2.parquet",
     ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]),
    ("countries_countries", "countries.parquet", "countries.parquet",
     ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]),
    ("points_in_countries_rev", "countries_points.parquet", "countries.parquet",
     ["within", "intersects", "coveredby", "touches"]),
]


@pytest.mark.parametrize("name,bf,pf,preds", FIXTURES)
def test_oracle_matches_reference_goldens_on_reference_fixtures(name, bf, pf, preds):
    B, P = O.Store(read_fixture(bf)), O.Store(read_fixture(pf))
    E = expected()
    for pred in preds:
        ob, op = O.join(B, P, pred, nthreads=4)
        assert np.array_equal(ob, E[f"{name}/{pred}/build"]), (name, pred)
        assert np.array_equal(op, E[f"{name}/{pred}/probe"]), (name, pred)


def test_oracle_matches_reference_on_adversarial_corpus():
    cb, cp = C.corpus()
    B, P = O.Store(cb), O.Store(cp)
    nb, np_ = B.n, P.n
    bi = np.repeat(np.arange(nb), np_).astype(np.uint32)
    pi = np.tile(np.arange(np_), nb).astype(np.uint32)
    E = expected()
    for pred in ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]:
        want = np.unpackbits(E[f"corpus/{pred}/mask"])[: len(bi)].astype(bool)
        assert np.array_equal(O.refine(B, P, pred, bi, pi), want), pred
        want2 = np.unpackbits(E[f"corpus_rev/{pred}/mask"])[: len(bi)].astype(bool)
        assert np.array_equal(O.refine(P, B, pred, pi, bi), want2), pred
    bb = np.repeat(np.arange(nb), nb).astype(np.uint32)
    pp = np.tile(np.arange(nb), nb).astype(np.uint32)
    want = np.unpackbits(E["corpus_polypoly/intersects/mask"])[: len(bb)].astype(bool)
    assert np.array_equal(O.refine(B, B, "intersects", bb, pp), want)


def test_oracle_hole_containment_matches_reference_on_hole_corpus():
    """Polygons with convex, concave and self-crossing (pentagram) holes x
    small polygons, every pair, both roles: TG decides holes in polygon x
    polygon intersects with tg_ring_contains_ring's convex/concave dispatch
    (tg.c:3119-3151).  Vertices inside a pentagram hole with edges crossing
    it are where a vertex-and-crossing test would differ."""
    hb, hp = C.hole_corpus()
    B, P = O.Store(hb), O.Store(hp)
    bi = np.repeat(np.arange(B.n), P.n).astype(np.uint32)
    pi = np.tile(np.arange(P.n), B.n).astype(np.uint32)
    E = expected()
    want = np.unpackbits(E["holes_polypoly/intersects/mask"])[: len(bi)].astype(bool)
    assert np.array_equal(O.refine(B, P, "intersects", bi, pi), want)
    want2 = np.unpackbits(E["holes_polypoly_rev/intersects/mask"])[: len(bi)].astype(bool)
    assert np.array_equal(O.refine(P, B, "intersects", pi, bi), want2)


RELATE = [("pp", "p", "p", ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"])] + \
    [(a + b, a, b, ["intersects"]) for a, b in (("l", "p"), ("p", "l"), ("l", "l"), ("q", "l"), ("l", "q"))]


def relate_cols():
    return {"p": C.column(C.relate_polygons()), "l": C.column(C.relate_lines()), "q": C.column(C.relate_points())}


@pytest.mark.parametrize("name,a,b,preds", RELATE)
def test_oracle_relate_corpus_matches_reference(name, a, b, preds):
    """(Multi)polygon x (multi)polygon under every predicate (TG's
    tg_poly_covers_poly, tg_poly_touches_poly and the multi-geometry
    dispatch, tg.c:4040-4249, 5925-6800, 15299-15330) and intersects with
    (multi)linestrings (tg.c:3194-3275, 3595-3611, 4016-4038), every pair of
    the relate corpus against the reference TG goldens."""
    cols = relate_cols()
    A, B = O.Store(cols[a]), O.Store(cols[b])
    ai = np.repeat(np.arange(A.n), B.n).astype(np.uint32)
    bi = np.tile(np.arange(B.n), A.n).astype(np.uint32)
    E = expected()
    for pred in preds:
        want = np.unpackbits(E[f"relate_{name}/{pred}/mask"])[: len(ai)].astype(bool)
        got = O.refine(A, B, pred, ai, bi)
        assert np.array_equal(got, want), (name, pred, int((got != want).sum()))


def test_corpus_reaches_every_ray_branch():
    """The corpus must exercise ON results and vertex-height nudges, or the
    pins above would not constrain those branches."""
    cb, cp = C.corpus()
    st_b, st_p = O.Store(cb).soa(), O.Store(cp).soa()
    pts = st_p["xy"][st_p["pt_off"][:-1]]
    xy = st_b["xy"]
    counts = {0: 0, 1: 0, 2: 0}
    vertex_height = 0
    for r in range(len(st_b["pt_off"]) - 1):
        ring = xy[st_b["pt_off"][r]:st_b["pt_off"][r + 1]]
        for i in range(len(ring) - 1):
            a, b = ring[i], ring[i + 1]
            for p in pts[::7]:
                rc = O.raycast(a[0], a[1], b[0], b[1], p[0], p[1])
                counts[rc] += 1
                if rc == 1 and (p[1] == a[1] or p[1] == b[1]):
                    vertex_height += 1
    assert counts[0] > 1000 and counts[1] > 100 and counts[2] > 50
    assert vertex_height > 0


def test_oracle_matches_reference_on_uniform_stars():
    from sjhip import datagen as D
    b, p = D.workload_pip_uniform(20_000, 200, seed=7, extent=100.0 * np.sqrt(200 / 10_000))
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), "intersects")
    E = expected()
    assert np.array_equal(ob, E["uniform_small/intersects/build"])
    assert np.array_equal(op, E["uniform_small/intersects/probe"])


@pytest.mark.skipif(not O.ref_available(), reason="reference TG library not built here")
@pytest.mark.parametrize("seed", [21, 22])
def test_oracle_matches_reference_tg_on_fresh_workloads(seed):
    from sjhip import datagen as D
    b, p = D.workload_pip_zipf(30_000, 300, seed=seed, max_vertices=512)
    B, P = O.Store(b.as_tuple()), O.Store(p.as_tuple())
    tb, tp = O.TgColumn(b.as_tuple(), index=1), O.TgColumn(p.as_tuple(), index=0)
    for pred in ["intersects", "contains"]:
        rb, rp = O.ref_join(tb, tp, B.rects(True), P.rects(True), pred)
        ob, op = O.join(B, P, pred)
        assert np.array_equal(ob, rb) and np.array_equal(op, rp)


@pytest.mark.skipif(not O.ref_available(), reason="reference TG library not built here")
@pytest.mark.parametrize("pred", ["touches", "equals", "intersects", "covers", "contains", "within", "coveredby"])
def test_oracle_touches_multipolygons_match_reference_tg(pred):
    """The multipolygon rules of tg_geom_touches (as the first argument: no
    other part may hold the point inside; as the second: any part's
    boundary), polygons with holes, both roles, against TG itself."""
    from touch_cases import touch_cases
    bcol, pcol = touch_cases()
    B, P = O.Store(bcol), O.Store(pcol)
    bi = np.repeat(np.arange(B.n), P.n).astype(np.uint32)
    pi = np.tile(np.arange(P.n), B.n).astype(np.uint32)
    tb, tp = O.TgColumn(bcol, index=1), O.TgColumn(pcol, index=0)
    assert np.array_equal(O.refine(B, P, pred, bi, pi), O.ref_refine(tb, tp, pred, bi, pi))
    assert np.array_equal(O.refine(P, B, pred, pi, bi), O.ref_refine(tp, tb, pred, pi, bi))
    if pred == "touches":
        m1, m2 = O.ref_refine(tb, tp, pred, bi, pi), O.ref_refine(tp, tb, pred, pi, bi)
        assert m1.sum() >= 8 and not np.array_equal(m1, m2)  # the multipolygon asymmetry is exercised


def _poly(rings):
    return C.polygon_wkb(rings)


def test_reference_known_answers_tg_predicates():
    """c/sedona-tg/src/tg.rs:299-359 (point / polygon cases)."""
    tri = _poly([[(0, 0), (5, 0), (0, 5), (0, 0)]])
    holed = _poly([[(0, 0), (3, 0), (3, 3), (0, 3), (0, 0)], [(1, 1), (1, 2), (2, 1), (1, 1)]])
    cases = [
        ("contains", tri, C.point_wkb(1, 1), True),
        ("contains", tri, C.point_wkb(-1, -1), False),
        ("covers", tri, C.point_wkb(1, 1), True),
        ("covers", tri, C.point_wkb(-1, -1), False),
        ("intersects", C.point_wkb(0, 1), C.point_wkb(0, 1), True),
        ("intersects", C.point_wkb(0, 1), C.point_wkb(0, 0), False),
        ("equals", C.point_wkb(0, 1), C.point_wkb(0, 1), True),
        ("within", C.point_wkb(1, 1), tri, True),
        ("within", C.point_wkb(-1, -1), tri, False),
        ("coveredby", C.point_wkb(1, 1), tri, True),
        ("coveredby", C.point_wkb(-1, -1), tri, False),
        # a point inside the hole is not contained
        ("contains", holed, C.point_wkb(1.25, 1.25), False),
        ("contains", holed, C.point_wkb(2.5, 2.5), True),
        # on the hole's boundary: covered, not contained
        ("covers", holed, C.point_wkb(1.0, 1.5), True),
        ("contains", holed, C.point_wkb(1.0, 1.5), False),
    ]
    for pred, a, b, want in cases:
        A, B = O.Store(C.column([a])), O.Store(C.column([b]))
        got = O.refine(A, B, pred, np.zeros(1, np.uint32), np.zeros(1, np.uint32))[0]
        assert got == want, (pred, want)


def test_reference_known_answer_touch_top_edge():
    """c/sedona-tg/src/tg.rs:419-445 (tidwall/tg#16)."""
    pts = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)] + [(0.0, i / 100.0) for i in range(99, -1, -1)]
    A = O.Store(C.column([_poly([pts])]))
    B = O.Store(C.column([C.point_wkb(0.5, 1.0)]))
    assert O.refine(A, B, "intersects", np.zeros(1, np.uint32), np.zeros(1, np.uint32))[0]


def test_reference_known_answer_intersects_vertex():
    """c/sedona-tg/src/tg.rs:381-417: a point on a vertex of a real polygon."""
    coords = [(-118.2936175, 34.026516), (-118.2935834, 34.026516), (-118.2935834, 34.0264634),
              (-118.2935834, 34.0263308), (-118.2935834, 34.0262071), (-118.2935834, 34.0260835),
              (-118.2935834, 34.0259598), (-118.2935835, 34.02592), (-118.2934886, 34.02592),
              (-118.2934886, 34.0259073), (-118.2934259, 34.0259073), (-118.2934259, 34.02592),
              (-118.2933447, 34.0259205), (-118.2933443, 34.0260829), (-118.2933442, 34.0262066),
              (-118.2933442, 34.0263303), (-118.2933442, 34.0264637), (-118.2933442, 34.0264907),
              (-118.2933259, 34.0264907), (-118.2933259, 34.0265669), (-118.2933035, 34.0265669),
              (-118.2931745, 34.0265669), (-118.2931745, 34.0265216), (-118.2931693, 34.0265191),
              (-118.2931876, 34.0263539), (-118.2931384, 34.0263608), (-118.2929819, 34.0263827),
              (-118.2929795, 34.0264157), (-118.292977, 34.0266764), (-118.2929762, 34.026758),
              (-118.2929659, 34.026758), (-118.2929658, 34.0267856), (-118.2931085, 34.0267856),
              (-118.2931085, 34.0267629), (-118.2931722, 34.0267629), (-118.2931722, 34.0267683),
              (-118.293303, 34.0267683), (-118.2934675, 34.0267684), (-118.2934675, 34.0267859),
              (-118.2934684, 34.0267859), (-118.2935959, 34.0267859), (-118.2935959, 34.0267078),
              (-118.2936175, 34.0267078), (-118.2936175, 34.0265897), (-118.2936175, 34.026516)]
    A = O.Store(C.column([_poly([coords])]))
    B = O.Store(C.column([C.point_wkb(-118.2934684, 34.0267859)]))
    assert O.refine(A, B, "intersects", np.zeros(1, np.uint32), np.zeros(1, np.uint32))[0]


def test_invalid_wkb_rejected_like_reference():
    for i, w in enumerate(C.invalid_wkbs()):
        with pytest.raises(ValueError):
            O.Store(C.column([C.point_wkb(0, 0), w]))


# ---------------------------------------------------------------------------
# Distance semantics (reference: rust/sedona-geo/src/st_dwithin.rs:114-175)


def _dw(a, b, d):
    A, B = O.Store(C.column([a])), O.Store(C.column([b]))
    return O.refine(A, B, "dwithin", np.zeros(1, np.uint32), np.zeros(1, np.uint32), distance=d)[0]


def test_dwithin_known_answers():
    assert _dw(C.point_wkb(0, 0), C.point_wkb(3, 4), 5.0)
    assert not _dw(C.point_wkb(0, 0), C.point_wkb(3, 4), 4.0)
    sq = _poly([[(0, 0), (0, 1), (1, 1), (1, 0), (0, 0)]])
    assert _dw(sq, C.point_wkb(0.5, 0.5), 1.0)
    assert not _dw(sq, C.point_wkb(5, 5), 1.0)
    line = C.linestring_wkb([(0, 0), (10, 0)])
    assert _dw(line, C.point_wkb(5, 0.05), 0.05)
    assert not _dw(line, C.point_wkb(5, 0.0500001), 0.05)
    # on the line: exact zero (robust orientation), also for d = 0
    assert _dw(C.linestring_wkb([(0, 0), (3, 1)]), C.point_wkb(1.5, 0.5), 0.0)


def _exact_seg_dist2(px, py, sx, sy, ex, ey):
    P, S, E = [Fraction(v) for v in (px, py)], [Fraction(v) for v in (sx, sy)], [Fraction(v) for v in (ex, ey)]
    dx, dy = E[0] - S[0], E[1] - S[1]
    d2 = dx * dx + dy * dy
    if d2 == 0:
        return (P[0] - S[0]) ** 2 + (P[1] - S[1]) ** 2
    r = ((P[0] - S[0]) * dx + (P[1] - S[1]) * dy) / d2
    r = min(max(r, Fraction(0)), Fraction(1))
    qx, qy = S[0] + r * dx, S[1] + r * dy
    return (P[0] - qx) ** 2 + (P[1] - qy) ** 2


def test_distance_against_exact_rationals():
    """Point-to-linestring distance within 1e-12 relative of the exact
    distance (closed form on rationals)."""
    rng = np.random.default_rng(5)
    for _ in range(300):
        pts = [tuple(rng.uniform(0, 10, 2)) for _ in range(int(rng.integers(2, 6)))]
        p = tuple(rng.uniform(0, 10, 2))
        A, B = O.Store(C.column([C.linestring_wkb(pts)])), O.Store(C.column([C.point_wkb(*p)]))
        got = O.distance(A, 0, B, 0)
        ex2 = min(_exact_seg_dist2(p[0], p[1], *pts[i], *pts[i + 1]) for i in range(len(pts) - 1))
        ex = float(ex2) ** 0.5
        assert abs(got - ex) <= 1e-12 * max(ex, 1e-300)


def _exact_ring_position(ring, px, py):
    """2 on the boundary, 1 inside, 0 outside, in exact rational arithmetic
    (crossing parity of the +x ray)."""
    inside = False
    for (ax, ay), (bx, by) in zip(ring, ring[1:]):
        ax, ay, bx, by = (Fraction(v) for v in (ax, ay, bx, by))
        cross = (bx - ax) * (py - ay) - (by - ay) * (px - ax)
        if cross == 0 and min(ax, bx) <= px <= max(ax, bx) and min(ay, by) <= py <= max(ay, by):
            return 2
        if (ay > py) != (by > py) and px < ax + (py - ay) * (bx - ax) / (by - ay):
            inside = not inside
    return 1 if inside else 0


def _exact_polygon_dist2(polys, x, y):
    """Squared distance from (x, y) to a (multi)polygon: 0 on a boundary or
    inside an exterior and no hole, else the least squared distance to any
    ring's segments (holes included)."""
    px, py = Fraction(x), Fraction(y)
    best = None
    for rings in polys:
        st = _exact_ring_position(rings[0], px, py)
        if st == 2:
            return Fraction(0)
        if st == 1:
            hs = [_exact_ring_position(h, px, py) for h in rings[1:]]
            if 2 in hs or 1 not in hs:
                return Fraction(0)
        for ring in rings:
            for (ax, ay), (bx, by) in zip(ring, ring[1:]):
                d = _exact_seg_dist2(x, y, ax, ay, bx, by)
                best = d if best is None or d < best else best
    return best


def _polygon_distance_cases():
    rng = np.random.default_rng(23)
    sq_hole = [[[(0, 0), (8, 0), (8, 8), (0, 8), (0, 0)], [(2, 2), (2, 6), (6, 6), (6, 2), (2, 2)]]]
    two_holes = [[[(0, 0), (12, 0), (12, 6), (0, 6), (0, 0)], [(1, 1), (5, 1), (5, 5), (1, 5), (1, 1)],
                  [(7, 1), (11, 1), (9, 5), (7, 1)]]]
    multi = [[[(0, 0), (4, 0), (4, 4), (0, 4), (0, 0)], [(1, 1), (3, 1), (3, 3), (1, 3), (1, 1)]],
             [[(6, 0), (10, 0), (8, 4), (6, 0)]]]
    n = 24
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    ext = [(5 + r * np.cos(a), 5 + r * np.sin(a)) for r, a in zip(rng.uniform(4, 5, n), ang)]
    hole = [(5 + r * np.cos(a), 5 + r * np.sin(a)) for r, a in zip(rng.uniform(1, 2.5, n), ang)]
    star = [[ext + [ext[0]], hole + [hole[0]]]]
    cases = []
    for polys in (sq_hole, two_holes, multi, star):
        pts = [(1.0, 1.0), (7.0, 4.0), (4.0, 4.0), (3.0, 3.0), (2.0, 4.0), (4.0, 2.0), (2.5, 5.5), (6.0, 6.0),
               (9.0, 2.0), (9.0, 4.5), (3.0, 0.5), (13.0, 3.0), (-1.0, 7.0), (5.0, 5.0), (2.0, 3.0), (8.0, 1.0)]
        for rings in polys:
            for ring in rings:
                for (ax, ay), (bx, by) in zip(ring, ring[1:]):
                    pts.append(((ax + bx) / 2, (ay + by) / 2))      # on an edge (hole edges included)
                    pts.append((ax - 0.3, ay))                       # at a vertex's height
                    pts.append((ax + 0.01, ay + 0.02))
        pts += [tuple(rng.uniform(-2, 14, 2)) for _ in range(150)]
        cases.append((polys, pts))
    return cases


def test_polygon_distance_against_exact_rationals():
    """Point-to-(multi)polygon distance (the DWithin refine over polygon
    build rows, geo's utils.rs:231-271 semantics): exactly 0 inside an
    exterior and outside every hole, and on any boundary; otherwise within
    1e-12 relative of the exact distance to the nearest ring, hole rings
    included.  Points lie deep inside solid parts, inside holes, on hole
    edges and at vertex heights, so dropping the interior rule, the holes
    or the hole rings' segments each fails.  Tolerance: 1e-12 relative
    plus a few ulps of the coordinates absolute."""
    zero = inside_hole = 0
    for polys, pts in _polygon_distance_cases():
        wkb = (C.polygon_wkb(polys[0]) if len(polys) == 1 else C.multipolygon_wkb(polys))
        A = O.Store(C.column([wkb]))
        B = O.Store(C.column([C.point_wkb(float(x), float(y)) for x, y in pts]))
        for i, (x, y) in enumerate(pts):
            got = O.distance(A, 0, B, i)
            ex2 = _exact_polygon_dist2(polys, float(x), float(y))
            if ex2 == 0:
                zero += 1
                assert got == 0.0, (x, y, got)
                continue
            ex = float(ex2) ** 0.5
            if any(_exact_ring_position(rings[0], Fraction(x), Fraction(y)) == 1 for rings in polys):
                inside_hole += 1
            # geo's formula divides a cross product by the segment length,
            # so points a few ulps off an edge carry an absolute error of a
            # few ulps of the coordinates (here below 16)
            assert abs(got - ex) <= 1e-12 * ex + 16 * 2.0**-52 * 16, (x, y, got, ex)
    assert zero > 50 and inside_hole > 20, (zero, inside_hole)


# ---------------------------------------------------------------------------
# KNN (reference ground truth: spatial_join_integration.rs:1373-1414)


def test_knn_matches_brute_force_with_id_ties():
    rng = np.random.default_rng(9)
    obj = np.round(rng.uniform(0, 20, size=(3000, 2)) * 4) / 4  # grid: many exact ties
    q = np.round(rng.uniform(0, 20, size=(400, 2)) * 4) / 4
    from sjhip import datagen as D
    B = O.Store(D.points_wkb(obj[:, 0], obj[:, 1]).as_tuple())
    P = O.Store(D.points_wkb(q[:, 0], q[:, 1]).as_tuple())
    for k in (1, 5, 8):
        ids, d = O.knn(B, P, k)
        dd = np.hypot(q[:, None, 0] - obj[None, :, 0], q[:, None, 1] - obj[None, :, 1])
        order = np.lexsort((np.broadcast_to(np.arange(len(obj)), dd.shape), dd), axis=1)[:, :k]
        assert np.array_equal(ids, order.astype(np.uint32))
        assert np.array_equal(d, np.take_along_axis(dd, order, axis=1))


def test_knn_brute_tie_breakers_match_numpy():
    """include_tie_breakers (default_spatial_index.rs:379-472): the k nearest
    by (distance, id), then every further object within 1e-9 of the k-th
    distance; without it at most k.  Grid points make many exact ties and
    near-ties a few ulps apart."""
    rng = np.random.default_rng(29)
    obj = np.round(rng.uniform(0, 10, size=(2000, 2)) * 4) / 4
    obj[:50] += rng.choice([-1e-12, 0.0, 1e-12], size=(50, 2))  # ties within the tolerance
    q = np.round(rng.uniform(0, 10, size=(300, 2)) * 4) / 4
    from sjhip import datagen as D
    B = O.Store(D.points_wkb(obj[:, 0], obj[:, 1]).as_tuple())
    P = O.Store(D.points_wkb(q[:, 0], q[:, 1]).as_tuple())
    dd = np.hypot(q[:, None, 0] - obj[None, :, 0], q[:, None, 1] - obj[None, :, 1])
    for k in (1, 4):
        for ties in (False, True):
            off, ids, d = O.knn_brute(B, P, k, include_tie_breakers=ties)
            for i in range(len(q)):
                order = np.lexsort((np.arange(len(obj)), dd[i]))
                keep = k
                if ties:
                    thr = dd[i, order[k - 1]] + 1e-9
                    keep = int(np.sum(dd[i] <= thr))
                assert off[i + 1] - off[i] == keep
                assert np.array_equal(ids[off[i]:off[i + 1]], order[:keep].astype(np.uint32))
                assert np.array_equal(d[off[i]:off[i + 1]], dd[i, order[:keep]])
    off, _, _ = O.knn_brute(B, P, 4, include_tie_breakers=True)
    assert np.any(np.diff(off) > 4)


def test_rect_rounding_is_outward_and_tight():
    """bounds.rs:147-162: lo rounds down, hi rounds up, at most one ulp."""
    rng = np.random.default_rng(3)
    for v in rng.uniform(-1e4, 1e4, 2000):
        lo, hi = O.f64_interval_to_f32(v, v)
        assert lo <= v <= hi
        assert np.nextafter(np.float32(lo), np.float32(np.inf)) >= np.float32(v) or lo == v
        assert hi == lo or np.nextafter(np.float32(lo), np.float32(np.inf)) == np.float32(hi)
