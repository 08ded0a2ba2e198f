"""Adversarial geometry corpus for the relation refine.

Grid-aligned polygons (axis-parallel and diagonal edges, holes,
multipolygons, unclosed rings, collinear vertices) probed by points on
vertices, on edge midpoints, at vertex heights, one ulp off edges and in
general position.  Every branch of the reference's ray-cast classifier
(c/sedona-tg/src/tg/tg.c:1199-1275) and of its collinearity test (861-901)
is reached by some pair.  Deterministic for a given seed.
"""
import struct

import numpy as np


def _ring_bytes(pts):
    return struct.pack("<I", len(pts)) + b"".join(struct.pack("<dd", float(x), float(y)) for x, y in pts)


def polygon_wkb(rings, big_endian=False):
    if big_endian:
        out = b"\x00" + struct.pack(">I", 3) + struct.pack(">I", len(rings))
        for r in rings:
            out += struct.pack(">I", len(r)) + b"".join(struct.pack(">dd", float(x), float(y)) for x, y in r)
        return out
    return b"\x01" + struct.pack("<I", 3) + struct.pack("<I", len(rings)) + b"".join(_ring_bytes(r) for r in rings)


def multipolygon_wkb(polys):
    return b"\x01" + struct.pack("<I", 6) + struct.pack("<I", len(polys)) + b"".join(polygon_wkb(p) for p in polys)


def point_wkb(x, y, big_endian=False):
    if big_endian:
        return b"\x00" + struct.pack(">I", 1) + struct.pack(">dd", x, y)
    return b"\x01" + struct.pack("<I", 1) + struct.pack("<dd", x, y)


def linestring_wkb(pts):
    return b"\x01" + struct.pack("<I", 2) + _ring_bytes(pts)


def column(wkbs):
    off = np.zeros(len(wkbs) + 1, np.int64)
    off[1:] = np.cumsum([len(w) for w in wkbs])
    data = np.frombuffer(b"".join(wkbs), dtype=np.uint8).copy() if wkbs else np.zeros(0, np.uint8)
    return off, data, None


def build_polygons(seed=11):
    rng = np.random.default_rng(seed)
    polys = []
    # axis-aligned squares and rectangles
    polys.append(polygon_wkb([[(0, 0), (4, 0), (4, 4), (0, 4), (0, 0)]]))
    polys.append(polygon_wkb([[(1, 1), (7, 1), (7, 3), (1, 3), (1, 1)]]))
    # square with a hole
    polys.append(polygon_wkb([[(0, 0), (8, 0), (8, 8), (0, 8), (0, 0)],
                              [(2, 2), (2, 6), (6, 6), (6, 2), (2, 2)]]))
    # diamond (diagonal edges)
    polys.append(polygon_wkb([[(4, 0), (8, 4), (4, 8), (0, 4), (4, 0)]]))
    # concave L with collinear vertices and horizontal runs
    polys.append(polygon_wkb([[(0, 0), (3, 0), (6, 0), (6, 2), (2, 2), (2, 6), (0, 6), (0, 3), (0, 0)]]))
    # triangle with vertex at the probes' y levels
    polys.append(polygon_wkb([[(1, 1), (7, 4), (1, 7), (1, 1)]]))
    # big-endian polygon
    polys.append(polygon_wkb([[(0, 0), (5, 1), (6, 6), (1, 5), (0, 0)]], big_endian=True))
    # multipolygon: two squares sharing an edge and one apart
    polys.append(multipolygon_wkb([[[(0, 0), (2, 0), (2, 2), (0, 2), (0, 0)]],
                                   [[(2, 0), (4, 0), (4, 2), (2, 2), (2, 0)]],
                                   [[(5, 5), (7, 5), (7, 7), (5, 7), (5, 5)],
                                    [(5.5, 5.5), (6.5, 5.5), (6.5, 6.5), (5.5, 6.5), (5.5, 5.5)]]]))
    # spiky star with vertices on a 0.5 grid
    for _ in range(6):
        n = int(rng.integers(5, 14))
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = rng.integers(2, 9, n) * 0.5
        cx, cy = rng.integers(2, 7, 2)
        pts = [(cx + np.round(ri * np.cos(a) * 2) / 2, cy + np.round(ri * np.sin(a) * 2) / 2) for ri, a in zip(r, ang)]
        pts.append(pts[0])
        polys.append(polygon_wkb([pts]))
    # general-position star with irrational coordinates
    for _ in range(6):
        n = int(rng.integers(6, 40))
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = rng.uniform(1.5, 4.0, n)
        cx, cy = rng.uniform(2, 6, 2)
        pts = [(cx + ri * np.cos(a), cy + ri * np.sin(a)) for ri, a in zip(r, ang)]
        pts.append(pts[0])
        polys.append(polygon_wkb([pts]))
    # degenerate: empty polygon, polygon with an empty ring
    polys.append(b"\x01" + struct.pack("<I", 3) + struct.pack("<I", 0))
    polys.append(polygon_wkb([[]]))
    return polys


def invalid_wkbs():
    """Rows the reference refuses to parse (tg.c:8380-8399) or that are not
    2-D WKB: each must raise an error naming the row."""
    return [
        polygon_wkb([[(1, 1), (7, 1), (7, 3), (1, 3)]]),  # ring not closed
        polygon_wkb([[(1, 1), (2, 2)]]),  # ring with two positions
        linestring_wkb([(1, 1)]),  # line with one position
        b"\x01" + struct.pack("<I", 1001) + struct.pack("<ddd", 1, 2, 3),  # Z point
        b"\x01" + struct.pack("<I", 3),  # truncated
        b"\x02" + struct.pack("<I", 1) + struct.pack("<dd", 1, 2),  # bad byte order
    ]


def probe_points(build_wkbs, seed=12, n_random=600):
    """Points on a 0.25 grid, on every vertex and edge midpoint of the build
    polygons, one ulp either side of those, and random points."""
    rng = np.random.default_rng(seed)
    pts = []
    g = np.arange(-0.5, 8.75, 0.25)
    for x in g[::2]:
        for y in g:
            pts.append((float(x), float(y)))
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "..", "oracle"))
    import oracle as O
    st = O.Store(column(build_wkbs)).soa()
    xy = st["xy"]
    vo = st["pt_off"]
    for r in range(len(vo) - 1):
        ring = xy[vo[r]:vo[r + 1]]
        for i in range(len(ring)):
            a = ring[i]
            b = ring[(i + 1) % len(ring)]
            m = (a + b) / 2
            for p in (a, m):
                px, py = float(p[0]), float(p[1])
                pts.append((px, py))
                pts.append((np.nextafter(px, np.inf), py))
                pts.append((px, np.nextafter(py, -np.inf)))
                pts.append((np.nextafter(px, -np.inf), np.nextafter(py, np.inf)))
            # a point at the vertex's height well left and right of it
            pts.append((float(a[0]) - 0.37, float(a[1])))
            pts.append((float(a[0]) + 0.37, float(a[1])))
    for _ in range(n_random):
        pts.append((float(rng.uniform(-0.5, 8.5)), float(rng.uniform(-0.5, 8.5))))
    wk = [point_wkb(x, y, big_endian=(i % 7 == 3)) for i, (x, y) in enumerate(pts)]
    wk.append(point_wkb(float("nan"), float("nan")))  # empty point
    return wk


def corpus(seed=11):
    b = build_polygons(seed)
    p = probe_points(b, seed + 1)
    return column(b), column(p)


def _star(rng, cx, cy, n, rmin, rmax, grid=0.5, ccw=True):
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    r = rng.uniform(rmin, rmax, n)
    pts = [(cx + np.round(ri * np.cos(a) / grid) * grid, cy + np.round(ri * np.sin(a) / grid) * grid)
           for ri, a in zip(r, ang)]
    pts = [(float(x), float(y)) for x, y in pts]
    dedup = [pts[0]]
    for q in pts[1:]:
        if q != dedup[-1]:
            dedup.append(q)
    if not ccw:
        dedup = dedup[::-1]
    dedup.append(dedup[0])
    return dedup


def _square(x0, y0, x1, y1, ccw=True):
    pts = [(x0, y0), (x1, y0), (x1, y1), (x0, y1), (x0, y0)]
    return pts if ccw else pts[::-1]


def _pentagram(cx, cy, r, ccw=True):
    """The star polygon {5/2}: the five points of a circle joined every
    second one (a self-crossing ring)."""
    pts = []
    for k in range(5):
        a = np.pi / 2 + (2 if ccw else -2) * k * 2 * np.pi / 5
        pts.append((cx + r * np.cos(a), cy + r * np.sin(a)))
    return pts + [pts[0]]


def hole_polygons(seed=17):
    """Polygons with holes over [0, 16]^2 (vertices on a 0.5 grid): convex
    holes in both orientations, concave holes (L shape, stars), a hole whose
    first three points are collinear (TG then treats the convex hole as
    concave), two holes, holes inside multipolygon parts, and small
    polygons without holes."""
    rng = np.random.default_rng(seed)
    polys = []
    outer = _square(0, 0, 16, 16)
    polys.append(polygon_wkb([outer, _square(4, 4, 12, 12)]))
    polys.append(polygon_wkb([outer, _square(4, 4, 12, 12, ccw=False)]))
    polys.append(polygon_wkb([outer, [(4, 4), (12, 4), (12, 8), (8, 8), (8, 12), (4, 12), (4, 4)]]))
    polys.append(polygon_wkb([outer, [(4, 4), (4, 8), (4, 12), (12, 12), (12, 4), (4, 4)]]))
    polys.append(polygon_wkb([outer, [(4, 4), (8, 4), (12, 4), (12, 12), (4, 12), (4, 4)]]))
    polys.append(polygon_wkb([outer, [(8, 3), (13, 8), (8, 13), (3, 8), (8, 3)]]))
    polys.append(polygon_wkb([outer, _square(2, 2, 7, 7), [(9, 9), (14, 9), (14, 14), (11.5, 11.5), (9, 14), (9, 9)]]))
    for k in range(6):
        polys.append(polygon_wkb([outer, _star(rng, 8, 8, int(rng.integers(5, 12)), 2.0, 6.5, ccw=bool(k % 2))]))
    polys.append(multipolygon_wkb([[_square(0, 0, 7, 7), _square(2, 2, 5, 5)],
                                   [_square(9, 9, 16, 16), [(10, 10), (15, 10), (12.5, 12.5), (15, 15), (10, 15), (10, 10)]]]))
    for _ in range(8):
        cx, cy = (float(v) for v in rng.integers(2, 15, 2))
        polys.append(polygon_wkb([_star(rng, cx, cy, int(rng.integers(3, 9)), 0.5, 3.0)]))
    # pentagram holes: every turn has the same sign, so TG flags them convex
    # (vertex test only) although their edges cross; a bowtie hole (concave)
    for ccw in (True, False):
        polys.append(polygon_wkb([outer, _pentagram(8.0, 8.0, 6.0, ccw)]))
    polys.append(polygon_wkb([outer, [(4, 4), (12, 12), (12, 4), (4, 12), (4, 4)]]))
    # general position: an irrational star exterior with a star hole
    for _ in range(4):
        cx, cy = rng.uniform(6, 10, 2)
        n = int(rng.integers(8, 30))
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        ext = [(cx + r * np.cos(a), cy + r * np.sin(a)) for r, a in zip(rng.uniform(6, 8, n), ang)]
        hole = [(cx + r * np.cos(a), cy + r * np.sin(a)) for r, a in zip(rng.uniform(1.5, 4.5, n), ang)]
        polys.append(polygon_wkb([ext + [ext[0]], hole + [hole[0]]]))
    return polys


def hole_probe_polygons(seed=18, n_random=260):
    """Small polygons around the holes of hole_polygons(): inside, touching
    at a vertex or along an edge, equal to a hole, crossing a hole's edge
    or notch, covering a hole; polygons with holes of their own; random
    grid stars."""
    rng = np.random.default_rng(seed)
    polys = []
    polys.append(polygon_wkb([_square(5, 5, 11, 11)]))
    polys.append(polygon_wkb([_square(4, 4, 12, 12)]))
    polys.append(polygon_wkb([_square(4, 4, 12, 12, ccw=False)]))
    polys.append(polygon_wkb([_square(4, 5, 6, 7)]))
    polys.append(polygon_wkb([_square(3, 5, 6, 7)]))
    polys.append(polygon_wkb([_square(3.5, 3.5, 12.5, 12.5)]))
    polys.append(polygon_wkb([[(5, 5), (11, 5), (5, 11), (5, 5)]]))
    polys.append(polygon_wkb([[(5, 5), (11, 6), (6, 11), (5, 5)]]))
    polys.append(polygon_wkb([[(9, 9), (11, 9), (11, 11), (9, 11), (9, 9)]]))
    polys.append(polygon_wkb([[(5, 9), (9, 5), (7, 11), (5, 9)]]))
    polys.append(polygon_wkb([[(8, 4), (12, 8), (8, 12), (4, 8), (8, 4)]]))
    polys.append(polygon_wkb([_square(-1, -1, 17, 17), _square(-0.5, -0.5, 16.5, 16.5)]))
    polys.append(polygon_wkb([_square(-1, -1, 17, 17), _square(1, 1, 15, 15)]))
    polys.append(polygon_wkb([_square(1, 1, 15, 15), _square(2, 2, 14, 14)]))
    polys.append(polygon_wkb([_square(6, 6, 10, 10), _square(7, 7, 9, 9)]))
    # triangles and quads joining points in the tips of the pentagram holes
    # (each tip has crossing parity 1, the centre 0): every vertex is inside
    # the hole while the edges cross it
    tips = [(8.0 + 4.8 * np.cos(np.pi / 2 + k * 2 * np.pi / 5), 8.0 + 4.8 * np.sin(np.pi / 2 + k * 2 * np.pi / 5))
            for k in range(5)]
    for ks in ((0, 2, 3), (0, 1, 2), (1, 3, 4), (0, 1, 2, 3), (4, 2, 0)):
        ring = [tips[k] for k in ks]
        polys.append(polygon_wkb([ring + [ring[0]]]))
    small = [(tips[0][0] + dx, tips[0][1] + dy) for dx, dy in ((-0.2, -0.3), (0.2, -0.3), (0.0, 0.2))]
    polys.append(polygon_wkb([small + [small[0]]]))
    for _ in range(n_random):
        cx, cy = (float(v) for v in rng.integers(0, 33, 2) * 0.5)
        polys.append(polygon_wkb([_star(rng, cx, cy, int(rng.integers(3, 8)), 0.4, 3.5,
                                        ccw=bool(rng.integers(0, 2)))]))
    for _ in range(20):
        cx, cy = rng.uniform(1, 15, 2)
        n = int(rng.integers(3, 12))
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = rng.uniform(0.3, 3.0, n)
        pts = [(cx + ri * np.cos(a), cy + ri * np.sin(a)) for ri, a in zip(r, ang)]
        polys.append(polygon_wkb([pts + [pts[0]]]))
    return polys


def hole_corpus(seed=17):
    return column(hole_polygons(seed)), column(hole_probe_polygons(seed + 1))


def relate_polygons(seed=19, n_random=90):
    """(Multi)polygons on a 1.0 grid over [0, 12]^2 for the polygon x polygon
    relations: tiles sharing edges and corners, nested squares (touching the
    outer boundary from inside or not), equal rings in both orientations
    and with a rotated start, squares with holes holding other squares,
    multipolygons of tiles, L shapes, and random grid stars."""
    rng = np.random.default_rng(seed)
    polys = []
    for x0, y0 in ((0, 0), (2, 0), (0, 2), (2, 2), (4, 4), (1, 1)):
        polys.append(polygon_wkb([_square(x0, y0, x0 + 2, y0 + 2)]))
    polys.append(polygon_wkb([_square(0, 0, 4, 4)]))
    polys.append(polygon_wkb([_square(0, 0, 4, 4, ccw=False)]))
    polys.append(polygon_wkb([[(4, 0), (4, 4), (0, 4), (0, 0), (4, 0)]]))
    polys.append(polygon_wkb([[(0, 0), (2, 0), (4, 0), (4, 4), (0, 4), (0, 0)]]))
    polys.append(polygon_wkb([_square(1, 1, 3, 3)]))
    polys.append(polygon_wkb([_square(0, 1, 2, 3)]))
    polys.append(polygon_wkb([_square(0, 0, 4, 4), _square(1, 1, 3, 3)]))
    polys.append(polygon_wkb([_square(0, 0, 4, 4), _square(1, 1, 2, 2)]))
    polys.append(polygon_wkb([_square(-1, -1, 5, 5), _square(0, 0, 4, 4)]))
    polys.append(polygon_wkb([[(0, 0), (4, 0), (4, 2), (2, 2), (2, 4), (0, 4), (0, 0)]]))
    polys.append(polygon_wkb([[(2, 2), (4, 2), (4, 4), (2, 4), (2, 2)]]))
    polys.append(polygon_wkb([[(0, 0), (4, 0), (2, 2), (0, 0)]]))
    polys.append(polygon_wkb([[(0, 0), (2, 2), (0, 4), (0, 0)]]))
    polys.append(polygon_wkb([[(2, 2), (6, 2), (6, 6), (2, 6), (2, 2)]]))
    # triangles with every vertex on or inside the L above and one edge
    # across its notch (endpoints on two different edges of the L)
    polys.append(polygon_wkb([[(3, 2), (2, 3), (1, 1), (3, 2)]]))
    polys.append(polygon_wkb([[(3, 2), (1, 1), (2, 3), (3, 2)]]))
    polys.append(polygon_wkb([[(0, 0), (0, 4), (2, 4), (2, 2), (4, 2), (4, 0), (0, 0)]]))
    polys.append(multipolygon_wkb([[_square(0, 0, 2, 2)], [_square(2, 2, 4, 4)]]))
    polys.append(multipolygon_wkb([[_square(0, 0, 2, 2)], [_square(2, 0, 4, 2)]]))
    polys.append(multipolygon_wkb([[_square(0, 0, 2, 2)], [_square(6, 6, 8, 8)]]))
    polys.append(multipolygon_wkb([[_square(0, 0, 4, 4), _square(1, 1, 3, 3)], [_square(1, 1, 3, 3)]]))
    polys.append(multipolygon_wkb([[_square(0, 0, 2, 2)]]))
    polys.append(multipolygon_wkb([[_square(4, 4, 6, 6)], [_square(0, 0, 1, 1)], [_square(8, 0, 9, 1)]]))
    # concave stars and polygons whose vertices sit on the star's vertices
    # and edge midpoints (segments from one boundary edge to another, inside
    # or across the star)
    for _ in range(12):
        cx, cy = (float(v) for v in rng.integers(3, 10, 2))
        star = _star(rng, cx, cy, int(rng.integers(5, 10)), 1.0, 4.0, grid=1.0, ccw=bool(rng.integers(0, 2)))
        polys.append(polygon_wkb([star]))
        cand = star[:-1] + [((a[0] + b[0]) / 2, (a[1] + b[1]) / 2) for a, b in zip(star, star[1:])]
        for _ in range(3):
            k = rng.choice(len(cand), size=int(rng.integers(3, 5)), replace=False)
            ring = [cand[i] for i in k]
            polys.append(polygon_wkb([ring + [ring[0]]]))
    for _ in range(n_random):
        cx, cy = (float(v) for v in rng.integers(0, 13, 2))
        kind = rng.integers(0, 4)
        if kind == 0:
            w, h = (float(v) for v in rng.integers(1, 5, 2))
            polys.append(polygon_wkb([_square(cx, cy, cx + w, cy + h, ccw=bool(rng.integers(0, 2)))]))
        elif kind == 1:
            polys.append(polygon_wkb([_star(rng, cx, cy, int(rng.integers(3, 9)), 1.0, 4.0, grid=1.0,
                                            ccw=bool(rng.integers(0, 2)))]))
        elif kind == 2:
            w = float(rng.integers(3, 6))
            polys.append(polygon_wkb([_square(cx, cy, cx + w, cy + w), _square(cx + 1, cy + 1, cx + w - 1, cy + w - 1)]))
        else:
            parts = []
            for _ in range(int(rng.integers(2, 4))):
                x, y = (float(v) for v in rng.integers(0, 12, 2))
                parts.append([_square(x, y, x + 1, y + 1)])
            polys.append(multipolygon_wkb(parts))
    return polys


def multilinestring_wkb(lines):
    return b"\x01" + struct.pack("<I", 5) + struct.pack("<I", len(lines)) + b"".join(linestring_wkb(l) for l in lines)


def relate_lines(seed=20, n_random=80):
    """(Multi)linestrings on a 0.5 grid over [0, 12]^2: along polygon
    edges, through vertices, inside and across holes, ending on a
    boundary, closed, zero-length segments, and random grid walks."""
    rng = np.random.default_rng(seed)
    lines = [
        [(0, 0), (4, 0)], [(1, 0), (3, 0)], [(-1, 0), (5, 0)], [(0, 0), (4, 4)], [(2, -1), (2, 5)],
        [(1.5, 1.5), (2.5, 2.5)], [(1.5, 1.5), (1.5, 2.5), (2.5, 2.5)], [(5, 5), (6, 6)],
        [(4, 4), (5, 6)], [(0, 4), (4, 4), (4, 0)], [(1, 1), (3, 1), (3, 3), (1, 3), (1, 1)],
        [(0.5, 0.5), (0.5, 0.5), (3.5, 0.5)], [(2, 2), (2, 2)], [(6, 0), (8, 0), (8, 2)],
        [(-2, 2), (-1, 2)], [(4, 2), (6, 2)], [(3, 2), (5, 2)],
    ]
    out = [linestring_wkb(l) for l in lines]
    out.append(multilinestring_wkb([[(0, 0), (1, 0)], [(10, 10), (11, 11)]]))
    out.append(multilinestring_wkb([[(-3, -3), (-2, -2)], [(20, 20), (21, 21)]]))
    for _ in range(n_random):
        n = int(rng.integers(2, 7))
        p = rng.integers(0, 25, 2) * 0.5
        pts = [tuple(float(v) for v in p)]
        for _ in range(n - 1):
            p = p + rng.integers(-4, 5, 2) * 0.5
            pts.append(tuple(float(v) for v in p))
        out.append(linestring_wkb(pts))
    return out


def relate_points(seed=21, n=150):
    """Points on a 0.5 grid over [-1, 13]^2 (on vertices, edges and
    interiors of the relate corpora)."""
    rng = np.random.default_rng(seed)
    pts = rng.integers(-2, 27, size=(n, 2)) * 0.5
    return [point_wkb(float(x), float(y)) for x, y in pts]
