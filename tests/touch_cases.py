"""Shared inputs of the touches tests (oracle pinning and GPU parity)."""
import corpus as C


def touch_cases():
    """Polygons and multipolygons (one with overlapping parts, one with a
    hole) and points on their vertices and edges, on a hole's boundary,
    inside, and on one part's boundary inside another part."""
    sq = lambda x0, y0, w: [[x0, y0], [x0 + w, y0], [x0 + w, y0 + w], [x0, y0 + w], [x0, y0]]
    polys = [
        C.polygon_wkb([sq(0, 0, 4)]),
        C.polygon_wkb([sq(10, 0, 4), sq(11, 1, 2)]),  # with a hole
        C.multipolygon_wkb([[sq(20, 0, 4)], [sq(22, 2, 4)]]),  # overlapping parts
        C.multipolygon_wkb([[sq(30, 0, 2)], [sq(32, 0, 2)]]),  # parts sharing an edge
    ]
    pts = [(0, 0), (2, 0), (4, 4), (2, 2), (5, 5), (10, 2), (11, 2), (12, 1), (12, 2), (11.5, 1.5),
           (20, 2), (22, 3), (24, 3), (26, 4), (23, 3), (32, 1), (31, 1), (34, 2), (30, 0), (33, 2)]
    return C.column(polys), C.column([C.point_wkb(x, y) for x, y in pts])
