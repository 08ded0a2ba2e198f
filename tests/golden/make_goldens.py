"""Mints the golden vectors in tests/golden/expected.npz with the
reference's own relation-refine library (tg.c compiled from the reference
tree into oracle/_ref/ by `make -C oracle ref`).

Inputs (all committed as data):
  data/*.parquet  -- the reference's GPU test fixtures
                     (c/sedona-libgpuspatial/libgpuspatial/test/data/:
                     synthetic_pip/{points,polygons}.parquet,
                     synthetic_poly/polygons{1,2}.parquet (polygons1 is
                     byte-identical to synthetic_pip/polygons.parquet),
                     countries/{natural-earth_countries_geo,
                     generated_points}.parquet)
  corpus.py       -- the adversarial corpus, the polygon-with-holes corpus
                     and the relate corpus of polygons, lines and points
                     (seeded)
  a small uniform star-polygon workload (seeded, sedona-db_amd/sjhip/datagen.py)

Outputs: for every (build, probe, predicate) the refined (build, probe)
pair set sorted by (probe, build), computed with the reference's filter
semantics (FP32 rect intersection) and TG's predicate, for each TG index
mode (they must agree; the script asserts it).

Run from the repository root:  python tests/golden/make_goldens.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "sedona-db_amd"))
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
import corpus as C  # noqa: E402
from sjhip import datagen as D  # noqa: E402

FIXTURE_JOINS = [
    # (name, build file, probe file, predicates)
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


def read_geom(name):
    import pyarrow.parquet as pq
    return pq.read_table(os.path.join(HERE, "data", name)).column("geometry").combine_chunks()


def ref_pairs(build, probe, pred):
    st_b, st_p = O.Store(build), O.Store(probe)
    br = st_b.rects(point_nearest=True)
    pr = st_p.rects(point_nearest=True)
    out = None
    for ix in (0, 1, 2):
        tb, tp = O.TgColumn(build, index=ix), O.TgColumn(probe, index=ix)
        b, p = O.ref_join(tb, tp, br, pr, pred)
        if out is None:
            out = (b, p)
        else:
            assert np.array_equal(out[0], b) and np.array_equal(out[1], p), "TG index modes disagree"
    return out


def main():
    res = {}
    for name, bf, pf, preds in FIXTURE_JOINS:
        b, p = read_geom(bf), read_geom(pf)
        for pred in preds:
            ob, op = ref_pairs(b, p, pred)
            res[f"{name}/{pred}/build"] = ob
            res[f"{name}/{pred}/probe"] = op
            print(name, pred, len(ob))
    # corpus: every pair, both roles
    cb, cp = C.corpus()
    nb, np_ = len(cb[0]) - 1, len(cp[0]) - 1
    bi = np.repeat(np.arange(nb), np_).astype(np.uint32)
    pi = np.tile(np.arange(np_), nb).astype(np.uint32)
    tb, tp = O.TgColumn(cb, index=1), O.TgColumn(cp, index=0)
    for pred in ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"]:
        m = O.ref_refine(tb, tp, pred, bi, pi)
        res[f"corpus/{pred}/mask"] = np.packbits(m)
        m2 = O.ref_refine(tp, tb, pred, pi, bi)
        res[f"corpus_rev/{pred}/mask"] = np.packbits(m2)
        print("corpus", pred, int(m.sum()), int(m2.sum()))
    bb = np.repeat(np.arange(nb), nb).astype(np.uint32)
    pp = np.tile(np.arange(nb), nb).astype(np.uint32)
    res["corpus_polypoly/intersects/mask"] = np.packbits(O.ref_refine(tb, tb, "intersects", bb, pp))
    # polygons with holes x small polygons, every pair, both roles
    hb, hp = C.hole_corpus()
    nhb, nhp = len(hb[0]) - 1, len(hp[0]) - 1
    hbi = np.repeat(np.arange(nhb), nhp).astype(np.uint32)
    hpi = np.tile(np.arange(nhp), nhb).astype(np.uint32)
    for ix in (0, 1, 2):
        thb, thp = O.TgColumn(hb, index=ix), O.TgColumn(hp, index=ix)
        m = O.ref_refine(thb, thp, "intersects", hbi, hpi)
        m2 = O.ref_refine(thp, thb, "intersects", hpi, hbi)
        if ix == 0:
            res["holes_polypoly/intersects/mask"] = np.packbits(m)
            res["holes_polypoly_rev/intersects/mask"] = np.packbits(m2)
        else:
            assert np.array_equal(np.packbits(m), res["holes_polypoly/intersects/mask"]), "TG index modes disagree"
            assert np.array_equal(np.packbits(m2), res["holes_polypoly_rev/intersects/mask"])
    print("holes_polypoly", int(m.sum()), int(m2.sum()))
    # relate corpus: (multi)polygons x (multi)polygons for every predicate,
    # and intersects between (multi)linestrings and polygons, lines, points
    cols = {"p": C.column(C.relate_polygons()), "l": C.column(C.relate_lines()), "q": C.column(C.relate_points())}
    cases = [("pp", "p", "p", ["intersects", "contains", "covers", "within", "coveredby", "touches", "equals"])]
    cases += [(a + b, a, b, ["intersects"]) for a, b in (("l", "p"), ("p", "l"), ("l", "l"), ("q", "l"), ("l", "q"))]
    for name, a, b, preds in cases:
        na, nb = len(cols[a][0]) - 1, len(cols[b][0]) - 1
        ai = np.repeat(np.arange(na), nb).astype(np.uint32)
        bi_ = np.tile(np.arange(nb), na).astype(np.uint32)
        for pred in preds:
            key = f"relate_{name}/{pred}/mask"
            for ix in (0, 1, 2):
                ta, tb_ = O.TgColumn(cols[a], index=ix), O.TgColumn(cols[b], index=ix)
                m = np.packbits(O.ref_refine(ta, tb_, pred, ai, bi_))
                if ix == 0:
                    res[key] = m
                else:
                    assert np.array_equal(m, res[key]), ("TG index modes disagree", key)
            print(key, int(np.unpackbits(res[key]).sum()))
    # small uniform star workload (1/50 of the 1M x 10k case, same densities)
    build, probe = D.workload_pip_uniform(20_000, 200, seed=7, extent=100.0 * np.sqrt(200 / 10_000))
    ob, op = ref_pairs(build.as_tuple(), probe.as_tuple(), "intersects")
    res["uniform_small/intersects/build"] = ob
    res["uniform_small/intersects/probe"] = op
    print("uniform_small", len(ob))
    np.savez_compressed(os.path.join(HERE, "expected.npz"), **res)


if __name__ == "__main__":
    main()
