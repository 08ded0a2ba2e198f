"""Build sides larger than a device-memory budget (SjhipPartitionedJoin):
K-D-B partitions of the build rows (rust/sedona-spatial-join/src/partitioning/kdb.rs)
built and joined one at a time, the reference's memory-quota batching of
the build side (c/sedona-libgpuspatial/libgpuspatial/src/relate_engine.cu:468-479).
Every pair must come from exactly one partition: the union of the
partitions' pairs equals the oracle's pair set, with no repeats."""
import sys

import numpy as np
import pytest

import oracle as O
from sjhip import datagen as D
from sjhip import gpuspatial as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sorted_pairs(b, p):
    k = (p.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64)
    return np.sort(k)


def _with_gaps(col, every=53):
    """Nulls in every `every`-th build row (they meet nothing)."""
    v = np.ones(len(col), bool)
    v[::every] = False
    return D.WkbColumn(col.offsets, col.data, v)


@pytest.fixture(scope="module")
def zipf():
    b, p = D.workload_pip_zipf(200_000, 2_000, seed=17)
    return _with_gaps(b), p


@pytest.mark.parametrize("budget_mb", [4096, 8, 2])
@pytest.mark.parametrize("pred", ["contains", "intersects"])
def test_partitioned_join_matches_oracle(zipf, budget_mb, pred):
    b, p = zipf
    dev = G.DeviceWkb.from_host(p)
    gb, gp, st = G.partitioned_join(b.to_arrow(), dev, pred, budget_mb << 20)
    say("budget", budget_mb, "MB:", st)
    keys = _sorted_pairs(gb, gp)
    assert len(np.unique(keys)) == len(keys), "a pair came from two partitions"
    ob, op = O.join(O.Store(b.as_tuple()), O.Store(p.as_tuple()), pred, nthreads=8)
    assert np.array_equal(keys, _sorted_pairs(ob, op))
    assert st["pairs"] == len(gb)
    if budget_mb == 4096:
        assert st["partitions"] == 1
    else:
        assert st["partitions"] > 1
        assert st["max_partition_bytes"] <= (budget_mb << 20) or st["max_partition_rows"] == 1


def test_partitioned_join_polygon_probes_and_dwithin():
    """Polygon probes (a probe row may meet build rows of several
    partitions) and ST_DWithin over host probe columns."""
    rng = np.random.default_rng(23)
    n = 600
    centers = rng.uniform(0, 60, size=(n, 2))
    counts = rng.integers(4, 200, size=n)
    build = D.star_polygons(rng, n, counts, centers, np.log(1.5))
    polys = D.star_polygons(rng, 400, rng.integers(4, 60, size=400), rng.uniform(0, 60, size=(400, 2)), np.log(1.0))
    xy = rng.uniform(0, 60, size=(50_000, 2))
    points = D.points_wkb(xy[:, 0], xy[:, 1])
    B = O.Store(build.as_tuple())
    for pred, probe, d in (("intersects", polys, 0.0), ("dwithin", points, 0.3)):
        gb, gp, st = G.partitioned_join(build.to_arrow(), probe.to_arrow(), pred, 1 << 20, distance=d)
        say(pred, st)
        assert st["partitions"] > 1
        keys = _sorted_pairs(gb, gp)
        assert len(np.unique(keys)) == len(keys)
        ob, op = O.join(B, O.Store(probe.as_tuple()), pred, distance=d, nthreads=8)
        assert len(ob) > 0
        assert np.array_equal(keys, _sorted_pairs(ob, op)), (pred, len(keys), len(ob))


def test_partitioned_join_errors_and_stop(zipf):
    b, p = zipf
    with pytest.raises(G.GpuSpatialError, match="at least 1 MiB"):
        G.partitioned_join(b.to_arrow(), p.to_arrow(), "contains", 1000)
    with pytest.raises(G.GpuSpatialError, match="rc=5"):
        G.partitioned_join(b.to_arrow(), p.to_arrow(), "contains", 2 << 20, callback_rc=5)
    # an all-null build side: no partition, one empty delivery
    nb = D.WkbColumn(b.offsets, b.data, np.zeros(len(b), bool))
    gb, gp, st = G.partitioned_join(nb.to_arrow(), p.to_arrow(), "contains", 2 << 20)
    assert len(gb) == 0 and st["partitions"] == 0
