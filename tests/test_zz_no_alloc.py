"""SjhipJoinDeviceNoAlloc (the join without allocation) against
SjhipJoinDevice: the same pair sets, on two calls in a row with the second
batch smaller, for point-in-polygon and distance predicates; its limits
give errors, not wrong pairs."""
import sys

import numpy as np
import pytest

from sjhip import datagen as D
from sjhip import gpuspatial as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    opts = G.GpuSpatialOptions(device_id=0)
    b, _ = D.workload_pip_zipf(1, 3_000, seed=31)
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(b.to_arrow())
    rf.push_build(b.to_arrow())
    rf.finish_building()
    ix = G.GpuSpatialIndex(opts)
    ix.build_from_refiner(rf)
    rng = np.random.default_rng(32)
    extent = 1000.0 * np.sqrt(3_000 / 1_000_000)
    xy = D.clustered_xy(rng, 120_000, extent)
    col = D.points_wkb(xy[:, 0], xy[:, 1])
    dev = G.DeviceWkb.from_host(col)
    cap = 4 << 20
    ob, op = G.DeviceBuffer(4 * cap), G.DeviceBuffer(4 * cap)
    yield dict(ix=ix, rf=rf, dev=dev, cap=cap, ob=ob, op=op)
    ob.free()
    op.free()


def _keys(b, p):
    return np.sort((p.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64))


def _eager(S, probe, pred, d):
    r = G.join_device(S["ix"], S["rf"], probe, pred, distance=d, sort_output=False)
    gb, gp = r.to_host()
    r.release()
    return _keys(gb, gp)


def _no_alloc(S, probe, pred, d):
    n = G.join_device_no_alloc(S["ix"], S["rf"], probe, pred, S["ob"], S["op"], S["cap"], distance=d)
    return _keys(S["ob"].download(np.uint32, n), S["op"].download(np.uint32, n))


@pytest.mark.parametrize("pred,d", [("contains", 0.0), ("intersects", 0.0), ("dwithin", 0.05)])
def test_no_alloc_join_equals_join_device(setup, pred, d):
    S = setup
    # two calls in a row on the shared array, the second over fewer rows
    for lo, hi in ((0, 65_536), (70_000, 71_000), (100_000, 100_001)):
        probe = S["dev"].slice(lo, hi)
        want = _eager(S, probe, pred, d)
        got = _no_alloc(S, probe, pred, d)
        say("no-alloc", pred, hi - lo, len(got), len(want))
        assert np.array_equal(got, want), (pred, lo, hi, len(got), len(want))


def test_no_alloc_arrow_offset_slice(setup):
    S = setup
    want = _eager(S, S["dev"].slice(5_000, 9_000), "contains", 0.0)
    got = _no_alloc(S, S["dev"].arrow_slice(5_000, 9_000), "contains", 0.0)
    assert np.array_equal(got, want)


def test_no_alloc_limits_are_errors(setup):
    S = setup
    with pytest.raises(G.GpuSpatialError, match="at most 65536 probe rows"):
        G.join_device_no_alloc(S["ix"], S["rf"], S["dev"].slice(0, 65_537), "contains", S["ob"], S["op"], S["cap"])
    probe = S["dev"].slice(0, 20_000)
    want = _eager(S, probe, "contains", 0.0)
    assert len(want) > 10
    with pytest.raises(G.GpuSpatialError, match="output capacity") as e:
        G.join_device_no_alloc(S["ix"], S["rf"], probe, "contains", S["ob"], S["op"], 10)
    assert e.value.pairs == len(want)
    # the array still serves after a refused call
    assert np.array_equal(_no_alloc(S, probe, "contains", 0.0), want)
