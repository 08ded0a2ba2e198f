"""The request loop (SjhipServer*): probe batches of different sizes and
predicates joined concurrently against one resident index give the same
pair sets as each batch joined alone, and admission respects the memory
budget."""
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


@pytest.fixture(scope="module")
def resident():
    """A 2k-polygon Zipf index (vertex counts 4..4096) and one device column
    of clustered probe points that the requests slice."""
    opts = G.GpuSpatialOptions(device_id=0)
    b, _ = D.workload_pip_zipf(1, 2_000, seed=21)
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(b.to_arrow())
    rf.push_build(b.to_arrow())
    rf.finish_building()
    ix = G.GpuSpatialIndex(opts)
    ix.build_from_refiner(rf)
    rng = np.random.default_rng(22)
    extent = 1000.0 * np.sqrt(2_000 / 1_000_000)
    sizes = [100, 5_000, 1, 40_000, 700, 20_000, 3_000, 60_000, 10, 12_000, 250, 30_000]
    xy = D.clustered_xy(rng, int(sum(sizes)), extent)
    pcol = D.points_wkb(xy[:, 0], xy[:, 1])
    dev = G.DeviceWkb.from_host(pcol)
    starts = np.concatenate([[0], np.cumsum(sizes)])
    preds = [("intersects", 0.0), ("dwithin", 0.05), ("contains", 0.0)]
    reqs = []
    for k, n in enumerate(sizes):
        pred, d = preds[k % 3]
        reqs.append(dict(lo=int(starts[k]), hi=int(starts[k + 1]), pred=pred, d=d))
    return dict(ix=ix, rf=rf, b=b, pcol=pcol, dev=dev, reqs=reqs)


def alone(R, q):
    r = G.join_device(R["ix"], R["rf"], R["dev"].slice(q["lo"], q["hi"]), q["pred"], distance=q["d"],
                      sort_output=True)
    out = r.to_host()
    r.release()
    return out


@pytest.mark.parametrize("slots,budget", [(4, 0), (3, 64 << 20)])
def test_requests_in_the_mix_equal_requests_alone(resident, slots, budget):
    R = resident
    want = [alone(R, q) for q in R["reqs"]]
    say("alone done", [len(w[0]) for w in want])
    sv = G.GpuSpatialServer(R["ix"], R["rf"], slots=slots, memory_budget=budget)
    try:
        ids = [sv.submit(R["dev"].slice(q["lo"], q["hi"]), q["pred"], distance=q["d"], sort_output=True)
               for q in R["reqs"]]
        for k, rid in enumerate(ids):
            res, t = sv.wait(rid)
            gb, gp = res.to_host()
            res.release()
            assert t["submit_ms"] <= t["admit_ms"] <= t["done_ms"]
            assert np.array_equal(gb, want[k][0]) and np.array_equal(gp, want[k][1]), (k, len(gb), len(want[k][0]))
    finally:
        sv.close()


def test_mixed_requests_match_oracle(resident):
    R = resident
    sv = G.GpuSpatialServer(R["ix"], R["rf"], slots=4)
    try:
        picks = [1, 3, 5]  # one request of each predicate
        ids = [sv.submit(R["dev"].slice(R["reqs"][k]["lo"], R["reqs"][k]["hi"]), R["reqs"][k]["pred"],
                         distance=R["reqs"][k]["d"], sort_output=True) for k in picks]
        for k, rid in zip(picks, ids):
            q = R["reqs"][k]
            res, _ = sv.wait(rid)
            gb, gp = res.to_host()
            res.release()
            sub = D.take_rows(R["pcol"], np.arange(q["lo"], q["hi"]))
            ob, op = O.join(O.Store(R["b"].as_tuple()), O.Store(sub.as_tuple()), q["pred"], distance=q["d"],
                            nthreads=8)
            assert np.array_equal(gb, ob) and np.array_equal(gp, op), (k, len(gb), len(ob))
    finally:
        sv.close()


def test_budget_below_every_request_runs_them_one_at_a_time(resident):
    """With a 1-byte budget no two requests may hold device memory at once:
    each is admitted only after the previous one finished."""
    R = resident
    sv = G.GpuSpatialServer(R["ix"], R["rf"], slots=4, memory_budget=1)
    try:
        qs = R["reqs"][:6]
        ids = [sv.submit(R["dev"].slice(q["lo"], q["hi"]), q["pred"], distance=q["d"]) for q in qs]
        spans = []
        for rid in ids:
            res, t = sv.wait(rid)
            res.release()
            spans.append((t["admit_ms"], t["done_ms"]))
        spans.sort()
        for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
            assert a1 >= d0, (spans,)
    finally:
        sv.close()


def test_request_error_is_reported_by_wait(resident):
    R = resident
    bad = D.points_wkb(np.array([1.0, 2.0]), np.array([1.0, 2.0]))
    data = bytearray(bad.data)
    data[1 + 21] = 9  # geometry type 9 in row 1
    bad = D.WkbColumn(bad.offsets, np.frombuffer(bytes(data), dtype=np.uint8))
    sv = G.GpuSpatialServer(R["ix"], R["rf"], slots=2)
    try:
        rid = sv.submit(bad.to_arrow(), "intersects")
        with pytest.raises(G.GpuSpatialError, match="row 1"):
            sv.wait(rid)
        # the server keeps serving after a failed request
        q = R["reqs"][1]
        res, _ = sv.wait(sv.submit(R["dev"].slice(q["lo"], q["hi"]), q["pred"], distance=q["d"], sort_output=True))
        gb, gp = res.to_host()
        res.release()
        want = alone(R, q)
        assert np.array_equal(gb, want[0]) and np.array_equal(gp, want[1])
    finally:
        sv.close()


def test_graph_replays_equal_eager_joins(resident):
    """With captured graphs, every replay (three rounds over the same size
    buckets, with eager joins in between) returns each request's pairs as
    the eager join does."""
    R = resident
    want = [alone(R, q) for q in R["reqs"]]
    sv = G.GpuSpatialServer(R["ix"], R["rf"], slots=2, graphs=True)
    try:
        for rnd in range(3):
            ids = [sv.submit(R["dev"].slice(q["lo"], q["hi"]), q["pred"], distance=q["d"], sort_output=True)
                   for q in R["reqs"]]
            for k, rid in enumerate(ids):
                res, _ = sv.wait(rid)
                gb, gp = res.to_host()
                res.release()
                assert np.array_equal(gb, want[k][0]) and np.array_equal(gp, want[k][1]), (rnd, k, len(gb))
            # eager work between the replays
            q = R["reqs"][rnd]
            gb, gp = alone(R, q)
            assert np.array_equal(gb, want[rnd][0]) and np.array_equal(gp, want[rnd][1])
            say("graph round", rnd, "ok")
    finally:
        sv.close()


def test_graph_replays_honour_arrow_offsets(resident):
    """Requests given as Arrow slices (array offset != 0) replay the captured
    sequence over rows [offset, offset + length)."""
    R = resident
    picks = [0, 4, 8, 2]
    want = [alone(R, R["reqs"][k]) for k in picks]
    sv = G.GpuSpatialServer(R["ix"], R["rf"], slots=2, graphs=True)
    try:
        for rnd in range(2):
            ids = [sv.submit(R["dev"].arrow_slice(R["reqs"][k]["lo"], R["reqs"][k]["hi"]), R["reqs"][k]["pred"],
                             distance=R["reqs"][k]["d"], sort_output=True) for k in picks]
            for j, rid in enumerate(ids):
                res, _ = sv.wait(rid)
                gb, gp = res.to_host()
                res.release()
                assert np.array_equal(gb, want[j][0]) and np.array_equal(gp, want[j][1]), (rnd, picks[j], len(gb))
            say("offset round", rnd, "ok")
    finally:
        sv.close()
