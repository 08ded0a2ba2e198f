"""The operator's concurrency contract at the C ABI: `probe` and `refine`
are called from several threads at once, one index context per thread,
against one index and one refiner
(c/sedona-libgpuspatial/libgpuspatial/src/gpuspatial_c.cc:191-214;
c/sedona-libgpuspatial/libgpuspatial/include/gpuspatial/index/spatial_index.hpp:51-53).
Each thread runs the whole batched join of the reference's flow (rects,
probe, refine per batch of rows) and its pairs are checked against the
oracle.  Also the pipelined join of a host column (SjhipJoinHostStream),
whose slots are threads of the library."""
import sys
import threading

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
def opts():
    return G.GpuSpatialOptions(concurrency=4, device_id=0)


@pytest.fixture(scope="module")
def workload(opts):
    b, p = D.workload_pip_zipf(200_000, 2_000, seed=13)
    barr = b.to_arrow()
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema(barr)
    rf.push_build(barr)
    rf.finish_building()
    ix = G.GpuSpatialIndex(opts)
    ix.push_build(G.compute_rects(barr, point_nearest=True))
    ix.finish_building()
    return b, p, ix, rf


def batched_join(ix, rf, parr, pred, batch):
    """The operator's flow over batches of probe rows; probe indices global."""
    outs_b, outs_p = [], []
    for a in range(0, len(parr), batch):
        part = parr.slice(a, min(batch, len(parr) - a))
        cb, cp = ix.probe(G.compute_rects(part, point_nearest=True))
        kb, kp = rf.refine(part, pred, cb, cp)
        outs_b.append(kb)
        outs_p.append(kp.astype(np.uint64) + a)
    b = np.concatenate(outs_b)
    p = np.concatenate(outs_p)
    order = np.lexsort((b, p))
    return b[order].astype(np.uint32), p[order].astype(np.uint32)


def test_probe_refine_from_four_threads_match_oracle(workload):
    b, p, ix, rf = workload
    parr = p.to_arrow()
    B, P = O.Store(b.as_tuple()), O.Store(p.as_tuple())
    # (predicate, batch rows) per thread: two threads share a predicate with
    # different batch sizes
    plan = [("contains", 25_000), ("intersects", 40_000), ("covers", 30_000), ("contains", 60_000)]
    want = {pred: O.join(B, P, pred, nthreads=8) for pred in {q for q, _ in plan}}
    results = [None] * len(plan)
    errors = []
    start = threading.Barrier(len(plan))

    def work(k):
        try:
            start.wait()
            results[k] = batched_join(ix, rf, parr, plan[k][0], plan[k][1])
        except BaseException as e:  # re-raised below
            errors.append((k, e))

    say("threads", len(plan), "points", len(parr))
    th = [threading.Thread(target=work, args=(k,)) for k in range(len(plan))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k, (pred, batch) in enumerate(plan):
        gb, gp = results[k]
        wb, wp = want[pred]
        say("  thread", k, pred, "batch", batch, "pairs", len(gb))
        assert len(gb) > 0
        assert np.array_equal(gb, wb) and np.array_equal(gp, wp), (k, pred, len(gb), len(wb))


def test_refine_rejects_out_of_range_indices(workload):
    """Candidate indices are checked on the device before any refine kernel
    reads with them: the first bad position is reported, its build index
    before its probe index."""
    b, p, ix, rf = workload
    part = p.to_arrow().slice(0, 1000)
    nb = len(b)
    cb = np.array([0, 1, nb + 5, 2], np.uint32)
    cp = np.array([0, 1000, 3, 2], np.uint32)
    with pytest.raises(G.GpuSpatialError, match="probe index out of range: 1000"):
        rf.refine(part, "contains", cb, cp)
    cp[1] = 1
    with pytest.raises(G.GpuSpatialError, match=f"build index out of range: {nb + 5}"):
        rf.refine(part, "contains", cb, cp)
    cb[2] = 3
    kb, kp = rf.refine(part, "contains", cb, cp)
    assert len(kb) == len(kp) <= 4


def test_probe_contexts_keep_their_own_errors(workload):
    """A failing probe (its callback's error code) in one thread leaves the
    other threads' probes intact."""
    b, p, ix, rf = workload
    rects = G.compute_rects(p.to_arrow().slice(0, 20_000), point_nearest=True)
    ref_b, ref_p = ix.probe(rects)
    out = {}

    def good(k):
        out[k] = ix.probe(rects)

    def bad():
        try:
            ix.probe(rects, callback_rc=7)
            out["bad"] = None
        except G.GpuSpatialError as e:
            out["bad"] = str(e)

    th = [threading.Thread(target=good, args=(k,)) for k in range(3)] + [threading.Thread(target=bad)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert out["bad"] is not None and "rc=7" in out["bad"]
    for k in range(3):
        assert np.array_equal(np.sort(out[k][0].astype(np.uint64) << 32 | out[k][1]),
                              np.sort(ref_b.astype(np.uint64) << 32 | ref_p))


def _with_nulls(col, every=97):
    v = np.ones(len(col), bool)
    v[::every] = False
    return D.WkbColumn(col.offsets, col.data, v)


@pytest.mark.parametrize("large", [False, True])
@pytest.mark.parametrize("slots,batch", [(1, 50_000), (2, 30_000), (3, 17_001)])
def test_host_stream_join_matches_device_join(workload, slots, batch, large):
    """SjhipJoinHostStream over a host column with nulls and a nonzero array
    offset (not a multiple of 8) gives the pairs of one SjhipJoinDevice call
    over the same rows, in batch order, probe indices global."""
    b, p, ix, rf = workload
    parr = _with_nulls(p).to_arrow(large=large).slice(5, len(p) - 11)
    say("stream", slots, batch, large, len(parr))
    gb, gp = G.join_host_stream(ix, rf, parr, "contains", batch_rows=batch, slots=slots, sort_output=True)
    r = G.join_device(ix, rf, parr, "contains", sort_output=True)
    wb, wp = r.to_host()
    r.release()
    assert len(gb) > 0
    # batches in order and each sorted: the whole output sorted by (probe, build)
    assert np.array_equal(gb, wb) and np.array_equal(gp, wp)
    B, P = O.Store(b.as_tuple()), O.Store(_with_nulls(p).slice(5, len(p) - 6).as_tuple())
    ob, op = O.join(B, P, "contains", nthreads=8)
    assert np.array_equal(gb, ob) and np.array_equal(gp, op)


def test_host_stream_join_callback_stops_and_errors(workload):
    b, p, ix, rf = workload
    parr = p.to_arrow()
    with pytest.raises(G.GpuSpatialError, match="rc=9"):
        G.join_host_stream(ix, rf, parr, "contains", batch_rows=20_000, slots=2, callback_rc=9)
    with pytest.raises(G.GpuSpatialError, match="slots"):
        G.join_host_stream(ix, rf, parr, "contains", batch_rows=20_000, slots=0)
    with pytest.raises(G.GpuSpatialError, match="batch_rows"):
        G.join_host_stream(ix, rf, parr, "contains", batch_rows=0)
    e_b, e_p = G.join_host_stream(ix, rf, parr.slice(0, 0), "contains")
    assert len(e_b) == len(e_p) == 0
