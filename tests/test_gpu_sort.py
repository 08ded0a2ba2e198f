"""The kernels-only radix sort that captured request sequences replay
(SjhipTestingGraphSortPairs): stable, over any bit range, at tile edges."""
import sys

import numpy as np
import pytest

from sjhip import _lib as L
from sjhip import gpuspatial as G

pytestmark = pytest.mark.gpu


def _sort(keys, vals, begin_bit, end_bit):
    n = len(keys)
    dk = G.DeviceBuffer.upload(keys)
    dv = G.DeviceBuffer.upload(vals)
    ok = G.DeviceBuffer(max(4 * n, 1))
    ov = G.DeviceBuffer(max(4 * n, 1))
    try:
        rc = L.lib().SjhipTestingGraphSortPairs(dk.ptr, dv.ptr, ok.ptr, ov.ptr, n, begin_bit, end_bit, 0)
        if rc:
            raise G.GpuSpatialError(L.lib().SjhipLastError().decode())
        k2, v2 = ok.download(np.uint32, n), ov.download(np.uint32, n)
        # the inputs are left as they were
        assert np.array_equal(dk.download(np.uint32, n), keys)
        return k2, v2
    finally:
        for b in (dk, dv, ok, ov):
            b.free()


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4095, 4096, 4097, 65536, 100_003])
@pytest.mark.parametrize("begin_bit,end_bit", [(0, 32), (0, 20), (4, 30)])
def test_graph_sort_is_a_stable_sort(n, begin_bit, end_bit):
    print("graph sort", n, begin_bit, end_bit, flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(n + 7 * begin_bit + end_bit)
    # few distinct digits in some passes (long runs of ties) and full-range keys
    keys = np.where(rng.random(n) < 0.5, rng.integers(0, 16, n) << 8,
                    rng.integers(0, 2**32, n, dtype=np.uint64)).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    mask = ((1 << (end_bit - begin_bit)) - 1) << begin_bit
    field = keys & np.uint32(mask)
    order = np.argsort(field, kind="stable").astype(np.uint32)
    k2, v2 = _sort(keys, vals, begin_bit, end_bit)
    assert np.array_equal(v2, order)
    assert np.array_equal(k2, keys[order])


def test_graph_sort_rejects_bad_arguments():
    keys = np.zeros(4, np.uint32)
    with pytest.raises(G.GpuSpatialError, match="bit range"):
        _sort(keys, keys, 8, 8)
