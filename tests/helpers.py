"""Shared helpers of the test suite."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def read_fixture(name):
    import pyarrow.parquet as pq
    return pq.read_table(os.path.join(GOLDEN, "data", name)).column("geometry").combine_chunks()


_expected = None


def expected():
    global _expected
    if _expected is None:
        _expected = dict(np.load(os.path.join(GOLDEN, "expected.npz")))
    return _expected


def pairs_equal(a, b):
    """Pair sets equal (order-insensitive); returns (equal, missing, extra)."""
    ka = (np.asarray(a[1], np.uint64) << 32) | np.asarray(a[0], np.uint64)
    kb = (np.asarray(b[1], np.uint64) << 32) | np.asarray(b[0], np.uint64)
    sa, sb = np.unique(ka), np.unique(kb)
    missing = np.setdiff1d(sb, sa)
    extra = np.setdiff1d(sa, sb)
    return len(missing) == 0 and len(extra) == 0 and len(sa) == len(ka), len(missing), len(extra)
