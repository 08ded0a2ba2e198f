"""Seeded synthetic workloads for the spatial join, written straight into
Arrow-style WKB columns (int64 offsets + uint8 data), little endian.

Workloads (sizes are the full ones; every generator takes counts so tests
can run the same distributions at small scale):

* ``points_uniform``     -- x, y ~ U[lo, hi]
* ``points_clustered``   -- 80 % from 64 isotropic Gaussians (sigma 2,
                            centres U[0, extent]^2), 20 % uniform
* ``star_polygons``      -- centre, n vertices at sorted uniform angles,
                            r_i = R * (0.7 + 0.3 u_i), R lognormal
* ``random_walks``       -- linestrings of n vertices, steps ~ N(0, s^2)

Star polygon vertex counts:
  1M points x 10k stars     : n ~ U{8..32}, R lognormal with mean 0.5
  100M points x 1M stars    : n = clip(Zipf(1.2), 4, 4096), R median 1
  5M x 5M stars             : n = floor(exp(U[ln 4, ln 1024])), R median 0.8
The lognormal sigma is not given by the workload text; 0.5 is used.
"""
import numpy as np

POINT_REC = np.dtype([("bo", "u1"), ("t", "<u4"), ("x", "<f8"), ("y", "<f8")], align=False)
assert POINT_REC.itemsize == 21

LOGN_SIGMA = 0.5


class WkbColumn:
    """Arrow Binary/LargeBinary-shaped WKB column held as numpy buffers."""

    def __init__(self, offsets, data, valid=None):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.data = np.ascontiguousarray(data, dtype=np.uint8)
        self.valid = valid

    def __len__(self):
        return len(self.offsets) - 1

    def as_tuple(self):
        return (self.offsets, self.data, self.valid)

    def to_arrow(self, large=None):
        import pyarrow as pa
        n = len(self)
        if large is None:
            large = self.offsets[-1] >= 2**31 - 1
        if large:
            typ, off = pa.large_binary(), self.offsets
        else:
            typ, off = pa.binary(), self.offsets.astype(np.int32)
        vbuf = None
        nulls = 0
        if self.valid is not None:
            v = np.asarray(self.valid, dtype=bool)
            nulls = int((~v).sum())
            vbuf = pa.py_buffer(np.packbits(v, bitorder="little"))
        return pa.Array.from_buffers(typ, n, [vbuf, pa.py_buffer(off), pa.py_buffer(self.data)], null_count=nulls)

    def slice(self, start, stop):
        o = self.offsets[start:stop + 1]
        v = None if self.valid is None else self.valid[start:stop]
        return WkbColumn(o - o[0], self.data[o[0]:o[-1]], v)


def points_wkb(x, y):
    n = len(x)
    rec = np.empty(n, dtype=POINT_REC)
    rec["bo"] = 1
    rec["t"] = 1
    rec["x"] = x
    rec["y"] = y
    return WkbColumn(np.arange(n + 1, dtype=np.int64) * 21, rec.view(np.uint8))


def rings_wkb(counts, xy, geom_type, ring_counts=None):
    """WKB for single-ring polygons (geom_type 3) or linestrings (2).

    counts: vertices per geometry (closing vertex included for polygons);
    xy: float64 (sum(counts), 2) coordinates in geometry order."""
    counts = np.asarray(counts, dtype=np.int64)
    n = len(counts)
    hdr = 13 if geom_type == 3 else 9
    sizes = hdr + 16 * counts
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    buf = np.zeros(int(off[-1]) + 16, dtype=np.uint8)  # +16 slack for aligned views
    # headers: byte order, type, (nrings=1,) npoints
    buf[off[:-1]] = 1
    _put_u32(buf, off[:-1] + 1, np.full(n, geom_type, np.uint32))
    if geom_type == 3:
        _put_u32(buf, off[:-1] + 5, np.ones(n, np.uint32))
        _put_u32(buf, off[:-1] + 9, counts.astype(np.uint32))
    else:
        _put_u32(buf, off[:-1] + 5, counts.astype(np.uint32))
    # coordinates: group records by alignment of their first coordinate byte
    first = off[:-1] + hdr
    vstart = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=vstart[1:])
    flat = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
    for a in range(8):
        sel = np.nonzero((first % 8) == a)[0]
        if len(sel) == 0:
            continue
        view = buf[a:a + ((len(buf) - a) // 8) * 8].view("<f8")
        # process in chunks to bound index memory
        CH = 1 << 22
        i = 0
        while i < len(sel):
            s = sel[i:i + CH]
            # choose chunk so that the float count stays bounded
            c2 = 2 * counts[s]
            tot = int(c2.sum())
            if tot > (1 << 27) and len(s) > 1:
                k = max(1, int(np.searchsorted(np.cumsum(c2), 1 << 27)))
                s = s[:k]
                c2 = c2[:k]
                tot = int(c2.sum())
            i += len(s)
            dst0 = (first[s] - a) // 8
            src0 = 2 * vstart[s]
            rep_dst = np.repeat(dst0, c2)
            rep_src = np.repeat(src0, c2)
            local = np.arange(tot, dtype=np.int64) - np.repeat(np.cumsum(c2) - c2, c2)
            view[rep_dst + local] = flat[rep_src + local]
    return WkbColumn(off, buf[:int(off[-1])])


def _put_u32(buf, pos, vals):
    v = np.asarray(vals, dtype="<u4").view(np.uint8).reshape(-1, 4)
    for k in range(4):
        buf[pos + k] = v[:, k]


def _sorted_angles(rng, counts):
    """Sorted uniform angles in [0, 2pi) per group via normalised cumulative
    exponential gaps (order statistics of uniforms)."""
    total = int(counts.sum())
    gaps = rng.exponential(size=total + len(counts))
    # one extra gap per group closes the circle
    grp_len = counts + 1
    start = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(grp_len, out=start[1:])
    cs = np.cumsum(gaps)
    base = np.concatenate([[0.0], cs[start[1:-1] - 1]])
    tot = cs[start[1:] - 1] - base
    gid = np.repeat(np.arange(len(counts)), grp_len)
    pos = np.arange(total + len(counts)) - start[gid]
    keep = pos < counts[gid]
    frac = (cs - base[gid]) / tot[gid]
    return frac[keep] * (2 * np.pi)


def star_polygons(rng, n_polys, counts, centers, radius_mu, radius_sigma=LOGN_SIGMA):
    """Star polygons: counts[i] distinct vertices (ring closed by repeating
    the first vertex), sorted angles, r = R * (0.7 + 0.3 u)."""
    counts = np.asarray(counts, dtype=np.int64)
    R = rng.lognormal(mean=radius_mu, sigma=radius_sigma, size=n_polys)
    # chunk over polygons to bound memory
    CH = max(1, int(2e7 // max(1.0, counts.mean())))
    offs = [np.zeros(1, np.int64)]
    datas = []
    base = 0
    for s in range(0, n_polys, CH):
        e = min(n_polys, s + CH)
        c = counts[s:e]
        ang = _sorted_angles(rng, c)
        gid = np.repeat(np.arange(s, e), c)
        r = R[gid] * (0.7 + 0.3 * rng.random(len(ang)))
        x = centers[gid, 0] + r * np.cos(ang)
        y = centers[gid, 1] + r * np.sin(ang)
        # close rings: append first vertex after each group
        cc = c + 1
        vstart = np.zeros(len(c) + 1, np.int64)
        np.cumsum(c, out=vstart[1:])
        idx = np.arange(int(cc.sum())) - np.repeat(np.cumsum(cc) - cc, cc)
        gsel = np.repeat(np.arange(len(c)), cc)
        src = vstart[gsel] + np.where(idx == c[gsel], 0, idx)
        xy = np.stack([x[src], y[src]], axis=1)
        col = rings_wkb(cc, xy, 3)
        offs.append(col.offsets[1:] + base)
        base += int(col.offsets[-1])
        datas.append(col.data)
    return WkbColumn(np.concatenate(offs), np.concatenate(datas) if datas else np.zeros(0, np.uint8))


def random_walks(rng, n_lines, counts, starts, step_sigma):
    counts = np.asarray(counts, dtype=np.int64)
    CH = max(1, int(2e7 // max(1.0, counts.mean())))
    offs = [np.zeros(1, np.int64)]
    datas = []
    base = 0
    for s in range(0, n_lines, CH):
        e = min(n_lines, s + CH)
        c = counts[s:e]
        tot = int(c.sum())
        steps = rng.normal(0.0, step_sigma, size=(tot, 2))
        vstart = np.zeros(len(c) + 1, np.int64)
        np.cumsum(c, out=vstart[1:])
        steps[vstart[:-1]] = starts[s:e]
        cs = np.cumsum(steps, axis=0)
        prev = np.concatenate([[[0.0, 0.0]], cs[vstart[1:-1] - 1]])
        gid = np.repeat(np.arange(len(c)), c)
        xy = cs - prev[gid]
        col = rings_wkb(c, xy, 2)
        offs.append(col.offsets[1:] + base)
        base += int(col.offsets[-1])
        datas.append(col.data)
    return WkbColumn(np.concatenate(offs), np.concatenate(datas) if datas else np.zeros(0, np.uint8))


def clustered_xy(rng, n, extent, n_clusters=64, sigma=2.0, frac_clustered=0.8, centers=None):
    if centers is None:
        centers = rng.uniform(0, extent, size=(n_clusters, 2))
    nc = int(round(n * frac_clustered))
    which = rng.integers(0, n_clusters, size=nc)
    pts = np.empty((n, 2))
    pts[:nc] = centers[which] + rng.normal(0, sigma, size=(nc, 2))
    pts[nc:] = rng.uniform(0, extent, size=(n - nc, 2))
    perm = rng.permutation(n)
    return pts[perm]


# ---------------------------------------------------------------------------
# Named workloads


def workload_pip_uniform(n_points=1_000_000, n_polys=10_000, seed=1, extent=100.0):
    """1M points x 10k stars (ST_Intersects), the CPU-runnable case."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, extent, size=(n_points, 2))
    probe = points_wkb(pts[:, 0], pts[:, 1])
    centers = rng.uniform(0, extent, size=(n_polys, 2))
    counts = rng.integers(8, 33, size=n_polys)
    mu = np.log(0.5) - LOGN_SIGMA**2 / 2  # mean 0.5
    build = star_polygons(rng, n_polys, counts, centers, mu)
    return build, probe


def zipf_counts(rng, n, a=1.2, lo=4, hi=4096):
    return np.clip(rng.zipf(a, size=n), lo, hi).astype(np.int64)


def workload_pip_zipf(n_points=100_000_000, n_polys=1_000_000, seed=2, extent=None, max_vertices=4096):
    """100M clustered points x 1M Zipf(1.2) stars (ST_Contains).  At reduced
    counts the extent shrinks so both densities stay those of the full size."""
    if extent is None:
        extent = 1000.0 * np.sqrt(n_polys / 1_000_000)
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0, extent, size=(n_polys, 2))
    counts = zipf_counts(rng, n_polys, hi=max_vertices)
    build = star_polygons(rng, n_polys, counts, centers, 0.0)
    n_clusters = max(1, int(round(64 * n_polys / 1_000_000))) if n_polys < 1_000_000 else 64
    pts = clustered_xy(rng, n_points, extent, n_clusters=n_clusters)
    probe = points_wkb(pts[:, 0], pts[:, 1])
    return build, probe


def workload_dwithin_walks(n_points=50_000_000, n_lines=5_000_000, seed=3, extent=None):
    """50M uniform points x 5M random-walk linestrings (ST_DWithin 0.05)."""
    if extent is None:
        extent = 1000.0 * np.sqrt(n_lines / 5_000_000)
    rng = np.random.default_rng(seed)
    counts = rng.integers(2, 257, size=n_lines)
    starts = rng.uniform(0, extent, size=(n_lines, 2))
    build = random_walks(rng, n_lines, counts, starts, 0.5)
    pts = rng.uniform(0, extent, size=(n_points, 2))
    probe = points_wkb(pts[:, 0], pts[:, 1])
    return build, probe


def workload_poly_poly(n_a=5_000_000, n_b=5_000_000, seed=4, extent=None):
    """Two sets of clustered star polygons (ST_Intersects)."""
    if extent is None:
        extent = 1000.0 * np.sqrt(n_a / 5_000_000)
    rng = np.random.default_rng(seed)
    ccent = rng.uniform(0, extent, size=(64, 2))
    cols = []
    for n in (n_a, n_b):
        counts = np.floor(np.exp(rng.uniform(np.log(4), np.log(1024), size=n))).astype(np.int64)
        centers = clustered_xy(rng, n, extent, centers=ccent)
        cols.append(star_polygons(rng, n, counts, centers, np.log(0.8)))
    return cols[0], cols[1]


def workload_knn(n_queries=20_000_000, n_objects=200_000_000, seed=5, extent=None):
    """KNN k=8: clustered query points vs clustered object points."""
    if extent is None:
        extent = 1000.0 * np.sqrt(n_objects / 200_000_000)
    rng = np.random.default_rng(seed)
    ccent = rng.uniform(0, extent, size=(64, 2))
    obj = clustered_xy(rng, n_objects, extent, centers=ccent)
    q = clustered_xy(rng, n_queries, extent, centers=ccent)
    return points_wkb(obj[:, 0], obj[:, 1]), points_wkb(q[:, 0], q[:, 1])


def serving_requests(n_requests, seed=6, lo=1_000, hi=10_000_000, extent=1000.0, mix=("intersects", "dwithin")):
    """Request sizes log-uniform in [lo, hi] with a fixed predicate mix."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_requests))).astype(np.int64)
    preds = [mix[i % len(mix)] for i in range(n_requests)]
    return sizes, preds, rng


# ---------------------------------------------------------------------------
# Parallel generation for the full-size workloads: per-row parameters are
# drawn in the parent; vertex-level work runs in forked workers (one child
# RNG per chunk, numpy.random.SeedSequence.spawn) writing straight into
# shared memory.  Same distributions as the sequential generators above,
# different random streams.  Call before any GPU runtime is initialised in
# the process (the workers are forked).


def _shm_array(shape, dtype):
    from multiprocessing import shared_memory
    nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
    shm = shared_memory.SharedMemory(create=True, size=max(nbytes, 1))
    arr = np.ndarray(shape, dtype=dtype, buffer=shm.buf)
    return shm, arr


def _run_parallel(fn, tasks, workers):
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        for _ in pool.imap_unordered(fn, tasks):
            pass


_PAR = {}


def _star_chunk(task):
    lo, hi, seed = task
    g = _PAR["star"]
    rng = np.random.default_rng(seed)
    counts = g["counts"][lo:hi]
    ang = _sorted_angles(rng, counts)
    gid = np.repeat(np.arange(lo, hi), counts)
    r = g["R"][gid] * (0.7 + 0.3 * rng.random(len(ang)))
    x = g["centers"][gid, 0] + r * np.cos(ang)
    y = g["centers"][gid, 1] + r * np.sin(ang)
    cc = counts + 1
    vstart = np.zeros(len(counts) + 1, np.int64)
    np.cumsum(counts, out=vstart[1:])
    idx = np.arange(int(cc.sum())) - np.repeat(np.cumsum(cc) - cc, cc)
    gsel = np.repeat(np.arange(len(counts)), cc)
    src = vstart[gsel] + np.where(idx == counts[gsel], 0, idx)
    col = rings_wkb(cc, np.stack([x[src], y[src]], axis=1), 3)
    off = g["offsets"]
    g["data"][off[lo]:off[hi]] = col.data


def star_polygons_parallel(seed, counts, centers, radius_mu, radius_sigma=LOGN_SIGMA, workers=16):
    """Star polygons (as star_polygons) generated by forked workers; returns
    (WkbColumn over shared memory, keepalive handles)."""
    counts = np.asarray(counts, dtype=np.int64)
    n = len(counts)
    ss = np.random.SeedSequence(seed)
    prm_rng = np.random.default_rng(ss.spawn(1)[0])
    R = prm_rng.lognormal(mean=radius_mu, sigma=radius_sigma, size=n)
    sizes = 13 + 16 * (counts + 1)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    shm, data = _shm_array((int(off[-1]),), np.uint8)
    _PAR["star"] = dict(counts=counts, R=R, centers=centers, offsets=off, data=data)
    per = max(1, int(4e6 // max(1.0, counts.mean())))
    bounds = list(range(0, n, per)) + [n]
    seeds = ss.spawn(len(bounds))[1:]
    tasks = [(bounds[i], bounds[i + 1], seeds[i]) for i in range(len(bounds) - 1)]
    _run_parallel(_star_chunk, tasks, workers)
    _PAR.pop("star", None)
    return WkbColumn(off, data), (shm,)


def _walk_chunk(task):
    lo, hi, seed = task
    g = _PAR["walk"]
    rng = np.random.default_rng(seed)
    c = g["counts"][lo:hi]
    steps = rng.normal(0.0, g["sigma"], size=(int(c.sum()), 2))
    vstart = np.zeros(len(c) + 1, np.int64)
    np.cumsum(c, out=vstart[1:])
    steps[vstart[:-1]] = g["starts"][lo:hi]
    cs = np.cumsum(steps, axis=0)
    prev = np.concatenate([[[0.0, 0.0]], cs[vstart[1:-1] - 1]])
    gid = np.repeat(np.arange(len(c)), c)
    col = rings_wkb(c, cs - prev[gid], 2)
    off = g["offsets"]
    g["data"][off[lo]:off[hi]] = col.data


def random_walks_parallel(seed, counts, starts, step_sigma, workers=16):
    """Random-walk linestrings (as random_walks) generated by forked
    workers; returns (WkbColumn over shared memory, keepalive handles)."""
    counts = np.asarray(counts, dtype=np.int64)
    n = len(counts)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(9 + 16 * counts, out=off[1:])
    shm, data = _shm_array((int(off[-1]),), np.uint8)
    _PAR["walk"] = dict(counts=counts, starts=starts, sigma=step_sigma, offsets=off, data=data)
    ss = np.random.SeedSequence(seed)
    per = max(1, int(4e6 // max(1.0, counts.mean())))
    bounds = list(range(0, n, per)) + [n]
    seeds = ss.spawn(len(bounds))
    tasks = [(bounds[i], bounds[i + 1], seeds[i]) for i in range(len(bounds) - 1)]
    _run_parallel(_walk_chunk, tasks, workers)
    _PAR.pop("walk", None)
    return WkbColumn(off, data), (shm,)


def workload_dwithin_walks_parallel(n_points=50_000_000, n_lines=5_000_000, seed=3, workers=16):
    """50M uniform points x 5M random-walk linestrings (2..256 vertices,
    steps N(0, 0.5^2) per axis, starts uniform), generated in parallel.
    Returns dict(build, probe, starts, counts, probe_xy, extent, keep)."""
    extent = 1000.0 * np.sqrt(n_lines / 5_000_000)
    rng = np.random.default_rng(seed)
    counts = rng.integers(2, 257, size=n_lines)
    starts = rng.uniform(0, extent, size=(n_lines, 2))
    build, k1 = random_walks_parallel(seed + 1000, counts, starts, 0.5, workers=workers)
    probe, xy, k2 = points_parallel(seed + 2000, n_points, extent, centers=None, workers=workers)
    return dict(build=build, probe=probe, starts=starts, centers=starts, counts=counts, probe_xy=xy, extent=extent,
                keep=k1 + k2)


def _points_chunk(task):
    lo, hi, seed = task
    g = _PAR["pts"]
    rng = np.random.default_rng(seed)
    n = hi - lo
    nc = int(round(n * g["frac"]))
    pts = np.empty((n, 2))
    if g["centers"] is not None:
        which = rng.integers(0, len(g["centers"]), size=nc)
        pts[:nc] = g["centers"][which] + rng.normal(0, g["sigma"], size=(nc, 2))
    else:
        nc = 0
    pts[nc:] = rng.uniform(0, g["extent"], size=(n - nc, 2))
    pts = pts[rng.permutation(n)]
    g["xy"][lo:hi] = pts
    rec = np.ndarray((n,), dtype=POINT_REC, buffer=g["data"], offset=21 * lo)
    rec["bo"] = 1
    rec["t"] = 1
    rec["x"] = pts[:, 0]
    rec["y"] = pts[:, 1]


def points_parallel(seed, n, extent, centers=None, sigma=2.0, frac_clustered=0.8, workers=16):
    """Points (clustered like clustered_xy when centers is given, uniform
    otherwise) by forked workers: (WkbColumn, xy array, keepalive)."""
    shm_d, data = _shm_array((21 * n,), np.uint8)
    shm_x, xy = _shm_array((n, 2), np.float64)
    _PAR["pts"] = dict(centers=centers, sigma=sigma, frac=frac_clustered if centers is not None else 0.0,
                       extent=extent, xy=xy, data=data)
    ss = np.random.SeedSequence(seed)
    per = 2_000_000
    bounds = list(range(0, n, per)) + [n]
    seeds = ss.spawn(len(bounds))
    tasks = [(bounds[i], bounds[i + 1], seeds[i]) for i in range(len(bounds) - 1)]
    _run_parallel(_points_chunk, tasks, workers)
    _PAR.pop("pts", None)
    return WkbColumn(np.arange(n + 1, dtype=np.int64) * 21, data), xy, (shm_d, shm_x)


def workload_knn_parallel(n_queries=20_000_000, n_objects=200_000_000, seed=5, workers=16):
    """KNN k=8: 20M query points vs 200M object points, both drawn from the
    100M-point mixture (64 Gaussians, sigma 2, centres U[0,1000]^2, 20 %
    uniform).  Returns dict(build, probe, probe_xy, build_xy, extent, keep)."""
    extent = 1000.0 * np.sqrt(n_objects / 200_000_000)
    rng = np.random.default_rng(seed)
    ccent = rng.uniform(0, extent, size=(64, 2))
    obj, oxy, k1 = points_parallel(seed + 1000, n_objects, extent, centers=ccent, workers=workers)
    q, qxy, k2 = points_parallel(seed + 2000, n_queries, extent, centers=ccent, workers=workers)
    return dict(build=obj, probe=q, build_xy=oxy, probe_xy=qxy, extent=extent, keep=k1 + k2)


def workload_pip_zipf_parallel(n_points=100_000_000, n_polys=1_000_000, seed=2, workers=16, max_vertices=4096):
    """The 100M clustered points x 1M Zipf(1.2) stars workload, generated in
    parallel.  Returns dict(build, probe, centers, probe_xy, keep)."""
    extent = 1000.0 * np.sqrt(n_polys / 1_000_000)
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0, extent, size=(n_polys, 2))
    counts = zipf_counts(rng, n_polys, hi=max_vertices)
    n_clusters = max(1, int(round(64 * n_polys / 1_000_000))) if n_polys < 1_000_000 else 64
    ccent = rng.uniform(0, extent, size=(n_clusters, 2))
    build, k1 = star_polygons_parallel(seed + 1000, counts, centers, 0.0, workers=workers)
    probe, xy, k2 = points_parallel(seed + 2000, n_points, extent, centers=ccent, workers=workers)
    return dict(build=build, probe=probe, centers=centers, counts=counts, probe_xy=xy, extent=extent,
                cluster_centers=ccent, keep=k1 + k2)


def release_shared(keep):
    for shm in keep:
        try:
            shm.close()
            shm.unlink()
        except Exception:
            pass


def take_rows(col, rows):
    """WkbColumn of the given rows (in order)."""
    rows = np.asarray(rows, dtype=np.int64)
    lens = col.offsets[rows + 1] - col.offsets[rows]
    off = np.zeros(len(rows) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    if len(rows) and np.all(lens == lens[0]):
        w = int(lens[0])
        base = col.offsets[rows][:, None] + np.arange(w)[None, :]
        return WkbColumn(off, col.data[base.reshape(-1)])
    data = np.concatenate([col.data[col.offsets[r]:col.offsets[r + 1]] for r in rows]) if len(rows) else \
        np.zeros(0, np.uint8)
    return WkbColumn(off, data)
