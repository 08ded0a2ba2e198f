"""Benchmark: refined join pairs/s of the spatial join on one MI355X.

Workload (default): 100M probe points (80 % from 64 isotropic Gaussians,
sigma 2, centres U[0,1000]^2; 20 % uniform) x 1M build star polygons
(vertex counts clip(Zipf(1.2), 4, 4096), radius lognormal median 1),
ST_Contains(build polygon, probe point), FP64 WKB columns, seeded.

One step = one probe batch through the device-resident join: device WKB
decode of the 100M probe points + FP32 rects, Morton order of the probes,
filter over the resident polygon index, refine (certified FP32 + FP64
fallback), compaction of the refined pairs.  The build side (decode, ring
index, rect tree) is built once before the timed region and reported as
build_ms.  Inputs are resident in HBM when the timed region starts.

The CPU baseline is the reference's own TG refine (oracle/_ref, built from
the reference sources) behind a grid filter, run over the whole workload on
this host's usable CPUs (at most 16); its pair set must equal the GPU's pair
for pair.  A one-thread rate is timed on a narrow vertical strip.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sedona-db_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
FLOPS_PER_SEGMENT_TEST = 8  # two subtractions x2, cross product (3), compare


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the full workload (both sides)")
    ap.add_argument("--precision", type=int, default=0, help="0 = FP32 certified + FP64 fallback, 1 = all FP64")
    ap.add_argument("--cpu-strip", type=float, default=0.12,
                    help="CPU sample of the other workloads (all threads): strip width / extent")
    ap.add_argument("--cpu-strip-1t", type=float, default=0.01,
                    help="pip_zipf_contains: one-thread CPU sample, strip width / extent")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-full", dest="cpu_full", action="store_false",
                    help="pip_zipf_contains: skip the reference CPU join of the whole workload (all threads)")
    ap.add_argument("--no-plain", action="store_true", help="skip the all-FP64 comparison step")
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--refine-order", type=int, default=None,
                    help="0 = refine in filter order, 1 = group candidates by build row first, 2 = (DWithin) fused "
                         "into the filter walk; default 1 for pip_zipf_contains, 0 for dwithin_walks")
    ap.add_argument("--workload", default="pip_zipf_contains",
                    choices=["pip_zipf_contains", "dwithin_walks", "knn", "poly_poly"],
                    help="pip_zipf_contains: 100M points x 1M Zipf stars, ST_Contains (the headline); "
                         "dwithin_walks: 50M points x 5M random-walk lines, ST_DWithin 0.05; "
                         "knn: 20M queries x 200M clustered points, k=8; "
                         "poly_poly: 2 x 5M clustered stars, ST_Intersects (use --scale)")
    ap.add_argument("--knn-sample", type=int, default=200_000, help="knn: queries checked against the oracle")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="pip_zipf_contains: skip the autotuner (use --refine-order and the default kernels)")
    ap.add_argument("--seg-leaf", type=int, default=0,
                    help="pip_zipf_contains: segments per ring-tree leaf (16 or 32) set before the joins (0 = keep)")
    ap.add_argument("--pip-waves", type=int, default=0,
                    help="pip_zipf_contains: resident waves of the point-in-polygon kernel (4, 6, 8) set before the "
                         "joins (0 = keep)")
    ap.add_argument("--probe-batch", type=int, default=5_000_000, help="dwithin_walks: probe points per batch")
    ap.add_argument("--poly-batch", type=int, default=500_000,
                    help="poly_poly: probe polygons per join call (0: all; each call's candidates must stay below 2^32)")
    ap.add_argument("--serve-requests", type=int, default=96,
                    help="requests in the serving measurement (0 = skip it)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--no-host", dest="host_inclusive", action="store_false",
                    help="pip_zipf_contains: skip the host-inclusive (PCIe) measurement")
    ap.add_argument("--abi-batch", type=int, default=1_000_000,
                    help="pip_zipf_contains: probe rows per batch in the reference-ABI host flow")
    ap.add_argument("--stream-batch", type=int, default=10_000_000,
                    help="pip_zipf_contains: probe rows per batch of the pipelined host-column join")
    ap.add_argument("--stream-slots", type=int, default=2,
                    help="pip_zipf_contains: batches in flight in the pipelined host-column join")
    ap.add_argument("--partition-budget-gb", type=float, default=0.0,
                    help="pip_zipf_contains: also run the partitioned join with this device budget for the build side "
                         "(0 = skip)")
    ap.add_argument("--abi-threads", type=int, default=4,
                    help="pip_zipf_contains: threads of the concurrent reference-ABI host flow")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only time the plain copy kernel (HBM calibration for the counter summaries)")
    args = ap.parse_args()
    if args.refine_order is None:
        args.refine_order = 1

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: every box in this build has one GPU, so --gpus must be 1 "
              "(a multi-GPU run would shard the probe side: see the design notes)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "dwithin_walks":
        run_dwithin_walks(args)
        return
    if args.workload == "knn":
        run_knn(args)
        return
    if args.workload == "poly_poly":
        run_poly_poly(args)
        return

    from sjhip import datagen as D

    n_points = int(round(100_000_000 * args.scale))
    n_polys = int(round(1_000_000 * args.scale))
    t0 = time.time()
    wl = D.workload_pip_zipf_parallel(n_points, n_polys, seed=args.seed, workers=args.workers)
    build, probe = wl["build"], wl["probe"]
    log(f"generated {n_polys} polygons ({int(wl['counts'].sum()) + n_polys} vertices, "
        f"{build.offsets[-1] / 1e9:.1f} GB WKB) and {n_points} points in {time.time() - t0:.1f}s")

    # CPU baseline sample (strip) is cut before any GPU work
    strip = None
    if not args.no_cpu:
        ext = wl["extent"]
        x0 = 0.5 * ext
        w = args.cpu_strip_1t * ext
        brows = np.nonzero((wl["centers"][:, 0] >= x0) & (wl["centers"][:, 0] < x0 + w))[0]
        prows = np.nonzero((wl["probe_xy"][:, 0] >= x0) & (wl["probe_xy"][:, 0] < x0 + w))[0]
        strip = dict(brows=brows, prows=prows, bcol=D.take_rows(build, brows), pcol=D.take_rows(probe, prows),
                     x0=x0, w=w)
        log(f"cpu strip: {len(brows)} polygons, {len(prows)} points")

    # serving requests (generated by forked workers, so before the GPU is used)
    serve = None
    if args.serve_requests > 0:
        sizes, _, _ = D.serving_requests(args.serve_requests, seed=args.seed + 4)
        scol, _, skeep = D.points_parallel(args.seed + 3000, int(sizes.sum()), wl["extent"],
                                           centers=wl["cluster_centers"], workers=args.workers)
        serve = dict(sizes=sizes, col=scol, keep=skeep)
        log(f"serving: {len(sizes)} requests, {int(sizes.sum())} points")

    from sjhip import gpuspatial as G

    opts = G.GpuSpatialOptions(device_id=0)
    tb = time.time()
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema("Z")
    bdev = G.DeviceWkb.from_host(build)
    rf.push_build_device(bdev)
    rf.finish_building()
    del bdev
    ix = G.GpuSpatialIndex(opts)
    ix.build_from_refiner(rf)
    build_ms = (time.time() - tb) * 1000
    log(f"build side resident: {build_ms:.0f} ms (includes host->device copy)")
    pdev = G.DeviceWkb.from_host(probe)
    if args.seg_leaf:
        rf.set_segment_leaf(args.seg_leaf)
    if args.pip_waves:
        ix.set_tuning(kind=0, refine_order=-1, pip_waves=args.pip_waves)
    tuned = None
    if args.autotune:
        ta = time.time()
        # the tuner times joins of the whole resident probe column (about
        # 2 x 12 joins, a few seconds): a strip of a tenth of the points had
        # the workload's density but no longer ranked the candidate orders
        # as the full join does (filter order won on the strip and lost by
        # 16 % at full size)
        tuned = G.autotune(ix, rf, pdev, "contains")
        tuned["tune_s"] = time.time() - ta
        tuned["sample_points"] = int(n_points)
        log(f"autotune (all {n_points} points): {tuned}")

    def step(precision, profile=True, sort_output=False):
        return G.join_device(ix, rf, pdev, "contains", precision=precision, sort_output=sort_output,
                             sort_probes=True, refine_order=-1 if args.autotune else args.refine_order,
                             profile=profile)

    for i in range(args.warmup):
        r = step(args.precision)
        log(f"warmup {i}: {r.size} pairs, {r.candidates} candidates, stages {list(r.res.stage_ms)}")
        r.release()
    stage = np.zeros(6)
    sizes = []
    t = time.time()
    for i in range(args.steps):
        r = step(args.precision)
        stage += np.array(list(r.res.stage_ms))
        sizes.append((r.size, r.candidates, r.segment_tests, r.fallback_tests, r.refine_candidates))
        if i < args.steps - 1:
            r.release()
    elapsed = time.time() - t
    last = r
    ms_per_step = elapsed * 1000 / args.steps
    pairs, cands, seg_tests, fb, walked = sizes[-1]
    assert all(s[0] == pairs for s in sizes), "pair count changed between steps"
    stage /= args.steps
    value = pairs / (ms_per_step / 1000)
    log(f"timed: {ms_per_step:.2f} ms/step, {pairs} pairs, stages(ms) decode {stage[0]:.2f} morton {stage[1]:.2f} "
        f"filter {stage[2]:.2f} refine {stage[3]:.2f}")

    last.release()
    # untimed: the probe-major filter (one lane per Morton-ordered point,
    # the rects of its grid cell; then the raster split and the regrouping
    # sort) in place of the build-rect pass, for its own roofline
    ix.set_filter_mode(3)
    pmaj = []
    for _ in range(2):
        r = step(args.precision)
        pmaj.append((list(r.res.stage_ms), r.size))
        r.release()
    ix.set_filter_mode(0)
    assert all(x[1] == pairs for x in pmaj), "probe-major pairs differ"
    pmaj_stage = np.mean([x[0] for x in pmaj], axis=0)
    log(f"probe-major filter: {pmaj_stage[2]:.2f} ms, split and regroup {pmaj_stage[5]:.2f} ms")

    # untimed: the same join with its pairs sorted by (probe, build) on the
    # device, for the exact pair-set comparisons below
    t_sorted = time.time()
    r = step(args.precision, profile=False, sort_output=True)
    sorted_ms = (time.time() - t_sorted) * 1000
    gb_pairs, gp_pairs = r.to_host()
    r.release()
    assert len(gb_pairs) == pairs

    plain = None
    if not args.no_plain:
        r1 = step(1 - args.precision)
        t1 = time.time()
        r1b = step(1 - args.precision)
        plain_ms = (time.time() - t1) * 1000
        plain = dict(precision=1 - args.precision, ms_per_step=plain_ms, pairs=int(r1b.size),
                     stage_ms=list(r1b.res.stage_ms))
        r1.release()
        r1b.release()
        r1c = step(1 - args.precision, profile=False, sort_output=True)
        pb, pp = r1c.to_host()
        r1c.release()
        plain["pair_set_diff"] = pair_set_diff(gb_pairs, gp_pairs, pb, pp)
        del pb, pp
        log(f"all-FP64 plain: {plain_ms:.1f} ms/step, pair-set diff {plain['pair_set_diff']}")

    host = None
    if args.host_inclusive:
        host = host_inclusive_measurement(G, ix, rf, probe, pairs, args)

    # a build side split under a device-memory budget (SjhipPartitionedJoin):
    # K-D-B partitions built and joined one at a time against the resident
    # probe column
    partitioned = None
    if args.partition_budget_gb > 0:
        t = time.time()
        xb, xp, st = G.partitioned_join(build.to_arrow(large=True), pdev, "contains",
                                        int(args.partition_budget_gb * (1 << 30)))
        part_ms = (time.time() - t) * 1000
        partitioned = dict(budget_bytes=int(args.partition_budget_gb * (1 << 30)), ms=part_ms, stats=st,
                           pairs=int(len(xb)), same_pairs=bool(len(xb) == pairs and pair_checksum(xb, xp) ==
                                                                pair_checksum(gb_pairs, gp_pairs)))
        log(f"partitioned join: {st['partitions']} partitions under {args.partition_budget_gb} GiB, {part_ms:.0f} ms "
            f"(build {st['build_ms']:.0f} ms, joins {st['join_ms']:.0f} ms), same pairs {partitioned['same_pairs']}")
        del xb, xp

    # Rooflines.  Algorithmic bytes per launch:
    #   filter: 16 B per probe rect read + 8 B per candidate pair written;
    #   refine: 8 B per candidate it reads + 8 B per kept pair (every pair of
    #           the join: an upper bound where a raster split wrote part of
    #           them before) + the FP32 vertex runs of the build polygons
    #           (8 B per vertex, each distinct polygon counted once).
    # With polygon-major order (1) the cell-raster split runs before the
    # regrouping (in its stage), and the refine stage is the certified FP32
    # walk alone over the candidates the split left open: k_walk_rings for
    # rows of one ring (all of this workload's), k_refine_pip for the rest.
    # In filter order (0) the refine stage is three launches on one stream,
    # timed together: the raster pass (k_raster_pass), the copy of the open
    # candidates (k_open_scatter) and the walk (no profiled traffic).  The
    # refine's arithmetic roof (FP32 VALU, 8 flops per segment test) is
    # reported beside its HBM roof.
    order_used = tuned["refine_order"] if tuned else args.refine_order
    filt_ms, ref_ms = stage[2], stage[3]
    filter_bytes = 16 * n_points + 8 * cands
    verts32 = 8 * (int(wl["counts"].sum()) + n_polys)
    refine_bytes = 8 * (walked if order_used == 1 else cands) + 8 * pairs + verts32
    refine_flops = FLOPS_PER_SEGMENT_TEST * seg_tests
    traffic, traffic_source = read_traffic(os.path.join(ROOT, "profiles", "traffic_pip_zipf_100Mx1M_contains.md"))
    full = args.scale == 1.0 and args.precision == 0
    # the filter stage of the timed joins is the build-rect pass (k_pm_*:
    # each build rect tests the Morton cells it covers, the cell-raster
    # split decided in the same pass); its algorithmic bytes are the
    # filter's (it also replaces the raster pass and the regrouping sort)
    roof_filter = dict(bound="hbm", achieved=filter_bytes / (filt_ms / 1e3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                       traffic=traffic.get("k_pm_split") if full and order_used == 1 else None,
                       kernel="k_pm_ends+k_pm_count+k_pm_fill+k_pm_test+k_pm_emit", ms=filt_ms,
                       algorithmic_bytes=filter_bytes)
    roof_filter["frac"] = roof_filter["achieved"] / HBM_PEAK_GBS
    pm_ms = pmaj_stage[2]
    roof_filter_pm = dict(bound="hbm", achieved=filter_bytes / (pm_ms / 1e3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                          traffic=traffic.get("k_filter_grid") if full else None,
                          kernel="k_filter_grid", ms=pm_ms, algorithmic_bytes=filter_bytes,
                          split_and_regroup_ms=pmaj_stage[5])
    roof_filter_pm["frac"] = roof_filter_pm["achieved"] / HBM_PEAK_GBS
    ach_tf = refine_flops / (ref_ms / 1e3) / 1e12
    roof_refine = dict(bound="hbm", achieved=refine_bytes / (ref_ms / 1e3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                       traffic=traffic.get("k_walk_rings") if full and order_used == 1 else None,
                       kernel="k_walk_rings" if order_used == 1 else "k_raster_pass+k_open_scatter+k_walk_rings",
                       ms=ref_ms,
                       algorithmic_bytes=refine_bytes,
                       valu_fp32={"achieved": ach_tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": ach_tf / FP32_PEAK_TFLOPS, "flops": refine_flops})
    roof_refine["frac"] = roof_refine["achieved"] / HBM_PEAK_GBS
    dominant = roof_refine if ref_ms >= filt_ms else roof_filter

    serving = None
    if serve is not None:
        try:
            serving = serving_measurement(G, ix, rf, serve, args)
        except G.GpuSpatialError as e:
            # reported in the result line; the headline above stands
            import traceback
            traceback.print_exc()
            log(f"serving measurement failed: {e}")
            serving = {"error": str(e)}
        D.release_shared(serve["keep"])

    cpu = None
    if strip is not None:
        cpu = cpu_baseline_c2(build, probe, strip, gb_pairs, gp_pairs, args)
    del gb_pairs, gp_pairs

    out = {
        "metric": "refined join pairs/s (ST_Contains, 100M points x 1M Zipf star polygons, device-resident)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 1 else "f32+f64",
        "data": "synthetic (seeded, generated on the host)",
        "config": {"workload": "pip_zipf_contains", "probe_points": n_points, "build_polygons": n_polys,
                   "build_vertices": int(wl["counts"].sum()) + n_polys, "predicate": "ST_Contains",
                   "precision": "fp32-certified+fp64-fallback" if args.precision == 0 else "fp64-plain",
                   "output": "pairs in refine order (an inner join needs no order; the reference refiner "
                             "sorts by probe: see sorted_output_ms_per_step)",
                   "parallelism": "single-gpu"},
        "sorted_output_ms_per_step": sorted_ms,
        "roofline": {k: dominant[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_kernel": dominant["kernel"],
        "roofline_traffic_source": traffic_source,
        "roofline_valu_fp32": dominant.get("valu_fp32"),
        "roofline_stages": {"filter_split": roof_filter, "filter_probe_major": roof_filter_pm, "refine": roof_refine},
        "cpu_baseline": cpu,
        "pairs_per_step": int(pairs),
        "candidates_per_step": int(cands),
        "walked_candidates_per_step": int(walked),
        "probe_points_per_s": n_points / (ms_per_step / 1000),
        "stage_ms": {"decode_rects": stage[0], "morton": stage[1], "filter": stage[2],
                     "raster_split_and_group": stage[5], "refine": stage[3]},
        "segment_tests": int(seg_tests),
        "fallback_rate": fb / max(1, seg_tests),
        "build_ms": build_ms,
        "plain_fp64": plain,
        "autotune": tuned,
        "serving": serving,
        "host_inclusive": host,
        "partitioned": partitioned,
    }
    if plain and args.precision == 0:
        out["speedup_vs_plain_fp64"] = plain["ms_per_step"] / ms_per_step
        out["refine_speedup_vs_plain_fp64"] = plain["stage_ms"][3] / stage[3]
    print(json.dumps(out), flush=True)
    D.release_shared(wl["keep"])


def pair_set_diff(b1, p1, b2, p2):
    """Size of the symmetric difference of two pair sets, each sorted by
    (probe, build) without repeats."""
    if len(b1) == len(b2) and np.array_equal(p1, p2) and np.array_equal(b1, b2):
        return 0
    k1 = (p1.astype(np.uint64) << np.uint64(32)) | b1.astype(np.uint64)
    k2 = (p2.astype(np.uint64) << np.uint64(32)) | b2.astype(np.uint64)
    return int(len(np.setxor1d(k1, k2, assume_unique=True)))


def host_cpu_info():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail}


def cpu_threads():
    """Worker threads of the CPU baseline: the CPUs this process may use, at
    most 16 (one GPU's share of the host)."""
    info = host_cpu_info()
    return max(1, min(16, info["cpus_available"] or 1))


def cpu_reference_join(bcol, pcol, predicate, threads, cap_hint=0, what="polygons"):
    """The reference CPU join on host columns: WKB parse (the oracle's SoA
    for the FP32 rects, the reference TG library for the geometries), the
    grid filter and the reference TG refine.  (pairs, seconds, kind)."""
    import oracle as O
    kind = "reference" if O.ref_available() else "port"
    t = time.time()
    if kind == "reference":
        st_b, st_p = O.Store(bcol.as_tuple()), O.Store(pcol.as_tuple())
        br, pr = st_b.rects(True), st_p.rects(True)
        del st_b, st_p
        tb = O.TgColumn(bcol.as_tuple(), index=1, nthreads=threads)
        tp = O.TgColumn(pcol.as_tuple(), index=1 if what == "polygons" and predicate == "intersects" else 0,
                        nthreads=threads)
        cb, cp = O.ref_join(tb, tp, br, pr, predicate, nthreads=threads, cap_hint=cap_hint)
        secs = time.time() - t
        del tb, tp
    else:
        st_b, st_p = O.Store(bcol.as_tuple()), O.Store(pcol.as_tuple())
        cb, cp = O.join(st_b, st_p, predicate, nthreads=threads)
        secs = time.time() - t
    return cb, cp, secs, kind


def cpu_baseline_c2(build, probe, strip, gb, gp, args):
    """Reference CPU join of the whole workload on all usable host CPUs (its
    pair set against the GPU's, pair for pair), and a one-thread rate on a
    narrow vertical strip of the same workload."""
    threads = cpu_threads()
    res = host_cpu_info()
    if args.cpu_full:
        cb, cp, secs, kind = cpu_reference_join(build, probe, "contains", threads, cap_hint=len(gb) + (1 << 20))
        res.update({"value": len(cb) / secs, "unit": "pairs/s", "cores": threads, "kind": kind,
                    "sample": f"the whole workload: {len(build)} polygons x {len(probe)} points, {len(cb)} pairs, "
                              f"{secs:.2f} s (WKB parse + filter + refine)",
                    "cpu_pairs": int(len(cb)), "gpu_pairs": int(len(gb)),
                    "pair_set_diff": pair_set_diff(gb, gp, cb, cp)})
        log(f"cpu reference, whole workload, {threads} threads: {len(cb)} pairs in {secs:.1f} s, "
            f"pair-set diff {res['pair_set_diff']}")
        del cb, cp
    cb, cp, secs, kind = cpu_reference_join(strip["bcol"], strip["pcol"], "contains", 1)
    one = {"value": len(cb) / secs, "unit": "pairs/s", "cores": 1, "kind": kind,
           "sample": f"vertical strip x in [{strip['x0']:.1f}, {strip['x0'] + strip['w']:.1f}): "
                     f"{len(strip['brows'])} polygons x {len(strip['prows'])} points, {len(cb)} pairs, {secs:.2f} s"}
    one["pair_set_diff"] = strip_diff(strip, gb, gp, cb, cp)
    log(f"cpu reference, strip, 1 thread: {len(cb)} pairs in {secs:.1f} s")
    res["one_thread"] = one
    if not args.cpu_full:
        res.update({k: one[k] for k in ("value", "unit", "cores", "kind", "sample", "pair_set_diff")})
    return res


def strip_diff(strip, gb, gp, cb, cp):
    """Pair-set difference between the GPU's full pairs restricted to the
    strip's rows and the CPU pairs of the strip (local row numbers)."""
    bmap = np.full(int(gb.max()) + 1 if len(gb) else 1, -1, np.int64)
    sel_b = strip["brows"]
    bmap[sel_b[sel_b < len(bmap)]] = np.nonzero(sel_b < len(bmap))[0]
    pset = np.full(int(gp.max()) + 1 if len(gp) else 1, -1, np.int64)
    sel_p = strip["prows"]
    pset[sel_p[sel_p < len(pset)]] = np.nonzero(sel_p < len(pset))[0]
    lb = bmap[gb]
    lp = pset[gp]
    keep = (lb >= 0) & (lp >= 0)
    gk = np.sort((lp[keep].astype(np.uint64) << 32) | lb[keep].astype(np.uint64))
    ck = np.sort((cp.astype(np.uint64) << 32) | cb.astype(np.uint64))
    return int(len(np.setxor1d(gk, ck)))


def pair_checksum(b, p, p0=0):
    """Order-independent checksum of (build, probe + p0) pairs (sum mod 2^64
    of a mix of both ids)."""
    k = (b.astype(np.uint64) << np.uint64(32)) | (p.astype(np.uint64) + np.uint64(p0))
    k *= np.uint64(0x9E3779B97F4A7C15)
    k ^= k >> np.uint64(29)
    return int(np.sum(k, dtype=np.uint64))


def host_inclusive_measurement(G, ix, rf, probe, pairs, args):
    """The same join with the probe WKB in host memory and the pairs copied
    back to host memory (PCIe included), two ways:
      fused: one SjhipJoinDevice call over the host column, then the copy
             of its pairs;
      reference ABI: per batch of rows (the operator's probe batches,
             stream.rs:1166-1269), compute the rects, probe the index with
             the host rects (pairs come back through the callback,
             gpu_spatial_index.rs:163-229), then refine the host candidates
             against the host WKB batch (gpuspatial_c.h:154-211)."""
    parr = probe.to_arrow(large=True)
    n = len(parr)
    t = time.time()
    r = G.join_device(ix, rf, parr, "contains", sort_output=False, refine_order=-1 if args.autotune else args.refine_order)
    gb, gp = r.to_host()
    fused_ms = (time.time() - t) * 1000
    r.release()
    fused_sum = pair_checksum(gb, gp)
    fused = dict(ms=fused_ms, pairs=int(len(gb)), pairs_per_s=len(gb) / (fused_ms / 1000),
                 h2d_bytes=int(probe.offsets[-1] + 8 * (n + 1)), d2h_bytes=int(8 * len(gb)))
    del gb, gp
    log(f"host-inclusive fused join: {fused_ms:.1f} ms, {fused['pairs']} pairs")

    import threading

    bsz = args.abi_batch
    starts = list(range(0, n, bsz))

    def abi_flow(threads):
        """Batches dealt round-robin to `threads` threads, each with its own
        index context per probe call (the operator's partitions run
        concurrently; the ABI is thread-safe with one context per thread)."""
        acc = [[0, 0, 0] for _ in range(threads)]  # kept, candidates, checksum
        errors = []

        def work(k):
            try:
                for a in starts[k::threads]:
                    part = parr.slice(a, min(bsz, n - a))
                    rects = G.compute_rects(part)
                    cb, cp = ix.probe(rects)
                    acc[k][1] += len(cb)
                    kb, kp = rf.refine(part, "contains", cb, cp)
                    acc[k][0] += len(kb)
                    acc[k][2] = (acc[k][2] + pair_checksum(kb, kp, a)) % (1 << 64)
            except BaseException as e:
                errors.append(e)

        th = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        t = time.time()
        for x in th:
            x.start()
        for x in th:
            x.join()
        ms = (time.time() - t) * 1000
        if errors:
            raise errors[0]
        kept = sum(a[0] for a in acc)
        res = dict(threads=threads, batch_rows=bsz, batches=len(starts), ms=ms, pairs=int(kept),
                   candidates=int(sum(a[1] for a in acc)), pairs_per_s=kept / (ms / 1000))
        log(f"host-inclusive reference-ABI flow, {threads} thread(s): {ms:.1f} ms, {kept} pairs in "
            f"{len(starts)} batches")
        return res, kept, sum(a[2] for a in acc) % (1 << 64)

    abi, kept1, sum1 = abi_flow(1)
    abi_mt, kept_t, sum_t = abi_flow(args.abi_threads)
    # the pipelined host-column join: batches copied in through pinned
    # blocks while other batches join (SjhipJoinHostStream)
    t = time.time()
    sb, sp = G.join_host_stream(ix, rf, parr, "contains", batch_rows=args.stream_batch, slots=args.stream_slots,
                                refine_order=-1 if args.autotune else args.refine_order)
    stream_ms = (time.time() - t) * 1000
    stream_sum = pair_checksum(sb, sp)
    stream = dict(batch_rows=args.stream_batch, slots=args.stream_slots, ms=stream_ms, pairs=int(len(sb)),
                  pairs_per_s=len(sb) / (stream_ms / 1000))
    log(f"host-inclusive pipelined join ({args.stream_slots} slots of {args.stream_batch} rows): {stream_ms:.1f} ms")
    same = (fused["pairs"] == pairs == kept1 == kept_t == len(sb) and fused_sum == sum1 == sum_t == stream_sum)
    del sb, sp
    return dict(fused=fused, pipelined=stream, reference_abi=abi, reference_abi_threads=abi_mt, same_pairs=bool(same))


SERVE_MIX = (("intersects", 0.0), ("dwithin", 0.05))


def serving_measurement(G, ix, rf, serve, args):
    """Latency of probe requests (log-uniform 1k..10M clustered points,
    ST_Intersects and ST_DWithin(0.05) alternating) arriving as a Poisson
    process against the resident polygon index.  Each arrival rate is run
    twice: through the request loop with `slots` concurrent requests, and
    one request at a time (a single slot).  Latency = finish - arrival."""
    import threading

    sizes = serve["sizes"]
    dev = G.DeviceWkb.from_host(serve["col"])
    starts = np.concatenate([[0], np.cumsum(sizes)])
    reqs = [(dev.slice(starts[k], starts[k + 1]),) + SERVE_MIX[k % len(SERVE_MIX)] for k in range(len(sizes))]
    # warm-up (code objects, memory pool) and the one-at-a-time service time
    for q, pred, d in reqs[:4]:
        G.join_device(ix, rf, q, pred, distance=d, sort_output=False).release()
    t = time.time()
    pairs = 0
    for q, pred, d in reqs:
        r = G.join_device(ix, rf, q, pred, distance=d, sort_output=False)
        pairs += r.size
        r.release()
    service_s = time.time() - t
    capacity = len(reqs) / service_s
    log(f"serving: one-at-a-time service {service_s * 1000:.1f} ms for {len(reqs)} requests ({pairs} pairs)")

    def run(rate, slots, seed, graphs=False):
        rng = np.random.default_rng(seed)
        gaps = rng.exponential(1.0 / rate, size=len(reqs))
        sv = G.GpuSpatialServer(ix, rf, slots=slots, graphs=graphs)
        # the server is closed (its queue drained, its workers joined) before
        # anything leaves this function, an exception included: the request
        # columns must outlive every join still running on the device
        try:
            if graphs:
                # capture every (slot, predicate, size bucket) graph before timing
                for q, pred, d in reqs:
                    for _ in range(slots):
                        sv.wait(sv.submit(q, pred, distance=d, sort_output=False))[0].release()
            ids = []
            t0 = time.perf_counter()
            due = 0.0
            for (q, pred, d), g in zip(reqs, gaps):
                due += g
                wait = due - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                ids.append(sv.submit(q, pred, distance=d, sort_output=False))
            lat, first, last, npairs = [], None, None, 0
            for rid in ids:
                res, tm = sv.wait(rid)
                npairs += res.size
                res.release()
                lat.append(tm["done_ms"] - tm["submit_ms"])
                first = tm["submit_ms"] if first is None else min(first, tm["submit_ms"])
                last = tm["done_ms"] if last is None else max(last, tm["done_ms"])
        finally:
            sv.close()
        span = (last - first) / 1000
        lat = np.array(lat)
        return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "throughput_rps": len(reqs) / span,
                "points_per_s": float(sizes.sum()) / span, "pairs_per_s": npairs / span}

    rates = []
    for rel in (0.5, 0.9, 1.5):
        rate = rel * capacity
        row = {"rel_to_one_at_a_time_capacity": rel, "rate_rps": rate,
               "batched": run(rate, args.serve_slots, 7 + int(rel * 10)),
               "one_at_a_time": run(rate, 1, 7 + int(rel * 10))}
        log(f"serving rate {rate:.1f}/s: batched p50 {row['batched']['p50_ms']:.2f} p99 {row['batched']['p99_ms']:.2f} ms;"
            f" one-at-a-time p50 {row['one_at_a_time']['p50_ms']:.2f} p99 {row['one_at_a_time']['p99_ms']:.2f} ms")
        rates.append(row)
    # captured-graph replay for requests of up to 65536 points (larger ones
    # take the eager path inside the same server), at the 0.9x rate
    rel = 0.9
    graphs = run(rel * capacity, args.serve_slots, 7 + int(rel * 10), graphs=True)
    graphs["rel_to_one_at_a_time_capacity"] = rel
    log(f"serving rate {rel * capacity:.1f}/s with captured graphs: p50 {graphs['p50_ms']:.2f} "
        f"p99 {graphs['p99_ms']:.2f} ms")
    del reqs, dev
    return {"requests": len(sizes), "points": int(sizes.sum()), "sizes": "log-uniform 1k..10M clustered points",
            "mix": "ST_Intersects / ST_DWithin(0.05) alternating", "slots": args.serve_slots,
            "memory_budget": "90% of free HBM", "one_at_a_time_service_ms": service_s * 1000,
            "one_at_a_time_capacity_rps": capacity, "rates": rates,
            "batched_graphs": graphs,
            "p50_ms": rates[1]["batched"]["p50_ms"], "p99_ms": rates[1]["batched"]["p99_ms"]}


def sources_digest():
    """First 16 hex digits of the SHA-256 over the library's HIP sources
    (sedona-db_amd/csrc, files in name order): the counter summaries record
    the digest of the sources they were measured on."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "sedona-db_amd", "csrc")
    for name in sorted(os.listdir(d)):
        h.update(name.encode())
        with open(os.path.join(d, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def read_traffic(path):
    """HBM bytes per launch of each profiled kernel: lines 'traffic <kernel> <bytes>'
    of the committed counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration), and a note on where they come from.  The
    summary's 'sources <digest>' line must match the sources being run, else
    no traffic is reported (the counters describe other kernels)."""
    out, digest = {}, None
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic":
                out[f[1]] = int(f[2])
            elif len(f) == 2 and f[0] == "sources":
                digest = f[1]
    rel = os.path.relpath(path, ROOT)
    now = sources_digest()
    if digest != now:
        return {}, f"none: {rel} was measured on sources {digest}, these are {now}"
    return out, f"{rel} (rocprofv3 FETCH_SIZE + WRITE_SIZE, sources {now})"


def calibrate_copy(nbytes=1 << 30, reps=20):
    from sjhip import gpuspatial as G

    ms = G.copy_kernel_ms(nbytes, reps)
    print(json.dumps({"copy_bytes": nbytes, "ms": ms, "hbm_gbs": 2 * nbytes / (ms / 1e3) / 1e9}), flush=True)


def strip_pairs_mask(strip, gb, gp):
    """Local (strip-row) keys (probe << 32 | build) of the GPU pairs whose
    build and probe rows both lie in the strip sample."""
    bmap = np.full(int(gb.max()) + 1 if len(gb) else 1, -1, np.int64)
    sel_b = strip["brows"]
    bmap[sel_b[sel_b < len(bmap)]] = np.nonzero(sel_b < len(bmap))[0]
    pset = np.full(int(gp.max()) + 1 if len(gp) else 1, -1, np.int64)
    sel_p = strip["prows"]
    pset[sel_p[sel_p < len(pset)]] = np.nonzero(sel_p < len(pset))[0]
    lb, lp = bmap[gb], pset[gp]
    keep = (lb >= 0) & (lp >= 0)
    return (lp[keep].astype(np.uint64) << 32) | lb[keep].astype(np.uint64)


def cut_strip(D, wl, build, probe, frac):
    ext = wl["extent"]
    x0 = 0.5 * ext
    w = frac * ext
    brows = np.nonzero((wl["centers"][:, 0] >= x0) & (wl["centers"][:, 0] < x0 + w))[0]
    prows = np.nonzero((wl["probe_xy"][:, 0] >= x0) & (wl["probe_xy"][:, 0] < x0 + w))[0]
    return dict(brows=brows, prows=prows, bcol=D.take_rows(build, brows), pcol=D.take_rows(probe, prows), x0=x0, w=w)


def run_dwithin_walks(args):
    """50M uniform points x 5M random-walk linestrings (2..256 vertices,
    steps N(0, 0.5^2)), ST_DWithin(line, point, 0.05).  One step = the 50M
    probe points through the device-resident join in batches of
    --probe-batch points."""
    from sjhip import datagen as D

    dist = 0.05
    n_points = int(round(50_000_000 * args.scale))
    n_lines = int(round(5_000_000 * args.scale))
    t0 = time.time()
    wl = D.workload_dwithin_walks_parallel(n_points, n_lines, seed=args.seed + 1, workers=args.workers)
    build, probe = wl["build"], wl["probe"]
    log(f"generated {n_lines} lines ({int(wl['counts'].sum())} vertices, {build.offsets[-1] / 1e9:.1f} GB WKB) and "
        f"{n_points} points in {time.time() - t0:.1f}s")
    # the CPU port refines about 1.3e5 pairs/s on 16 threads: a strip of
    # 0.3 % of the extent (about 5e6 pairs) keeps its leg near 40 s
    strip = None if args.no_cpu else cut_strip(D, wl, build, probe, args.cpu_strip / 40)
    if strip is not None:
        log(f"cpu strip: {len(strip['brows'])} lines, {len(strip['prows'])} points")

    from sjhip import gpuspatial as G

    opts = G.GpuSpatialOptions(device_id=0)
    tb = time.time()
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema("Z")
    bdev = G.DeviceWkb.from_host(build)
    rf.push_build_device(bdev)
    rf.finish_building()
    del bdev
    ix = G.GpuSpatialIndex(opts)
    ix.build_from_refiner(rf)
    build_ms = (time.time() - tb) * 1000
    pdev = G.DeviceWkb.from_host(probe)
    log(f"build side resident: {build_ms:.0f} ms")
    bsz = args.probe_batch

    def step(precision, keep_pairs=False, per_batch=None):
        tot = np.zeros(5, np.int64)
        stage = np.zeros(6)
        kept = []
        for a in range(0, n_points, bsz):
            r = G.join_device(ix, rf, pdev.slice(a, min(n_points, a + bsz)), "dwithin", distance=dist,
                              precision=precision, sort_output=per_batch is not None,
                              refine_order=args.refine_order, profile=True)
            tot += (r.size, r.candidates, r.segment_tests, r.fallback_tests, r.refine_candidates)
            stage += np.array(list(r.res.stage_ms))
            if keep_pairs:
                gb, gp = r.to_host()
                kept.append(strip_pairs_mask(strip, gb, (gp + a).astype(np.int64)))
            if per_batch is not None:
                per_batch(a, r)
            r.release()
        return tot, stage, kept

    for i in range(args.warmup):
        tot, stage, _ = step(args.precision)
        log(f"warmup {i}: {tot[0]} pairs, {tot[1]} candidates, stages {list(stage)}")
    t = time.time()
    for i in range(args.steps):
        tot, stage, _ = step(args.precision)
    ms_per_step = (time.time() - t) * 1000 / args.steps
    pairs, cands, tests, scans, walked = (int(x) for x in tot)
    log(f"timed: {ms_per_step:.1f} ms/step, {pairs} pairs, {cands} candidates, stages(ms) {list(stage)}")
    plain = None
    if not args.no_plain:
        t1 = time.time()
        ptot, pstage, _ = step(1)
        plain = dict(ms_per_step=(time.time() - t1) * 1000, pairs=int(ptot[0]), refine_ms=float(pstage[3]))
        # the pair sets, batch by batch, both sorted by (probe, build) on the device
        mine = {}
        step(args.precision, per_batch=lambda a, r: mine.__setitem__(a, r.to_host()))
        diff = [0]

        def compare(a, r):
            pb, pp = r.to_host()
            gb, gp = mine.pop(a)
            diff[0] += pair_set_diff(gb, gp, pb, pp)

        step(1, per_batch=compare)
        plain["pair_set_diff"] = diff[0]
        log(f"plain fp64: {plain}")
    cpu = None
    if strip is not None:
        _, _, kept = step(args.precision, keep_pairs=True)
        gk = np.sort(np.concatenate(kept)) if kept else np.zeros(0, np.uint64)
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {len(strip['brows'])} lines x {len(strip['prows'])} points on {threads} threads")
        tc = time.time()
        cb, cp = O.join(O.Store(strip["bcol"].as_tuple()), O.Store(strip["pcol"].as_tuple()), "dwithin",
                        distance=dist, nthreads=threads)
        secs = time.time() - tc
        ck = np.sort((cp.astype(np.uint64) << 32) | cb.astype(np.uint64))
        # the distances behind the decisions: the library's FP64 distance of
        # every sampled pair against the oracle's
        gdist = rf.pair_distances(strip["pcol"].to_arrow(large=True), strip["brows"][cb], cp)
        odist = O.pair_distances(O.Store(strip["bcol"].as_tuple()), O.Store(strip["pcol"].as_tuple()), cb, cp,
                                 nthreads=threads)
        rel = np.abs(gdist - odist) / np.maximum(np.abs(odist), 1e-300)
        cpu = {"value": len(cb) / secs, "unit": "pairs/s", "cores": threads, "kind": "port",
               "max_distance_rel_error": float(rel.max()) if len(rel) else 0.0,
               "sample": f"vertical strip x in [{strip['x0']:.1f}, {strip['x0'] + strip['w']:.1f}) of line starts "
                         f"and points: {len(strip['brows'])} lines x {len(strip['prows'])} points, {len(cb)} pairs, "
                         f"{secs:.2f} s (WKB parse + grid filter + geo-semantics distance refine)",
               "gpu_pairs_in_sample": int(len(gk)), "pair_set_diff": int(len(np.setxor1d(gk, ck)))}
    filter_bytes = 16 * n_points + 8 * cands
    out = {
        "metric": "refined join pairs/s (ST_DWithin 0.05, 50M points x 5M random-walk lines, device-resident)",
        "value": pairs / (ms_per_step / 1000), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == 1 else "f32+f64",
        "data": "synthetic (seeded, generated on the host)",
        "config": {"workload": "dwithin_walks", "probe_points": n_points, "build_lines": n_lines,
                   "build_vertices": int(wl["counts"].sum()), "predicate": "ST_DWithin", "distance": dist,
                   "probe_batch": bsz, "parallelism": "single-gpu"},
        "pairs_per_step": pairs, "candidates_per_step": cands, "walked_candidates_per_step": walked,
        "probe_points_per_s": n_points / (ms_per_step / 1000),
        "stage_ms": {"decode_rects": stage[0], "morton": stage[1], "filter": stage[2], "group_by_line": stage[5],
                     "refine": stage[3]},
        "roofline_filter": {"bound": "hbm", "achieved": filter_bytes / (stage[2] / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": filter_bytes / (stage[2] / 1e3) / 1e9 / HBM_PEAK_GBS},
        "precision": "fp32-certified+fp64-fallback" if args.precision == 0 else "fp64-plain",
        "segment_tests": tests, "fp64_reevaluations": scans,
        "fallback_rate": scans / max(1, tests),
        "build_ms": build_ms, "plain_fp64": plain, "cpu_baseline": cpu,
    }
    if plain:
        out["speedup_vs_plain_fp64"] = plain["ms_per_step"] / ms_per_step
    print(json.dumps(out), flush=True)
    D.release_shared(wl["keep"])


def run_knn(args):
    """KNN join, k = 8: 20M query points x 200M object points (the 100M-point
    mixture), exact neighbour lists with ties by id, FP64 distances.  One
    step = the KNN kernel over all queries (device-resident queries and
    outputs; the object tree is built once, with the build side)."""
    from sjhip import datagen as D

    k = 8
    nq = int(round(20_000_000 * args.scale))
    no = int(round(200_000_000 * args.scale))
    t0 = time.time()
    wl = D.workload_knn_parallel(nq, no, seed=args.seed + 3, workers=args.workers)
    log(f"generated {no} objects and {nq} queries in {time.time() - t0:.1f}s")
    from sjhip import gpuspatial as G

    opts = G.GpuSpatialOptions(device_id=0)
    tb = time.time()
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema("Z")
    bdev = G.DeviceWkb.from_host(wl["build"])
    rf.push_build_device(bdev)
    rf.finish_building()
    del bdev
    qdev = G.DeviceWkb.from_host(wl["probe"])
    ids, dd, ms0 = rf.knn_device(qdev, k)  # builds the object tree (resident from here on)
    build_ms = (time.time() - tb) * 1000
    log(f"build side resident (incl. object tree): {build_ms:.0f} ms; first call kernel {ms0:.1f} ms")
    ids.free()
    dd.free()
    for i in range(args.warmup):
        a, b, ms = rf.knn_device(qdev, k)
        a.free()
        b.free()
    t = time.time()
    kms = 0.0
    for i in range(args.steps):
        ids, dd, ms = rf.knn_device(qdev, k)
        kms += ms
        if i < args.steps - 1:
            ids.free()
            dd.free()
    ms_per_step = (time.time() - t) * 1000 / args.steps
    kms /= args.steps
    log(f"timed: {ms_per_step:.1f} ms/step, kernel {kms:.1f} ms")
    # exactness on a sample of queries
    rng = np.random.default_rng(args.seed)
    sample = np.sort(rng.choice(nq, size=min(args.knn_sample, nq), replace=False))
    gi = ids.download(np.uint32, nq * k).reshape(nq, k)[sample]
    gd = dd.download(np.float64, nq * k).reshape(nq, k)[sample]
    ids.free()
    dd.free()
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    tc = time.time()
    sq = D.take_rows(wl["probe"], sample)
    oid, od = O.knn(O.Store(wl["build"].as_tuple()), O.Store(sq.as_tuple()), k, nthreads=threads)
    secs = time.time() - tc
    rel = np.abs(gd - od) / np.maximum(np.abs(od), 1e-300)
    cpu = {"value": len(sample) / secs, "unit": "queries/s", "cores": threads, "kind": "port",
           "sample": f"{len(sample)} random queries against all {no} objects (grid build included), {secs:.2f} s"}
    # neighbours examined: a query's k results; bytes: 16 B per query + 12 B per neighbour written
    out_bytes = 16 * nq + 12 * nq * k
    out = {
        "metric": "KNN join queries/s (k=8, 20M queries x 200M clustered points, device-resident)",
        "value": nq / (ms_per_step / 1000), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded, generated on the host)",
        "config": {"workload": "knn", "queries": nq, "objects": no, "k": k, "parallelism": "single-gpu"},
        "kernel_ms": kms, "neighbours_per_s": nq * k / (kms / 1000),
        "roofline": {"bound": "hbm", "achieved": out_bytes / (kms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": out_bytes / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "build_ms": build_ms, "cpu_baseline": cpu,
        "sample_exact_ids": bool(np.array_equal(gi, oid)),
        "sample_max_rel_distance_error": float(rel.max()) if rel.size else 0.0,
    }
    print(json.dumps(out), flush=True)
    D.release_shared(wl["keep"])


def run_poly_poly(args):
    """Polygon-polygon ST_Intersects: two sets of clustered star polygons
    (vertex counts log-uniform 4..1024, radius lognormal median 0.8,
    centres from the 100M-point mixture), at --scale of 5M x 5M."""
    from sjhip import datagen as D

    na = int(round(5_000_000 * args.scale))
    t0 = time.time()
    a, b = D.workload_poly_poly(na, na, seed=args.seed + 2)
    log(f"generated 2 x {na} polygons ({a.offsets[-1] / 1e9:.2f} + {b.offsets[-1] / 1e9:.2f} GB WKB) "
        f"in {time.time() - t0:.1f}s")
    strip = None
    if not args.no_cpu:
        import oracle as O
        ra, rb = O.Store(a.as_tuple()).rects(True), O.Store(b.as_tuple()).rects(True)
        ext = 1000.0 * np.sqrt(na / 5_000_000)
        x0, w = 0.5 * ext, args.cpu_strip * ext / 8
        ca = 0.5 * (ra[:, 0] + ra[:, 2])
        cb = 0.5 * (rb[:, 0] + rb[:, 2])
        brows = np.nonzero((ca >= x0) & (ca < x0 + w))[0]
        prows = np.nonzero((cb >= x0) & (cb < x0 + w))[0]
        strip = dict(brows=brows, prows=prows, bcol=D.take_rows(a, brows), pcol=D.take_rows(b, prows), x0=x0, w=w)
    from sjhip import gpuspatial as G

    opts = G.GpuSpatialOptions(device_id=0)
    rf = G.GpuSpatialRefiner(opts)
    rf.init_build_schema("Z")
    adev = G.DeviceWkb.from_host(a)
    rf.push_build_device(adev)
    rf.finish_building()
    del adev
    ix = G.GpuSpatialIndex(opts)
    ix.build_from_refiner(rf)
    pdev = G.DeviceWkb.from_host(b)

    # probe polygons per join call.  The clusters keep their width at every
    # scale, so candidates grow about as the square of the scale (5.5e8 at
    # 0.1 in one call); slices of the probe column keep each call's
    # candidates far below 2^32 (the pair-set comparisons run on one-call
    # scales)
    nb = na if args.poly_batch <= 0 else min(args.poly_batch, na)
    batched = nb < na
    if batched and strip is not None:
        log("probe side in slices: the CPU strip comparison runs at one-call scales (--scale 0.1)")
        strip = None

    def step():
        if not batched:
            r = G.join_device(ix, rf, pdev, "intersects", sort_output=False, refine_order=args.refine_order,
                              profile=True)
            return r, [int(r.size), int(r.candidates), int(r.refine_candidates)], list(r.res.stage_ms)
        tot, stage = [0, 0, 0], [0.0] * 6
        for s0 in range(0, na, nb):
            r = G.join_device(ix, rf, pdev.slice(s0, min(na, s0 + nb)), "intersects", sort_output=False,
                              refine_order=args.refine_order, profile=True)
            for k, v in enumerate((r.size, r.candidates, r.refine_candidates)):
                tot[k] += int(v)
            stage = [a + float(b) for a, b in zip(stage, r.res.stage_ms)]
            r.release()
            if (s0 // nb) % 10 == 9:
                log(f"  {min(na, s0 + nb)} of {na} probe polygons joined, {tot[0]} pairs so far")
        return None, tot, stage

    for i in range(args.warmup):
        r, tot, stage = step()
        log(f"warmup {i}: {tot[0]} pairs, {tot[1]} candidates, stages {stage}")
        if r is not None:
            r.release()
    t = time.time()
    for i in range(args.steps):
        r, tot, stage = step()
        if i < args.steps - 1 and r is not None:
            r.release()
    ms_per_step = (time.time() - t) * 1000 / args.steps
    pairs, cands, walked = tot
    log(f"timed: {ms_per_step:.1f} ms/step, {pairs} pairs, {cands} candidates ({walked} left to the wavefront "
        f"kernel by the raster pass), stages {stage}")
    # the certified kernels against the inline-precision kernel alone
    # (precision 1: every open candidate to k_refine_polypoly), both pair
    # sets sorted by (probe, build) on the device
    plain = None
    if not args.no_plain and not batched:
        r.release()
        r = G.join_device(ix, rf, pdev, "intersects", sort_output=True, refine_order=args.refine_order)
        gb, gp = r.to_host()
        r.release()
        t1 = time.time()
        r1 = G.join_device(ix, rf, pdev, "intersects", precision=1, sort_output=True, refine_order=args.refine_order,
                           profile=True)
        plain_ms = (time.time() - t1) * 1000
        pb, pp = r1.to_host()
        plain = {"ms_per_step": plain_ms, "refine_ms": float(r1.res.stage_ms[3]), "pairs": int(r1.size),
                 "pair_set_diff": pair_set_diff(gb, gp, pb, pp)}
        r1.release()
        del pb, pp
        log(f"inline-precision kernel alone: {plain_ms:.1f} ms, pair-set diff {plain['pair_set_diff']}")
    cpu = None
    if strip is not None:
        if plain is None:
            gb, gp = r.to_host()
        cpu = cpu_baseline(strip, gb, gp, args, predicate="intersects", what="polygons")
    if plain is None and r is not None:
        r.release()
    out = {
        "metric": "refined join pairs/s (ST_Intersects, clustered star polygons x star polygons, device-resident)",
        "value": pairs / (ms_per_step / 1000), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded, generated on the host)",
        "config": {"workload": "poly_poly", "polygons_each_side": na, "scale_of_5M": args.scale,
                   "predicate": "ST_Intersects", "probe_polygons_per_call": nb, "parallelism": "single-gpu"},
        "pairs_per_step": pairs, "candidates_per_step": cands, "walked_candidates_per_step": walked,
        "stage_ms": {"decode_rects": stage[0], "morton": stage[1], "filter": stage[2], "group_by_polygon": stage[5],
                     "refine": stage[3]},
        "inline_precision_kernel": plain,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(strip, gb, gp, args, predicate="contains", what="polygons"):
    """The reference's TG refine (oracle/_ref) behind a grid filter on the
    strip sample; pairs/s on this host's cores.  Also checks that the GPU's
    full-size pairs restricted to the strip equal the CPU pairs."""
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    kind = "reference" if O.ref_available() else "port"
    bcol, pcol = strip["bcol"], strip["pcol"]
    t = time.time()
    if kind == "reference":
        st_b, st_p = O.Store(bcol.as_tuple()), O.Store(pcol.as_tuple())
        br, pr = st_b.rects(True), st_p.rects(True)
        tb = O.TgColumn(bcol.as_tuple(), index=1, nthreads=threads)
        tp = O.TgColumn(pcol.as_tuple(), index=1 if what == "polygons" and predicate == "intersects" else 0,
                        nthreads=threads)
        cb, cp = O.ref_join(tb, tp, br, pr, predicate, nthreads=threads)
    else:
        st_b, st_p = O.Store(bcol.as_tuple()), O.Store(pcol.as_tuple())
        cb, cp = O.join(st_b, st_p, predicate, nthreads=threads)
    secs = time.time() - t
    res = {"value": len(cb) / secs, "unit": "pairs/s", "cores": threads, "kind": kind,
           "sample": f"vertical strip x in [{strip['x0']:.1f}, {strip['x0'] + strip['w']:.1f}): "
                     f"{len(strip['brows'])} polygons x {len(strip['prows'])} {'points' if predicate == 'contains' else 'polygons'}, {len(cb)} pairs, "
                     f"{secs:.2f} s (WKB parse + filter + refine)"}
    if gb is not None:
        bmap = np.full(int(gb.max()) + 1 if len(gb) else 1, -1, np.int64)
        sel_b = strip["brows"]
        bmap[sel_b[sel_b < len(bmap)]] = np.nonzero(sel_b < len(bmap))[0]
        pset = np.full(int(gp.max()) + 1 if len(gp) else 1, -1, np.int64)
        sel_p = strip["prows"]
        pset[sel_p[sel_p < len(pset)]] = np.nonzero(sel_p < len(pset))[0]
        lb = bmap[gb]
        lp = pset[gp]
        keep = (lb >= 0) & (lp >= 0)
        gk = np.sort((lp[keep].astype(np.uint64) << 32) | lb[keep].astype(np.uint64))
        ck = np.sort((cp.astype(np.uint64) << 32) | cb.astype(np.uint64))
        res["gpu_pairs_in_sample"] = int(len(gk))
        res["pair_set_diff"] = int(len(np.setxor1d(gk, ck)))
    return res


if __name__ == "__main__":
    main()
