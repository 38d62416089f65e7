"""Benchmark of the exact painter on one MI355X: receiver-source paths/s and Lden cells/s.

Workload: the tile-scale scene — 4096x4096 receivers at 10 m (4 m above a diamond-square DEM of
30 m lattice and 200 m relief), 200,000 point sources (road 0.05 m / rail 0.5 m, band levels
N(85, 8) dB, day/evening/night), 50,000 polygon buildings (4-12 edges, heights U[3, 60] m), every
source within the 2 km cutoff of each receiver. One step paints a window of it (default the
central 1024x1024 receivers, 1/16 of the grid, about 1.6e9 paths; --window 4096 paints the whole
grid) in one persistent launch; per-path work is the whole scene's. The scene, sources and
receivers are resident in HBM before the timed region.

Printed: one JSON line with the metric, the roofline of the painting kernel (algorithmic FP32
FLOP per launch over its measured duration), the FP64 CPU oracle timed on a bounded sample of the
same workload on the host's cores, and the painted Lden's error against that oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "receiver-source paths/s & Lden cells/s at ≤0.05 dB; surrogate vs exact, 1 GPU"
# Worker threads of the CPU baseline: the host's cores, at most the CPU share of one GPU's box.
CPU_SHARE = 16


def host_cpu():
    """The host's CPU model and core count, stated next to every CPU figure."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((line.split(":", 1)[1].strip() for line in f if line.startswith("model name")), model)
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def median_seconds(run, repeats=3):
    """The median wall time of `repeats` runs of `run()` (warm: the inputs are already built)."""
    times = []
    for _ in range(repeats):
        t = time.time()
        run()
        times.append(time.time() - t)
    return sorted(times)[len(times) // 2]


def threaded_chunks(fn, n, threads):
    """fn(begin, end) over `n` items split into `threads` chunks run on a thread pool (the oracle's C
    calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    bounds = [(n * k // threads, n * (k + 1) // threads) for k in range(threads)]
    with ThreadPoolExecutor(threads) as pool:
        list(pool.map(lambda b: fn(*b), bounds))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="side of the scene relative to 4096 receivers")
    ap.add_argument("--window", type=int, default=1024, help="side of the painted window (receivers)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-loudness", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--band-sources", type=int, default=200, help="sources whose paths measure the surrogate's band error")
    ap.add_argument("--no-flights", action="store_true")
    ap.add_argument("--no-doc29", action="store_true")
    ap.add_argument("--doc29-scale", type=float, default=1.0, help="side of the Doc 29 grid relative to 2048")
    ap.add_argument("--doc29-bands", type=int, default=2, help="launches (bands of rows) the Doc 29 grid takes")
    ap.add_argument("--loudness-cells", type=int, default=16 * 1024 * 1024)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the painter has no multi-GPU path (tiles of receivers shard "
              "across GPUs with no exchange). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from quietmap_amd import abi, hip, scenes
    from quietmap_amd.painter import RECEIVER_BYTES, Gpu, lden_db

    t0 = time.time()
    sc = scenes.tile_scene(scale=args.scale)
    gx, gy = sc.grid_shape
    points, nx, ny = sc.window(args.window)
    n = nx * ny
    print(f"[bench] scene {nx}x{ny} receivers, {sc.sources.size} sources, "
          f"{sc.scene.obstacles.outlines.size} buildings, generated in {time.time() - t0:.1f} s", file=sys.stderr,
          flush=True)

    gpu = Gpu(0)
    square = gpu.upload(sc.scene, sc.sources)
    stream = hip.Stream()
    d_points = hip.DeviceBuffer.from_array(points)
    d_receivers = hip.DeviceBuffer(n * RECEIVER_BYTES)
    d_energy = hip.DeviceBuffer(n * 3 * 8)
    d_paths = hip.DeviceBuffer(n * 4)
    d_failures = hip.DeviceBuffer(n * 4)
    gpu.receivers_device(square, d_points.ptr, n, d_receivers.ptr, stream.handle)

    def step():
        gpu.paint_receivers_device(square, d_receivers.ptr, nx, ny, sc.cutoff_m, d_energy.ptr, d_paths.ptr,
                                   d_failures.ptr, stream.handle)

    for k in range(args.warmup):
        step()
        stream.synchronize()
        print(f"[bench] warmup {k + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    hip.synchronize()
    times = []
    t_start = time.time()
    for k in range(args.steps):
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        step()
        e1.record(stream)
        e1.synchronize()
        times.append(e1.seconds_since(e0))
        print(f"[bench] step {k + 1}/{args.steps}: {times[-1]:.3f} s", file=sys.stderr, flush=True)
    hip.synchronize()
    wall = time.time() - t_start
    path_counts = d_paths.to_array(np.uint32, n).astype(np.int64)
    paths = int(path_counts.sum())
    failures = int(d_failures.to_array(np.uint32, n).astype(np.int64).sum())
    energy = d_energy.to_array(np.float64, (n, 3))
    kernel_s = sum(times) / len(times)
    value = paths / kernel_s
    cells = n / kernel_s

    # The surrogate leg: the same window painted with the network inside its envelope.
    sur = None
    params_file = os.path.join(ROOT, "quietmap_amd", "surrogate_params.npy")
    if not args.no_surrogate and os.path.exists(params_file):
        from quietmap_amd import surrogate

        net = surrogate.Mlp(gpu, np.load(params_file), max_batch=1)
        d_learned = hip.DeviceBuffer(n * 4)
        d_energy_s = hip.DeviceBuffer(n * 3 * 8)

        def surrogate_step():
            surrogate.paint_receivers_device(gpu, square, net, d_receivers.ptr, nx, ny, sc.cutoff_m, d_energy_s.ptr,
                                             d_paths.ptr, d_failures.ptr, d_learned.ptr, stream.handle)

        surrogate_step()
        stream.synchronize()
        st = []
        for k in range(args.steps):
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            surrogate_step()
            e1.record(stream)
            e1.synchronize()
            st.append(e1.seconds_since(e0))
            print(f"[bench] surrogate step {k + 1}/{args.steps}: {st[-1]:.3f} s", file=sys.stderr, flush=True)
        s_time = sum(st) / len(st)
        s_paths = int(d_paths.to_array(np.uint32, n).astype(np.int64).sum())
        learned = int(d_learned.to_array(np.uint32, n).astype(np.int64).sum())
        lden_exact = lden_db(energy)
        lden_sur = lden_db(d_energy_s.to_array(np.float64, (n, 3)))
        err_s = np.abs(lden_sur - lden_exact)
        err_s = err_s[np.isfinite(err_s)]
        # Per-band error: the network's A_ground + A_dif per band and state against the exact
        # kernel's on in-envelope paths of this scene (sources and receivers it was not trained on).
        samples, (spr, _) = surrogate.generate_samples(gpu, sc, args.band_sources, seed=2, stride=16)
        rows = samples.rows
        d_out = hip.DeviceBuffer(max(rows, 1) * surrogate.OUTPUTS * 4)
        chunk = 16384
        rows_net = surrogate.Mlp(gpu, np.load(params_file), max_batch=chunk)
        for at in range(0, rows, chunk):
            rows_net.infer_device(samples.inputs.ptr + at * surrogate.INPUTS * 4, min(chunk, rows - at),
                                  d_out.ptr + at * surrogate.OUTPUTS * 4)
        hip.synchronize()
        rows_net.close()
        band_err = np.abs(d_out.to_array(np.float32, (rows, surrogate.OUTPUTS)).astype(np.float64) -
                          samples.targets.to_array(np.float32, (rows, surrogate.OUTPUTS)).astype(np.float64))
        del samples, d_out
        sur = {"value": s_paths / s_time, "unit": "paths/s", "ms_per_step": s_time * 1e3,
               "speedup_vs_exact": (s_paths / s_time) / value, "network_paths_fraction": learned / max(s_paths, 1),
               "lden_vs_exact_db": {"mean": float(err_s.mean()), "p99": float(np.quantile(err_s, 0.99)),
                                    "p999": float(np.quantile(err_s, 0.999)), "max": float(err_s.max())},
               "band_attenuation_vs_exact_db": {
                   "paths": int(spr.size), "rows": int(rows), "mean": float(band_err.mean()),
                   "p99": float(np.quantile(band_err, 0.99)), "p999": float(np.quantile(band_err, 0.999)),
                   "max": float(band_err.max()),
                   "within_0_5_db": float((band_err <= 0.5).mean()),
                   "per_band_p99": [float(np.quantile(band_err[:, b], 0.99)) for b in range(surrogate.OUTPUTS)]},
               "targets_met": {"speedup_3x": bool((s_paths / s_time) / value >= 3.0),
                               "lden_0_2_db": bool(err_s.max() <= 0.2),
                               "band_0_5_db": bool(band_err.max() <= 0.5)},
               "mlp_flop_per_inference": 417792}
        net.close()
        print(f"[bench] surrogate {sur['value']:.3e} paths/s ({sur['speedup_vs_exact']:.2f}x exact), Lden vs exact "
              f"mean {sur['lden_vs_exact_db']['mean']:.3f} max {sur['lden_vs_exact_db']['max']:.3f} dB",
              file=sys.stderr, flush=True)

    # The aircraft-box leg: 20,000 boxes over 20 km heard at a 64 x 64 grid of points (every point
    # hearing all 16 tiles), inputs resident on the device, one flight_sums launch a step.
    fl = None
    if not args.no_flights:
        from quietmap_amd import flights

        boxes_f = flights.synthetic_boxes(seed=31, count=20000, extent_m=20000.0)
        side = 64
        jj, ii = np.mgrid[0:side, 0:side]
        xy_f = np.stack([(ii.reshape(-1) + 0.5) * 20000.0 / side, (jj.reshape(-1) + 0.5) * 20000.0 / side], 1)
        boxes_f, points_f, segments_f, n_sums_f = flights.flight_requests(boxes_f, xy_f, extent_m=20000.0)
        handle_f = gpu.flight_boxes(boxes_f)
        d_fp, d_fs = hip.DeviceBuffer.from_array(points_f), hip.DeviceBuffer.from_array(segments_f)
        d_sums_f = hip.DeviceBuffer(n_sums_f * 291 * 8)
        reads = int(segments_f["count"].astype(np.int64).sum())
        ft = []
        for k in range(4):
            hip.check(hip.runtime().hipMemset(d_sums_f.ptr, 0, n_sums_f * 291 * 8))
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            gpu.flight_sums_device(handle_f, d_fp.ptr, 0, d_fs.ptr, segments_f.size, d_sums_f.ptr, stream.handle)
            e1.record(stream)
            e1.synchronize()
            if k:
                ft.append(e1.seconds_since(e0))
        f_time = sum(ft) / len(ft)
        sums_f = d_sums_f.to_array(np.float64, (n_sums_f, 291))
        fl = {"value": reads / f_time, "unit": "box reads/s", "kernel_ms": f_time * 1e3, "boxes": int(boxes_f.size),
              "points": int(points_f.size), "segments": int(segments_f.size), "kernel": "flight_sums_kernel"}
        if not args.no_cpu:
            from oracle import oracle

            oracle.use_native()
            threads = args.cpu_threads or min(os.cpu_count() or 1, CPU_SHARE)
            pick = np.arange(0, points_f.size, 16)  # 256 points' segments
            keep = np.isin(segments_f["point"], pick)
            sel = segments_f[keep]
            want_f = oracle.flight_sums(boxes_f, points_f, None, sel, n_sums_f)

            def flights_cpu(begin, end):
                oracle.flight_sums(boxes_f, points_f, None, sel[begin:end], n_sums_f)

            reads_cpu = int(sel["count"].astype(np.int64).sum())
            cpu_f = reads_cpu / median_seconds(lambda: threaded_chunks(flights_cpu, sel.size, threads))
            one_f = reads_cpu / median_seconds(lambda: flights_cpu(0, sel.size))
            fl["cpu_baseline"] = {"value": cpu_f, "unit": "box reads/s", "cores": threads, "kind": "port",
                                  "one_thread": one_f, **host_cpu(), "median_of": 3, "build": oracle.NATIVE_FLAGS,
                                  "sample": f"the {int(keep.sum())} segments of {pick.size} points, FP64 C++ "
                                            "restatement of the reference's CPU path"}
            e_g, e_w = sums_f[pick, :3], want_f[pick, :3]
            ok = e_w > 0
            fl["max_abs_db_vs_oracle"] = float(np.abs(10 * np.log10(e_g[ok] / e_w[ok])).max())
            fl["classes_equal"] = bool(np.array_equal(sums_f[pick, 3 + 144:] > 0, want_f[pick, 3 + 144:] > 0))
        handle_f.close()
        print(f"[bench] flights {fl['value']:.3e} box reads/s", file=sys.stderr, flush=True)

    # The Doc 29 leg: the aircraft configuration's 2048 x 2048 grid at 50 m (20,000 tracks of 200
    # segments, 100 classes, 16 km reach) in bands of rows, one aircraft_kernel launch a band; Lden,
    # N65/N70 and the 0.5 dB LAmax histogram per receiver, checked on sampled receivers.
    d29 = None
    if not args.no_doc29:
        from quietmap_amd import scenes as scenes_a

        ac = scenes_a.aircraft_scene(scale=args.doc29_scale)
        nxa, nya = ac.grid_shape
        na = nxa * nya
        energy_a = np.zeros((na, 3))
        n65_a = np.zeros(na, np.uint32)
        n70_a = np.zeros(na, np.uint32)
        hist_a = np.zeros((na, abi.HIST_BINS), np.uint32)
        bands = max(1, min(args.doc29_bands, nya))
        edges = [nya * k // bands for k in range(bands + 1)]
        t_a = time.time()
        for k in range(bands):
            r0, r1 = edges[k], edges[k + 1]
            sub = scenes_a.AircraftScene(ac.points_xy[r0 * nxa:r1 * nxa], (nxa, r1 - r0), ac.classes, ac.tracks,
                                         ac.vertices, ac.receiver_altitude_m, ac.reach_m)
            tb = time.time()
            r = gpu.aircraft_grid(sub)
            sl = slice(r0 * nxa, r1 * nxa)
            energy_a[sl], n65_a[sl], n70_a[sl], hist_a[sl] = r["period_energy"], r["n65"], r["n70"], r["hist"]
            print(f"[bench] doc29 band {k + 1}/{bands}: rows {r0}-{r1}, {time.time() - tb:.2f} s", file=sys.stderr,
                  flush=True)
        wall_a = time.time() - t_a
        d29 = {"workload": f"Doc 29 grid {nxa}x{nya} at 50 m, {ac.tracks.size} tracks of 200 segments, "
                           f"{ac.classes.size} classes, {ac.reach_m:.0f} m reach",
               "receivers": na, "launches": bands, "wall_s": wall_a, "receivers_per_s": na / wall_a,
               "kernel": "aircraft_kernel", "timing": "host wall of the launches with their transfers"}
        if not args.no_cpu:
            from oracle import oracle

            oracle.use_native()
            threads = args.cpu_threads or min(os.cpu_count() or 1, CPU_SHARE)
            pick = np.random.default_rng(29).choice(na, 4 * threads, replace=False)
            want_a = oracle.aircraft_points(ac, ac.points_xy[pick], threads=threads)
            cpu_a = pick.size / median_seconds(lambda: oracle.aircraft_points(ac, ac.points_xy[pick], threads=threads))
            one_a = 8 / median_seconds(lambda: oracle.aircraft_points(ac, ac.points_xy[pick[:8]], threads=1))
            heard = float(want_a["segments"].mean())
            lden_g, lden_w = lden_db(energy_a[pick]), lden_db(want_a["period_energy"])
            ok = np.isfinite(lden_w)
            # About 160 FP32 operations a heard segment (closest point, NPD read, duration, lateral
            # and installation terms, energy), transcendentals counted as one.
            flop = 160.0 * heard * na / wall_a / 1e12
            d29.update({
                "segments_heard_per_receiver": heard,
                "segment_evaluations_per_s": heard * na / wall_a,
                "roofline": {"bound": "valu", "achieved": flop, "peak": 157.3, "unit": "TFLOP/s",
                             "frac": flop / 157.3, "traffic": None, "kernel": "aircraft_kernel",
                             "flop_per_heard_segment": 160},
                "parity_vs_fp64_oracle": {"receivers": int(pick.size),
                                          "max_abs_lden_db": float(np.abs(lden_g[ok] - lden_w[ok]).max()),
                                          "silent_equal": bool(np.array_equal(np.isfinite(lden_g), ok)),
                                          "n65_n70_equal": bool(np.array_equal(n65_a[pick], want_a["n65"]) and
                                                                np.array_equal(n70_a[pick], want_a["n70"])),
                                          "hist_equal": bool(np.array_equal(hist_a[pick], want_a["hist"]))},
                "cpu_baseline": {"value": cpu_a, "unit": "receivers/s", "cores": threads, "kind": "port",
                                 "one_thread": one_a, **host_cpu(), "median_of": 3, "build": oracle.NATIVE_FLAGS,
                                 "sample": f"{pick.size} random receivers of the grid (8 on one thread), FP64 C++ "
                                           "restatement of the reference's CPU path"},
            })
        print(f"[bench] doc29 {na / wall_a:.3e} receivers/s over {bands} launches", file=sys.stderr, flush=True)

    # The loudness leg: ISO 532-1 on band-major planes of 28 third-octave levels per cell.
    loud = None
    if not args.no_loudness:
        cells_l = args.loudness_cells
        rng_l = np.random.default_rng(4)
        planes = rng_l.uniform(20.0, 90.0, (28, cells_l)).astype(np.float32)
        d_levels = hip.DeviceBuffer.from_array(planes)
        d_sone = hip.DeviceBuffer(cells_l * 4)
        gpu.loudness_device(d_levels.ptr, cells_l, 1, cells_l, d_sone.ptr, stream.handle)
        stream.synchronize()
        lt = []
        for k in range(5):
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            gpu.loudness_device(d_levels.ptr, cells_l, 1, cells_l, d_sone.ptr, stream.handle)
            e1.record(stream)
            e1.synchronize()
            lt.append(e1.seconds_since(e0))
        ls = sum(lt) / len(lt)
        sone = d_sone.to_array(np.float32, cells_l)
        loud = {"cells": cells_l, "kernel_ms": ls * 1e3, "cells_per_s": cells_l / ls,
                "roofline": {"bound": "hbm", "achieved": 116.0 * cells_l / ls / 1e9, "peak": 8000.0,
                             "unit": "GB/s", "frac": 116.0 * cells_l / ls / 1e9 / 8000.0, "traffic": None,
                             "kernel": "zwicker_kernel"}}
        if not args.no_cpu:
            from oracle import oracle

            oracle.use_native()
            threads = args.cpu_threads or min(os.cpu_count() or 1, CPU_SHARE)
            k = 200000
            cells64 = np.ascontiguousarray(planes[:, :k].T.astype(np.float64))
            want = oracle.zwicker_sone(cells64)
            cpu_l = k / median_seconds(lambda: threaded_chunks(lambda a, b: oracle.zwicker_sone(cells64[a:b]), k,
                                                               threads))
            k1 = 50000
            one_l = k1 / median_seconds(lambda: oracle.zwicker_sone(cells64[:k1]))
            loud["cpu_baseline"] = {"value": cpu_l, "unit": "cells/s", "cores": threads, "kind": "port",
                                    "one_thread": one_l, **host_cpu(), "median_of": 3, "build": oracle.NATIVE_FLAGS,
                                    "sample": f"{k} cells ({k1} on one thread), FP64 C++ restatement of the "
                                              "reference's CPU path"}
            loud["max_abs_sone_vs_oracle"] = float(np.abs(sone[:k].astype(np.float64) - want).max())
        print(f"[bench] loudness {loud['cells_per_s']:.3e} cells/s", file=sys.stderr, flush=True)

    # Algorithmic work per path from the oracle's counters on a sample of this workload's paths.
    weights = json.load(open(os.path.join(ROOT, "quietmap_amd", "roofline.json")))
    cpu = None
    parity = None
    work = None
    if not args.no_cpu:
        from oracle import oracle

        oracle.use_native()
        threads = args.cpu_threads or min(os.cpu_count() or 1, CPU_SHARE)
        rng = np.random.default_rng(1234)
        probe = rng.choice(n, 4 * threads, replace=False)
        t1 = time.time()
        oracle.paint_points(sc.scene, points[probe], sc.sources, sc.cutoff_m, threads=threads)
        per_receiver = (time.time() - t1) / probe.size
        # Three runs of about a sixth of --cpu-seconds each on `threads` threads, then as long on one.
        count = int(min(max(args.cpu_seconds / 6.0 / max(per_receiver, 1e-6), 16), 20000))
        sample = rng.choice(n, count, replace=False)
        ref = oracle.paint_points(sc.scene, points[sample], sc.sources, sc.cutoff_m, threads=threads)
        cpu_paths = int(ref["paths"].sum())
        cpu_s = median_seconds(lambda: oracle.paint_points(sc.scene, points[sample], sc.sources, sc.cutoff_m,
                                                           threads=threads))
        count1 = max(count // threads, 2)
        paths1 = int(ref["paths"][:count1].sum())
        one_s = median_seconds(lambda: oracle.paint_points(sc.scene, points[sample[:count1]], sc.sources,
                                                           sc.cutoff_m, threads=1))
        cpu = {"value": cpu_paths / cpu_s, "unit": "paths/s", "cores": threads, "kind": "port",
               "one_thread": paths1 / one_s, **host_cpu(), "median_of": 3, "build": oracle.NATIVE_FLAGS,
               "sample": f"{count} random receivers of the workload, all {cpu_paths} paths within 2 km "
                         f"({cpu_s:.1f} s a run; the first {count1} on one thread); FP64 C++ restatement of "
                         "the reference's Rust CPU path (Rust toolchain unavailable)"}
        got = lden_db(energy[sample])
        want = lden_db(ref["period_energy"])
        err = np.abs(got - want)
        err = err[np.isfinite(err)]
        parity = {"receivers": int(count), "max_abs_lden_db": float(err.max()),
                  "p999_abs_lden_db": float(np.quantile(err, 0.999)),
                  "paths_equal": bool(np.array_equal(path_counts[sample], ref["paths"].astype(np.int64)))}
        # Work counters over a pair sample of the same receivers.
        pr = rng.choice(sample, min(64, count), replace=False)
        pairs_point, pairs_source = [], []
        for j, i in enumerate(pr):
            d = np.hypot(sc.sources["x_m"] - points["x_m"][i], sc.sources["y_m"] - points["y_m"][i])
            near = np.flatnonzero(d <= sc.cutoff_m)
            pairs_source.append(near)
            pairs_point.append(np.full(near.size, j, np.uint32))
        res = oracle.evaluate_pairs(sc.scene, points[pr], sc.sources, np.concatenate(pairs_point),
                                    np.concatenate(pairs_source), threads=threads, point_bands=False)
        m = sum(len(x) for x in pairs_source)
        work = {k: v / m for k, v in res["work"].items()}
        print(f"[bench] cpu baseline {cpu['value']:.0f} paths/s on {threads} threads; work per path {work}",
              file=sys.stderr, flush=True)

    roofline = None
    if work is not None:
        f = weights["flop_per_path"]
        flop_path = (f["V"] * work["vertices"] + f["C"] * work["cells"] + f["E"] * work["edges"] +
                     f["X"] * work["crossings"] + f["K"] * work["candidates"] + f["const"])
        b = weights["bytes_per_path"]
        bytes_path = (b["V"] * work["vertices"] + b["C"] * work["cells"] + b["C_occ"] * work["cells_occupied"] +
                      b["E"] * work["edges"] + b["X"] * work["crossings"] + b["const"])
        achieved = flop_path * paths / kernel_s / 1e12
        peak = weights["peaks"]["fp32_valu_tflops"]
        traffic = None
        tfile = os.path.join(ROOT, "profiles", "paint_kernel_traffic.json")
        if os.path.exists(tfile):
            traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
        roofline = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                    "frac": achieved / peak, "traffic": traffic, "kernel": "paint_kernel",
                    "flop_per_path": flop_path, "compulsory_bytes_per_path": bytes_path,
                    "kernel_ms": kernel_s * 1e3}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kernel_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 crossings, path differences and energy sums)",
        "data": "synthetic (seeded scene generator; no reference data)",
        "config": {"workload": f"exact painter at tile scale: {gx}x{gy} receivers at 10 m, {sc.sources.size} "
                               f"point sources, {sc.scene.obstacles.outlines.size} polygon buildings, 30 m DEM "
                               f"200 m relief, {sc.cutoff_m:.0f} m cutoff; a step paints the central "
                               f"{nx}x{ny} receivers",
                   "paths_per_step": paths, "receivers_per_step": n, "failed_paths": failures,
                   "scale": args.scale},
        "lden_cells_per_s": cells,
        "wall_s_timed": wall,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_fp64_oracle": parity,
        "loudness": loud,
        "surrogate": sur,
        "flights": fl,
        "doc29": d29,
        "gpu": gpu.name,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
