"""One launch of each kernel on a fixed workload, for rocprofv3 (kernel traces and counters).

    python -m quietmap_amd.profile_workload --kernels paint,loudness,aircraft,copy

`paint` paints the central `--window` x `--window` receivers of the tile-scale scene at `--scale`
in one paint_kernel launch (the defaults are the benchmark's step), `surrogate` paints the same
window with the trained network (quietmap_amd/surrogate_params.npy), `loudness` runs
zwicker_kernel over `--cells` cells of band-major planes, `aircraft` paints the aircraft scene at
`--aircraft-scale` (`passes` with its 48 pass classes per period), `pixels` runs the pixel
loudness kernel (pixel_loudness_kernel) over `--pixels` synthetic packed pixels,
`flights` sums `--flight-boxes` synthetic aircraft boxes over 20 km at `--flight-points` points
(every point hearing all 16 tiles: flight_sums_kernel), `horizons` builds the horizons of
`--horizon-points` receivers of the tile-scale scene at `--scale` 0.25 (horizons_kernel), and
`mosaic` paints the batched uneven tiles configuration at full size (`--mosaic-tiles` a side of
256 x 256 receivers, log-normal sources and buildings) through autotuned captured plans of
`--mosaic-batch` tiles (tiles_kernel), and
`phases` paints the `paint` window through the phase-clock build (libqmpaint_phases.so, alone in
its process) and prints how the waves' time splits over the phases of a path, and
`pairs` evaluates 2,000,000 point pairs of the tile-scale scene through Gpu::evaluate_pairs
(pair_order_kernel, the radix sort and pairs_kernel): the sources nearest the centre of its
256 x 256 central receivers, all of each source's receivers within 2 km in the caller's row-major
order (a candidate's pairs as the Batch hands them over), then the same pairs shuffled, and
`copy` makes one device-to-device copy of `--copy-mib` MiB, whose
FETCH_SIZE and WRITE_SIZE calibrate the counters against a known byte count. Each step prints a
line when it ends.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import hip, scenes
from .painter import RECEIVER_BYTES, Gpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernels", default="paint,loudness,aircraft,copy")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--window", type=int, default=1024)
    ap.add_argument("--cells", type=int, default=16 * 1024 * 1024)
    ap.add_argument("--aircraft-scale", type=float, default=0.05)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--pixels", type=int, default=4096)
    ap.add_argument("--flight-boxes", type=int, default=20000)
    ap.add_argument("--flight-points", type=int, default=4096)
    ap.add_argument("--horizon-points", type=int, default=4096)
    ap.add_argument("--mosaic-tiles", type=int, default=64, help="tiles a side of the batched mosaic")
    ap.add_argument("--mosaic-batch", type=int, default=256, help="tiles a captured plan paints")
    args = ap.parse_args()
    kernels = args.kernels.split(",")
    if "phases" in kernels:
        if kernels != ["phases"]:
            raise SystemExit("the phases leg runs alone (it loads the phase-clock build)")
        from . import painter

        painter.use_phase_clocks()
    gpu = Gpu(0)
    if "phases" in kernels:
        from . import painter

        sc = scenes.tile_scene(scale=args.scale)
        points, nx, ny = sc.window(args.window)
        n = nx * ny
        square = gpu.upload(sc.scene, sc.sources)
        d_points = hip.DeviceBuffer.from_array(points)
        d_rec = hip.DeviceBuffer(n * RECEIVER_BYTES)
        d_e, d_p, d_f = hip.DeviceBuffer(n * 24), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4)
        gpu.receivers_device(square, d_points.ptr, n, d_rec.ptr)
        hip.synchronize()
        painter.phase_clocks(reset=True)
        t = time.time()
        gpu.paint_receivers_device(square, d_rec.ptr, nx, ny, sc.cutoff_m, d_e.ptr, d_p.ptr, d_f.ptr)
        hip.synchronize()
        seconds = time.time() - t
        clocks = painter.phase_clocks(reset=True)
        paths = int(d_p.to_array(np.uint32, n).astype(np.int64).sum())
        total = sum(clocks.values())
        print(f"phases: {nx}x{ny} receivers, {paths} paths, {seconds:.2f} s with the clocks", flush=True)
        for name, cycles in clocks.items():
            print(f"phases: {name:10s} {100.0 * cycles / total:6.2f} %  {cycles * 64 / max(paths, 1):9.0f} "
                  f"wave cycles a round of 64 paths", flush=True)
    if "paint" in kernels:
        sc = scenes.tile_scene(scale=args.scale)
        points, nx, ny = sc.window(args.window)
        n = nx * ny
        square = gpu.upload(sc.scene, sc.sources)
        d_points = hip.DeviceBuffer.from_array(points)
        d_rec = hip.DeviceBuffer(n * RECEIVER_BYTES)
        d_e, d_p, d_f = hip.DeviceBuffer(n * 24), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4)
        gpu.receivers_device(square, d_points.ptr, n, d_rec.ptr)
        t = time.time()
        gpu.paint_receivers_device(square, d_rec.ptr, nx, ny, sc.cutoff_m, d_e.ptr, d_p.ptr, d_f.ptr)
        hip.synchronize()
        paths = int(d_p.to_array(np.uint32, n).astype(np.int64).sum())
        print(f"paint: {nx}x{ny} receivers, {paths} paths, {time.time() - t:.2f} s", flush=True)
    if "surrogate" in kernels:
        import os

        from . import surrogate

        sc = scenes.tile_scene(scale=args.scale)
        points, nx, ny = sc.window(args.window)
        n = nx * ny
        square = gpu.upload(sc.scene, sc.sources)
        d_points = hip.DeviceBuffer.from_array(points)
        d_rec = hip.DeviceBuffer(n * RECEIVER_BYTES)
        d_e, d_p, d_f, d_s = (hip.DeviceBuffer(n * 24), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4),
                              hip.DeviceBuffer(n * 4))
        gpu.receivers_device(square, d_points.ptr, n, d_rec.ptr)
        params = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_params.npy"))
        net = surrogate.Mlp(gpu, params, max_batch=1)
        t = time.time()
        surrogate.paint_receivers_device(gpu, square, net, d_rec.ptr, nx, ny, sc.cutoff_m, d_e.ptr, d_p.ptr, d_f.ptr,
                                         d_s.ptr)
        hip.synchronize()
        paths = int(d_p.to_array(np.uint32, n).astype(np.int64).sum())
        print(f"surrogate: {nx}x{ny} receivers, {paths} paths, {time.time() - t:.2f} s", flush=True)
    if "loudness" in kernels:
        planes = np.random.default_rng(4).uniform(20.0, 90.0, (28, args.cells)).astype(np.float32)
        d_levels = hip.DeviceBuffer.from_array(planes)
        d_sone = hip.DeviceBuffer(args.cells * 4)
        gpu.loudness_device(d_levels.ptr, args.cells, 1, args.cells, d_sone.ptr)
        hip.synchronize()
        print(f"loudness: {args.cells} cells, {args.cells * 116} algorithmic bytes", flush=True)
    if "aircraft" in kernels:
        ac = scenes.aircraft_scene(scale=args.aircraft_scale)
        t = time.time()
        r = gpu.aircraft_grid(ac)
        print(f"aircraft: {ac.grid_shape} receivers, {ac.tracks.size} tracks, events binned "
              f"{int(r['hist'].sum())}, {time.time() - t:.2f} s", flush=True)
    if "passes" in kernels:
        ac = scenes.aircraft_scene(scale=args.aircraft_scale)
        t = time.time()
        r = gpu.aircraft_grid(ac, passes=True)
        print(f"passes: {ac.grid_shape} receivers, {ac.tracks.size} tracks, classes held "
              f"{int((r['passes'][:, 1] > 0).sum())}, {time.time() - t:.2f} s", flush=True)
    if "pixels" in kernels:
        from . import pixels

        packed, offsets = pixels.concat(pixels.synthetic_pixels(seed=21, count=args.pixels))
        t = time.time()
        nd = gpu.pixel_loudness(packed, offsets)
        print(f"pixels: {args.pixels} pixels, place Nden mean {float(nd[:, 6].mean()):.2f} sone, "
              f"{time.time() - t:.2f} s", flush=True)
    if "pairs" in kernels:
        sc = scenes.tile_scene(scale=args.scale)
        points, nx, ny = sc.window(256)
        square = gpu.upload(sc.scene, sc.sources)
        px, py = points["x_m"].astype(np.float64), points["y_m"].astype(np.float64)
        sx, sy = sc.sources["x_m"], sc.sources["y_m"]
        order = np.argsort((sx - px.mean()) ** 2 + (sy - py.mean()) ** 2, kind="stable")
        pp, ps, total = [], [], 0
        for s in order:
            inside = np.flatnonzero((px - sx[s]) ** 2 + (py - sy[s]) ** 2 <= sc.cutoff_m ** 2)
            pp.append(inside.astype(np.uint32))
            ps.append(np.full(inside.size, s, np.uint32))
            total += inside.size
            if total >= 2_000_000:
                break
        pp, ps = np.concatenate(pp)[:2_000_000], np.concatenate(ps)[:2_000_000]
        shuffle = np.random.default_rng(5).permutation(pp.size)
        for name, a, b in (("candidate order", pp, ps), ("shuffled", pp[shuffle], ps[shuffle])):
            t = time.time()
            r = gpu.evaluate_pairs(square, points, a, b, point_bands=False)
            print(f"pairs ({name}): {a.size} pairs, {int((r['failed'] != 0).sum())} failed, "
                  f"{time.time() - t:.2f} s with transfers", flush=True)
    if "flights" in kernels:
        from . import flights

        boxes = flights.synthetic_boxes(seed=31, count=args.flight_boxes, extent_m=20000.0)
        side = int(round(args.flight_points ** 0.5))
        jj, ii = np.mgrid[0:side, 0:side]
        xy = np.stack([(ii.reshape(-1) + 0.5) * 20000.0 / side, (jj.reshape(-1) + 0.5) * 20000.0 / side], 1)
        boxes, fpoints, segments, n_sums = flights.flight_requests(boxes, xy, extent_m=20000.0)
        handle = gpu.flight_boxes(boxes)
        t = time.time()
        sums = gpu.flight_sums(handle, fpoints, None, segments, n_sums)
        print(f"flights: {boxes.size} boxes, {fpoints.size} points, {segments.size} segments, "
              f"{int(segments['count'].astype(np.int64).sum())} box reads, heard energy {float(sums[:, 0].mean()):.3e}, "
              f"{time.time() - t:.2f} s", flush=True)
        handle.close()
    if "horizons" in kernels:
        sc = scenes.tile_scene(scale=0.25)
        square = gpu.upload(sc.scene, sc.sources)
        pts = sc.points[np.random.default_rng(5).choice(sc.points.size, args.horizon_points, replace=False)]
        t = time.time()
        rec, failed = gpu.horizons(square, pts)
        print(f"horizons: {pts.size} receivers, {int((failed != 0).sum())} failed, {time.time() - t:.2f} s", flush=True)
    if "mosaic" in kernels:
        # The batched uneven tiles at full size: tiles of 256 x 256 receivers with log-normal
        # source (median 500) and building (median 300) counts, each against its own sources;
        # the autotuner picks (block_sources, groups_per_cu) on 8 sampled tiles, then captured
        # plans of --mosaic-batch tiles (each one persistent launch in a hipGraph) paint them all.
        from . import tiles as tiles_mod

        t = time.time()
        m = scenes.tile_mosaic_fast(tiles=(args.mosaic_tiles, args.mosaic_tiles))
        n = m.points.size
        table = m.tile_table()
        print(f"mosaic: {m.n_tiles} tiles, {n} receivers, {m.sources.size} sources, "
              f"{m.scene.obstacles.outlines.size} buildings, generated in {time.time() - t:.1f} s", flush=True)
        square = gpu.upload(m.scene, m.sources)
        d_points = hip.DeviceBuffer.from_array(m.points)
        d_rec = hip.DeviceBuffer(n * RECEIVER_BYTES)
        gpu.receivers_device(square, d_points.ptr, n, d_rec.ptr)
        hip.synchronize()
        del d_points
        tuned = tiles_mod.autotune(gpu, square, table, d_rec.ptr, n)
        print(f"mosaic: autotuned on {tuned['sample_tiles']} tiles: block_sources {tuned['block_sources']}, "
              f"groups_per_cu {tuned['groups_per_cu']}", flush=True)
        d_e, d_p, d_f = hip.DeviceBuffer(n * 24), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4)
        stream = hip.Stream()
        total_paths, total_s, counted, failed = 0, 0.0, 0, 0
        for at in range(0, m.n_tiles, args.mosaic_batch):
            plan = tiles_mod.TilePlan(gpu, square, table[at:at + args.mosaic_batch], d_rec.ptr, n,
                                      tuned["block_sources"], tuned["groups_per_cu"], d_e.ptr, d_p.ptr, d_f.ptr)
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            plan.run(stream.handle)
            e1.record(stream)
            e1.synchronize()
            s_ = e1.seconds_since(e0)
            total_paths += plan.paths
            total_s += s_
            # A run clears every receiver's outputs, then paints its own tiles'.
            counted += int(d_p.to_array(np.uint32, n).astype(np.int64).sum())
            failed += int(d_f.to_array(np.uint32, n).astype(np.int64).sum())
            print(f"mosaic: tiles {at}-{min(at + args.mosaic_batch, m.n_tiles) - 1}: {plan.paths} paths in "
                  f"{s_:.2f} s ({plan.paths / s_:.3e} paths/s)", flush=True)
            plan.close()
        print(f"mosaic: {total_paths} paths ({counted} counted on the card, {failed} over the per-path buffers) in "
              f"{total_s:.1f} s: {total_paths / total_s:.3e} paths/s", flush=True)
    if "copy" in kernels:
        nbytes = args.copy_mib << 20
        a, b = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
        hip.check(hip.runtime().hipMemcpy(b.ptr, a.ptr, nbytes, hip.D2D))
        hip.synchronize()
        print(f"copy: {nbytes} bytes device to device", flush=True)


if __name__ == "__main__":
    main()
