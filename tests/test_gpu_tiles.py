"""The batched tiles painter (one persistent launch in a hipGraph) against the FP64 oracle per
receiver (each tile against its own sources only, Lden within 0.05 dB, paths counted exactly),
graph replays between eager work all equal to the first, and the autotuner's choice."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


@pytest.fixture(scope="module")
def mosaic():
    from quietmap_amd import scenes

    return scenes.tile_mosaic(seed=4, tiles=(3, 2), receivers=40, source_median=30, building_median=40)


def setup(gpu, m):
    from quietmap_amd import hip
    from quietmap_amd.painter import RECEIVER_BYTES

    square = gpu.upload(m.scene, m.sources)
    n = m.points.size
    d_points = hip.DeviceBuffer.from_array(m.points)
    d_rec = hip.DeviceBuffer(n * RECEIVER_BYTES)
    gpu.receivers_device(square, d_points.ptr, n, d_rec.ptr)
    outs = (hip.DeviceBuffer(n * 24), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4))
    return square, d_points, d_rec, outs


def test_tiles_match_the_oracle_and_replay(gpu, mosaic):
    from quietmap_amd import hip
    from quietmap_amd.painter import lden_db
    from quietmap_amd.tiles import TilePlan

    print("case: tiles 3x2 of 40x40", flush=True)
    m = mosaic
    square, _, d_rec, (d_e, d_p, d_f) = setup(gpu, m)
    n = m.points.size
    tiles = m.tile_table()
    plan = TilePlan(gpu, square, tiles, d_rec.ptr, n, 16, 2, d_e.ptr, d_p.ptr, d_f.ptr)
    nx, ny = m.tile_shape
    assert plan.paths == int((tiles["source_count"].astype(np.int64) * nx * ny).sum())
    stream = hip.Stream()
    runs = []
    for k in range(3):
        plan.run(stream.handle)
        stream.synchronize()
        runs.append((d_e.to_array(np.float64, (n, 3)), d_p.to_array(np.uint32, n), d_f.to_array(np.uint32, n)))
        gpu.loudness(np.full((64, 28), 60.0))  # eager work between the replays
    for e, p, f in runs[1:]:
        assert np.array_equal(p, runs[0][1]) and np.array_equal(f, runs[0][2])
        assert np.allclose(e, runs[0][0], rtol=1e-12, atol=0)
    energy, paths, fails = runs[0]
    # The oracle: every receiver of a tile against that tile's sources.
    pp, ps = [], []
    for t in tiles:
        r = np.arange(t["receiver_first"], t["receiver_first"] + nx * ny, dtype=np.uint32)
        s = np.arange(t["source_first"], t["source_first"] + t["source_count"], dtype=np.uint32)
        pp.append(np.repeat(r, s.size))
        ps.append(np.tile(s, r.size))
    pp, ps = np.concatenate(pp), np.concatenate(ps)
    ref = oracle.evaluate_pairs(m.scene, m.points, m.sources, pp, ps, point_bands=False)
    want = np.zeros((n, 3))
    np.add.at(want, pp, ref["sums"][:, :3])
    want_paths = np.bincount(pp, minlength=n)
    want_fails = np.bincount(pp, weights=(ref["failed"] != 0), minlength=n)
    assert np.array_equal(paths, want_paths)
    assert np.array_equal(fails, want_fails.astype(np.uint32))
    err = np.abs(lden_db(energy) - lden_db(want))
    assert not np.isinf(err).any(), "silent on one side only"
    err = err[~np.isnan(err)]
    print(f"tiles Lden: n={err.size} median={np.median(err):.5f} p99.9={np.quantile(err, 0.999):.5f} max={err.max():.5f} dB",
          flush=True)
    assert err.max() <= 0.05
    plan.close()
    square.close()


def test_autotune_picks_a_candidate(gpu, mosaic):
    from quietmap_amd import tiles as T

    print("case: tiles autotune", flush=True)
    square, _, d_rec, _ = setup(gpu, mosaic)
    res = T.autotune(gpu, square, mosaic.tile_table(), d_rec.ptr, mosaic.points.size,
                     candidates=((16, 1), (64, 2)), sample=4, repeats=1)
    print(res, flush=True)
    assert (res["block_sources"], res["groups_per_cu"]) in ((16, 1), (64, 2))
    assert all(t["paths_per_s"] > 0 for t in res["timings"])
    square.close()


def test_hundreds_of_tiles_at_the_configured_counts(gpu):
    """256 tiles with the configuration's log-normal counts (sources median 500, buildings median
    300, clipped 10..50k) in one autotuned captured plan, on tiles of the configuration's side
    (2.4 km) with 24 x 24 receivers at 100 m; 160 sampled receivers against the oracle (each
    against its own tile's sources), paths and failures alike."""
    from quietmap_amd import hip, scenes
    from quietmap_amd import tiles as T
    from quietmap_amd.painter import RECEIVER_BYTES, lden_db
    from quietmap_amd.tiles import TilePlan

    print("case: tiles 16x16 of 24x24 at median 500 sources", flush=True)
    m = scenes.tile_mosaic_fast(seed=9, tiles=(16, 16), receivers=24, spacing=100.0)
    n = m.points.size
    square = gpu.upload(m.scene, m.sources)
    d_points = hip.DeviceBuffer.from_array(m.points)
    d_rec = hip.DeviceBuffer(n * RECEIVER_BYTES)
    gpu.receivers_device(square, d_points.ptr, n, d_rec.ptr)
    tiles = m.tile_table()
    counts = tiles["source_count"]
    print(f"tiles {tiles.size}: sources median {np.median(counts):.0f}, min {counts.min()}, max {counts.max()}",
          flush=True)
    assert 350 <= np.median(counts) <= 700 and counts.max() > 5 * np.median(counts)
    tuned = T.autotune(gpu, square, tiles, d_rec.ptr, n, sample=6, repeats=1)
    d_e, d_p, d_f = hip.DeviceBuffer(n * 24), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4)
    plan = TilePlan(gpu, square, tiles, d_rec.ptr, n, tuned["block_sources"], tuned["groups_per_cu"], d_e.ptr, d_p.ptr,
                    d_f.ptr)
    plan.run()
    hip.synchronize()
    energy = d_e.to_array(np.float64, (n, 3))
    paths = d_p.to_array(np.uint32, n)
    fails = d_f.to_array(np.uint32, n)
    assert int(paths.astype(np.int64).sum()) == plan.paths
    rng = np.random.default_rng(10)
    pick = rng.choice(n, 160, replace=False)
    tile_of = pick // (24 * 24)
    pp, ps = [], []
    for j, (i, t) in enumerate(zip(pick, tile_of)):
        s = np.arange(tiles["source_first"][t], tiles["source_first"][t] + tiles["source_count"][t], dtype=np.uint32)
        pp.append(np.full(s.size, i, np.uint32))
        ps.append(s)
    pp, ps = np.concatenate(pp), np.concatenate(ps)
    ref = oracle.evaluate_pairs(m.scene, m.points, m.sources, pp, ps, point_bands=False)
    want = np.zeros((n, 3))
    np.add.at(want, pp, ref["sums"][:, :3])
    assert np.array_equal(paths[pick].astype(np.int64), np.bincount(pp, minlength=n)[pick])
    # The card counts a path it cannot hold (more crossings, hull points or roofs than its
    # buffers: the densest tiles) as failed, where the oracle has no bound; every receiver the
    # card painted whole agrees with the oracle, which fails none of its paths.
    want_fails = np.bincount(pp, weights=(ref["failed"] != 0), minlength=n)[pick]
    whole = fails[pick] == 0
    assert np.all(want_fails[whole] == 0)
    share = fails.astype(np.int64).sum() / plan.paths
    err = np.abs(lden_db(energy[pick][whole]) - lden_db(want[pick][whole]))
    assert not np.isinf(err).any()
    print(f"tiles at the configured counts: {plan.paths} paths, {share:.2e} of them over the buffers; Lden over "
          f"{int(whole.sum())} whole receivers max {err.max():.5f} dB (block {tuned['block_sources']}, groups "
          f"{tuned['groups_per_cu']})", flush=True)
    assert share < 2e-2 and whole.mean() > 0.8
    assert err.max() <= 0.05
    plan.close()
    square.close()
