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
