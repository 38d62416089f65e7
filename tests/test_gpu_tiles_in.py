"""A square read from the reference's tile formats (four z12 qmterr2 and qmobs2 tiles about
Prague and the qmwthr2 world table, quietmap_amd/tiles_in.py) painted on the card against the
FP64 oracle on the same square: point pairs within 0.05 dB with failures alike, crossing lists
equal, and the grid's Lden with its paths."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi, scenes
from tests.tile_formats import four_tile_square

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def test_square_from_tiles(gpu):
    from quietmap_amd.painter import lden_db

    print("case: square read from tiles", flush=True)
    scene, (ox, oy, east_m, north_m), world, (tx, ty) = four_tile_square(buildings=120)
    # Receivers and sources about the tiles' shared corner, in the frame's metres.
    cx, cy = (tx + 1 - ox) * east_m, (oy - (ty + 1)) * north_m
    rng = np.random.default_rng(4)
    side, spacing = 48, 15.0
    points = scenes.receiver_grid(side, side, spacing, (cx - side * spacing / 2, cy - side * spacing / 2))
    src = np.zeros(200, abi.SOURCE_DTYPE)
    src["x_m"] = cx + rng.uniform(-900, 900, src.size)
    src["y_m"] = cy + rng.uniform(-900, 900, src.size)
    src["height_m"] = 0.05
    src["exclusion_radius_m"] = 2.0
    src["ground_percent"] = abi.GROUND_FROM_TERRAIN
    src["energy"] = 10.0 ** ((rng.normal(80, 6, (src.size, 1, 8)) + scenes.A_WEIGHTING_DB) / 10.0)
    square = gpu.upload(scene, src)
    n = 3000
    pp = rng.integers(0, points.size, n).astype(np.uint32)
    ps = rng.integers(0, src.size, n).astype(np.uint32)
    got = gpu.evaluate_pairs(square, points, pp, ps)
    want = oracle.evaluate_pairs(scene, points, src, pp, ps, point_bands=False)
    assert np.array_equal(got["failed"] != 0, want["failed"] != 0)
    ok = got["failed"] == 0
    assert ok.mean() > 0.99
    g, w = got["sums"][ok].astype(np.float64), want["sums"][ok]
    assert np.array_equal(g > 0, w > 0)
    err = np.abs(10 * np.log10(g[w > 0] / w[w > 0]))
    print(f"pairs from tiles: n={err.size} max {err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    # Crossing lists over the buildings read from several tiles each.
    seg = np.stack([points["x_m"][pp[:400]], points["y_m"][pp[:400]], src["x_m"][ps[:400]], src["y_m"][ps[:400]]], 1)
    gc = gpu.crossings(square, seg)
    crossed = 0
    for i in range(400):
        _, ref = oracle.crossings(scene.obstacles, seg[i, :2], seg[i, 2:])
        k = int(gc["count"][i])
        assert k == ref.size
        assert np.array_equal(gc["t"][i, :k], ref["t"]) and np.array_equal(gc["footprint"][i, :k], ref["footprint_id"])
        crossed += k
    print(f"crossings from tiles: {crossed} over 400 segments, lists equal", flush=True)
    assert crossed > 50
    r = gpu.paint_grid(square, points, side, side, 800.0)
    pick = rng.choice(points.size, 128, replace=False)
    ref = oracle.paint_points(scene, points[pick], src, 800.0)
    assert np.array_equal(r["paths"][pick].astype(np.int64), ref["paths"].astype(np.int64))
    got_l, want_l = lden_db(r["period_energy"][pick]), lden_db(ref["period_energy"])
    assert np.array_equal(np.isfinite(got_l), np.isfinite(want_l))
    e = np.abs(got_l - want_l)
    e = e[np.isfinite(e)]
    print(f"grid from tiles: Lden max {e.max():.5f} dB", flush=True)
    assert e.max() <= 0.05
    square.close()
