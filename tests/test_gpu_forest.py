"""Foliage on the card against the FP64 oracle: the road scene with woods (forest cover 40-100 %
in round stands, crowns 8-25 m high, ISO 9613-2 Table A.1 over the metres each state's ray spends
inside crowns; physics/src/foliage.rs:22-64 through popup/src/evaluate.rs and the card's
add_foliage), point pairs within 0.05 dB, failures alike, and the woods moving many pairs."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi, scenes

pytestmark = pytest.mark.gpu
TOLERANCE_DB = 0.05


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def wooded(sc, seed=19, stands=40):
    """The scene's terrain with round stands of trees: cover and crown height per node."""
    t = sc.scene.terrain
    rows, cols = t.nodes.shape
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:rows, 0:cols]
    x = t.origin_x_m + xx * t.spacing_m
    y = t.origin_y_m + yy * t.spacing_m
    cover = np.zeros((rows, cols))
    canopy = np.zeros((rows, cols), np.uint8)
    extent = cols * t.spacing_m
    for _ in range(stands):
        cx, cy = rng.uniform(0, extent, 2)
        r = rng.uniform(60.0, 260.0)
        inside = (x - cx) ** 2 + (y - cy) ** 2 <= r * r
        cover[inside] = np.maximum(cover[inside], rng.uniform(40, 100))
        canopy[inside] = np.maximum(canopy[inside], np.uint8(rng.uniform(8, 25)))
    height = -500.0 + (t.nodes & 0xFFFF) / 5.0
    impervious = (t.nodes >> 16) & 0xFF
    nodes = abi.Terrain.pack(height, impervious, np.round(cover).astype(np.uint32))
    terrain = abi.Terrain(t.origin_x_m, t.origin_y_m, t.spacing_m, nodes, canopy_m=canopy)
    return abi.Scene(terrain, sc.scene.obstacles, sc.scene.weather), float((cover > 0).mean())


def db(x):
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.asarray(x, np.float64))


def test_forest_pairs_within_0_05_db(gpu):
    print("case: forest pairs", flush=True)
    sc = scenes.road_scene(receivers=96, lines=60, buildings=120)
    scene, share = wooded(sc)
    print(f"wooded share of the lattice {share:.2f}", flush=True)
    rng = np.random.default_rng(23)
    n = 3000
    pp = rng.integers(0, sc.points.size, n).astype(np.uint32)
    ps = rng.integers(0, sc.sources.size, n).astype(np.uint32)
    square = gpu.upload(scene, sc.sources)
    got = gpu.evaluate_pairs(square, sc.points, pp, ps)
    want = oracle.evaluate_pairs(scene, sc.points, sc.sources, pp, ps)
    bare = oracle.evaluate_pairs(sc.scene, sc.points, sc.sources, pp, ps, point_bands=False)
    assert np.array_equal(got["failed"] != 0, want["failed"] != 0)
    ok = (got["failed"] == 0) & (want["failed"] == 0)
    assert ok.mean() > 0.99
    g, w = got["sums"][ok].astype(np.float64), want["sums"][ok]
    assert np.array_equal(g > 0, w > 0), "silent on one side only"
    err = np.abs(db(g) - db(w))
    err = err[~np.isnan(err)]
    moved = np.abs(db(want["sums"][ok, :3]) - db(bare["sums"][ok, :3]))
    moved = moved[~np.isnan(moved)]
    print(f"forest pairs: n={err.size} median={np.median(err):.5f} p99.9={np.quantile(err, 0.999):.5f} "
          f"max={err.max():.5f} dB; woods move {np.mean(moved > 0.5):.2f} of pairs by > 0.5 dB "
          f"(max {moved.max():.1f} dB)", flush=True)
    assert err.max() <= TOLERANCE_DB
    assert np.mean(moved > 0.5) > 0.1, "the woods moved too few pairs to test foliage"
    square.close()


def test_forest_grid_within_0_05_db(gpu):
    """The grid painter over the wooded scene (the painter's instantiation with the foliage term;
    a square without forest cover takes the one without) against the oracle's painted points."""
    print("case: forest grid", flush=True)
    sc = scenes.road_scene(receivers=96, lines=60, buildings=120)
    scene, _ = wooded(sc)
    nx, ny = sc.grid_shape
    square = gpu.upload(scene, sc.sources)
    got = gpu.paint_grid(square, sc.points, nx, ny, sc.cutoff_m)
    bare_square = gpu.upload(sc.scene, sc.sources)
    bare = gpu.paint_grid(bare_square, sc.points, nx, ny, sc.cutoff_m)
    sample = np.random.default_rng(29).choice(sc.points.size, 200, replace=False)
    want = oracle.paint_points(scene, sc.points[sample], sc.sources, sc.cutoff_m)
    from quietmap_amd.painter import lden_db

    assert np.array_equal(got["paths"][sample], want["paths"].astype(np.uint32))
    err = np.abs(lden_db(got["period_energy"][sample]) - lden_db(want["period_energy"]))
    moved = np.abs(lden_db(got["period_energy"]) - lden_db(bare["period_energy"]))
    print(f"forest grid: n={err.size} max={np.nanmax(err):.5f} dB; woods move {np.mean(moved > 0.1):.2f} of the "
          f"receivers by > 0.1 dB", flush=True)
    assert np.nanmax(err) <= TOLERANCE_DB
    assert np.mean(moved > 0.1) > 0.1
    square.close()
    bare_square.close()
