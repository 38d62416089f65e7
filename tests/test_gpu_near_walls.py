"""Receivers centimetres from a wall (a lattice receiver beside a facade, or inside a footprint its
own-building test missed) on the card against the FP64 oracle: the receiver-side mean plane then
spans the few centimetres from the wall to the receiver, fitted from the difference of two prefix
moments over a long terrain segment, and the pair energies stay within 0.05 dB of the oracle's
(a point on the segment rounded to FP32 in one of them once moved such pairs by 4-6 dB)."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi

from quietmap_amd import scenes

BASE_M = 250.0  # the terrain's altitude: an FP32-rounded point there is off by 1e-5 m


def wall_scene():
    """Terrain on a 100 m lattice rising 1 m a node eastward from 250 m, and a 12 m square
    building from x = 40 to 52 m, 15 m high."""
    n = 61
    xs = -3000.0 + 100.0 * np.arange(n)
    heights = np.tile(BASE_M + 0.01 * xs, (n, 1))
    terrain = abi.Terrain(-3000.0, -3000.0, 100.0, abi.Terrain.pack(heights, np.full((n, n), 50)))
    ring = [(40.0, -6.0), (52.0, -6.0), (52.0, 6.0), (40.0, 6.0)]
    obstacles = scenes._obstacles([ring], [15.0], [abi.KIND_BUILDING], np.array([7], np.uint64), (-600.0, -600.0),
                                  (1200.0, 1200.0))
    weather = scenes._weather((-1000.0, -1000.0), (2000.0, 2000.0), lambda r, c: np.full((3, 16), 50))
    return abi.Scene(terrain, obstacles, weather)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def test_receivers_beside_a_wall(gpu):
    print("case: receivers beside a wall", flush=True)
    scene = wall_scene()
    rng = np.random.default_rng(71)
    gaps = np.array([0.01, 0.02, 0.03, 0.05, 0.1, 0.5])
    points = np.zeros(gaps.size * 2, abi.POINT_DTYPE)
    points["x_m"][: gaps.size] = 40.0 - gaps  # outside, west of the near wall
    points["x_m"][gaps.size :] = 40.0 + gaps  # inside the box, its own footprint unset
    points["y_m"] = rng.uniform(-4.0, 4.0, points.size)
    sources = np.zeros(200, abi.SOURCE_DTYPE)
    # East of the box, rays to the receivers cross it whole (a roof); west of it, rays to the
    # receivers inside cross one wall centimetres before them (a top without a roof).
    sources["x_m"] = rng.uniform(120.0, 900.0, sources.size) * rng.choice([-1.0, 1.0], sources.size)
    sources["y_m"] = rng.uniform(-60.0, 60.0, sources.size)
    sources["height_m"] = rng.choice([0.05, 0.5, 4.0], sources.size)
    sources["ground_percent"] = abi.GROUND_FROM_TERRAIN
    sources["exclusion_radius_m"] = 2.0
    sources["energy"] = 1e8
    square = gpu.upload(scene, sources)
    pp = np.repeat(np.arange(points.size, dtype=np.uint32), sources.size)
    ps = np.tile(np.arange(sources.size, dtype=np.uint32), points.size)
    got = gpu.evaluate_pairs(square, points, pp, ps)
    want = oracle.evaluate_pairs(scene, points, sources, pp, ps)
    ok = (got["failed"] == 0) & (want["failed"] == 0)
    assert ok.mean() > 0.99
    with np.errstate(divide="ignore", invalid="ignore"):
        err = np.abs(10 * np.log10(got["sums"][ok].astype(np.float64) / want["sums"][ok]))
    assert not np.isinf(err).any(), "silent on one side only"
    err = err[~np.isnan(err)]
    print(f"pairs beside walls: n={err.size // 9} max={err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    square.close()
