"""The HIP painter against the FP64 oracle on the same seeded scenes, through the C ABI.

Bars: crossing lists bit-exact (t, footprint, outline, edge, flags); pair energies and painted
Lden within 0.05 dB of the oracle (the reference's own GPU-vs-CPU gate,
paint-gpu/tests/corner_rays.rs:150,298); paths counted exactly; no failed pairs."""
import sys

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
    print("gpu:", g.name, flush=True)
    yield g
    g.close()


@pytest.fixture(scope="module")
def road():
    return scenes.road_scene()


@pytest.fixture(scope="module")
def tile():
    return scenes.tile_scene(scale=0.06)


def db(x):
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.asarray(x, np.float64))


def report(name, err):
    err = np.asarray(err)
    assert not np.isinf(err).any(), f"{name}: silent on one side only"
    err = np.abs(err[~np.isnan(err)])
    print(f"{name}: n={err.size} median={np.median(err):.5f} p99.9={np.quantile(err, 0.999):.5f} "
          f"max={err.max():.5f} dB", flush=True)
    return err.max()


@pytest.mark.parametrize("which", ["road", "tile"])
def test_crossing_lists_bit_exact(gpu, road, tile, which):
    sc = road if which == "road" else tile
    print("case: crossings", which, flush=True)
    square = gpu.upload(sc.scene, sc.sources)
    rng = np.random.default_rng(11)
    ext = sc.grid_shape[0] * 10.0
    n = 3000
    seg = np.concatenate([rng.uniform(0, ext, (n, 2)), rng.uniform(0, ext, (n, 2))], axis=1)
    got = gpu.crossings(square, seg)
    total = 0
    for i in range(n):
        failed, want = oracle.crossings(sc.scene.obstacles, seg[i, :2], seg[i, 2:])
        assert failed == 0 and got["failed"][i] == 0
        k = int(got["count"][i])
        assert k == len(want), i
        total += k
        assert got["t"][i, :k].tobytes() == want["t"].tobytes(), i
        assert np.array_equal(got["footprint"][i, :k], want["footprint_id"])
        assert np.array_equal(got["outline"][i, :k], want["outline"])
        assert np.array_equal(got["edge"][i, :k], want["edge"])
        flags = want["building"].astype(np.uint32) | (want["bounds_roof"].astype(np.uint32) << 1)
        assert np.array_equal(got["flags"][i, :k], flags)
    assert total > 1000


@pytest.mark.parametrize("which", ["road", "tile"])
def test_pair_energies_within_0_05_db(gpu, road, tile, which):
    sc = road if which == "road" else tile
    print("case: pairs", which, flush=True)
    square = gpu.upload(sc.scene, sc.sources)
    rng = np.random.default_rng(3)
    n = 4000
    pp = rng.integers(0, sc.points.size, n).astype(np.uint32)
    ps = rng.integers(0, sc.sources.size, n).astype(np.uint32)
    if which == "tile":  # within the cutoff, as the painter evaluates them
        d = np.hypot(sc.sources["x_m"][ps] - sc.points["x_m"][pp], sc.sources["y_m"][ps] - sc.points["y_m"][pp])
        keep = d <= sc.cutoff_m
        pp, ps = pp[keep], ps[keep]
    got = gpu.evaluate_pairs(square, sc.points, pp, ps)
    want = oracle.evaluate_pairs(sc.scene, sc.points, sc.sources, pp, ps)
    assert not want["failed"].any()
    assert not got["failed"].any(), np.flatnonzero(got["failed"])[:10]
    worst = report(f"{which} mixed", db(got["sums"][:, :3]) - db(want["sums"][:, :3]))
    worst_states = report(f"{which} states", db(got["sums"][:, 3:]) - db(want["sums"][:, 3:]))
    lden_err = report(f"{which} pair Lden", db([oracle.lden_energy(s) for s in got["sums"][:, :3].astype(np.float64)])
                      - db([oracle.lden_energy(s) for s in want["sums"][:, :3]]))
    assert worst <= TOLERANCE_DB and worst_states <= TOLERANCE_DB and lden_err <= TOLERANCE_DB
    # Point bands: the per-point, per-layer spectra.
    gb, wb = got["point_bands"].astype(np.float64), want["point_bands"]
    mask = wb > 0
    assert np.array_equal(gb > 0, mask)
    assert report(f"{which} point bands", db(gb[mask]) - db(wb[mask])) <= TOLERANCE_DB


def test_painted_grid_lden_within_0_05_db(gpu, tile):
    print("case: paint grid tile", flush=True)
    sc = tile
    square = gpu.upload(sc.scene, sc.sources)
    nx, ny = sc.grid_shape
    got = gpu.paint_grid(square, sc.points, nx, ny, sc.cutoff_m)
    rng = np.random.default_rng(8)
    sample = rng.choice(sc.points.size, 64, replace=False)
    want = oracle.paint_points(sc.scene, sc.points[sample], sc.sources, sc.cutoff_m)
    assert np.array_equal(got["paths"][sample], want["paths"].astype(np.uint32))
    assert not got["failures"].any() and not want["failures"].any()
    from quietmap_amd.painter import lden_db

    err = lden_db(got["period_energy"][sample]) - lden_db(want["period_energy"])
    assert report("grid Lden", err) <= TOLERANCE_DB
    assert got["paths"].sum() > 100_000


def test_no_pair_fails_on_the_road_scene_grid(gpu, road):
    print("case: paint grid road", flush=True)
    square = gpu.upload(road.scene, road.sources)
    nx, ny = road.grid_shape
    # A 64x64 corner of the road grid, every source.
    sel = (np.arange(64)[:, None] * nx + np.arange(64)[None, :]).reshape(-1)
    got = gpu.paint_grid(square, road.points[sel], 64, 64, 1e9)
    assert np.all(got["paths"] == road.sources.size)
    assert not got["failures"].any()
    sample = np.random.default_rng(2).choice(sel.size, 24, replace=False)
    want = oracle.paint_points(road.scene, road.points[sel][sample], road.sources, 1e9)
    from quietmap_amd.painter import lden_db

    assert report("road grid Lden", lden_db(got["period_energy"][sample]) - lden_db(want["period_energy"])) <= TOLERANCE_DB
