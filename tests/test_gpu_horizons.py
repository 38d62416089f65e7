"""The card's horizon builder (qm_horizons: Horizons::build and write_f32, popup/src/aircraft/
horizons.rs:138-201) against the FP64 oracle, whose builder is pinned on a plane and a box by
geometry: every float of the record the same NaN / infinite pattern, ranges and tangents within
1e-6 relative (both FP64 until written as FP32), on the plane scene, on a tile of hilly
terrain with polygon buildings (receivers inside buildings drop their own walls), with and without
fine terrain, at given altitudes and at the receivers' own (ground + 4 m). The records then screen
the aircraft grid as the oracle's own records do."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi, scenes

from tests.test_oracle_horizons import plane_scene

pytestmark = pytest.mark.gpu

SECTOR_FLOATS = 1 + 2 * abi.HORIZON_BANDS


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def no_sources():
    s = np.zeros(1, abi.SOURCE_DTYPE)
    s["x_m"], s["y_m"], s["height_m"], s["ground_percent"] = 0.0, 0.0, 0.05, abi.GROUND_FROM_TERRAIN
    return s


def oracle_records(scene, points, altitudes):
    out = np.zeros((points.size, abi.HORIZON_FLOATS), np.float32)
    failed = np.zeros(points.size, np.uint32)
    for i, p in enumerate(points):
        out[i], failed[i] = oracle.horizons(scene, p["x_m"], p["y_m"], altitudes[i], int(p["own_footprint"]))
    return out, failed


def range_and_tangent_masks():
    """Per record float: is it a range (else a tangent or a steepest)."""
    is_range = np.zeros(abi.HORIZON_FLOATS, bool)
    for base, sectors in ((1, abi.TERRAIN_SECTORS), (abi.TERRAIN_HORIZON_FLOATS + 1, abi.BUILDING_SECTORS)):
        for s in range(sectors):
            at = base + s * SECTOR_FLOATS
            is_range[at + 1 : at + SECTOR_FLOATS : 2] = True
    return is_range


def steepest_is_the_max(rec):
    """A record's steepest values are the max of its band tangents (per sector, then of all)."""
    for base, sectors in ((0, abi.TERRAIN_SECTORS), (abi.TERRAIN_HORIZON_FLOATS, abi.BUILDING_SECTORS)):
        body = rec[:, base + 1 : base + 1 + sectors * SECTOR_FLOATS].reshape(len(rec), sectors, SECTOR_FLOATS)
        tangents = body[:, :, 2::2]
        with np.errstate(invalid="ignore"):
            most = np.where(np.isnan(tangents), -np.inf, tangents).max(axis=-1)
        assert np.array_equal(body[:, :, 0], most)
        assert np.array_equal(rec[:, base], most.max(axis=-1))


def compare(got, want, label, altitude_tol_m=0.0):
    """Equal NaN / infinite patterns, ranges within 1e-6 relative, tangents within 1e-6 relative
    (and `altitude_tol_m` / range, where the two sides' altitudes may differ by that)."""
    assert np.array_equal(np.isnan(got), np.isnan(want)), label
    assert np.array_equal(np.isinf(got), np.isinf(want)), label
    steepest_is_the_max(got)
    steepest_is_the_max(want)
    finite = np.isfinite(want)
    is_range = np.broadcast_to(range_and_tangent_masks(), want.shape)
    g, w = got.astype(np.float64), want.astype(np.float64)
    r = finite & is_range
    assert np.all(np.abs(g[r] - w[r]) <= 1e-6 * np.abs(w[r]) + 1e-6), label
    t = np.zeros_like(r)
    t[:, 1:] = r[:, :-1]  # each band's tangent follows its range
    bound = 1e-6 * np.abs(w[t]) + 1e-7 + altitude_tol_m / w[np.roll(t, -1, axis=1)]
    assert np.all(np.abs(g[t] - w[t]) <= bound), (label, np.abs(g[t] - w[t]).max())


def test_plane_scene_horizons(gpu):
    print("case: plane and box horizons", flush=True)
    scene, _ = plane_scene()
    square = gpu.upload(scene, no_sources())
    points = np.zeros(5, abi.POINT_DTYPE)
    points["x_m"] = [0.0, 20.0, -100.0, 46.0, 46.0]
    points["y_m"] = [0.0, 30.0, 2.0, 0.0, 0.0]
    points["own_footprint"] = [0, 0, 0, 77, 0]
    altitudes = np.array([4.0, 30.0, 2.5, 8.0, 8.0])
    got, failed = gpu.horizons(square, points, altitudes)
    want, want_failed = oracle_records(scene, points, altitudes)
    assert np.all(failed == 0) and np.all(want_failed == 0)
    compare(got, want, "plane")
    square.close()


@pytest.mark.parametrize("fine", [False, True])
def test_tile_horizons_match_oracle(gpu, fine):
    print(f"case: tile horizons fine={fine}", flush=True)
    sc = scenes.tile_scene(scale=0.06)
    scene = sc.scene
    if fine:
        T = scenes.with_fine_terrain(scene.terrain, share=0.6, amplitude_m=6.0)
        scene = abi.Scene(T, scene.obstacles, scene.weather)
    square = gpu.upload(scene, sc.sources)
    rng = np.random.default_rng(31)
    inside = np.flatnonzero(sc.points["own_footprint"] != 0)
    pick = np.concatenate([rng.choice(sc.points.size, 40, replace=False), inside[:16]])
    points = sc.points[pick]
    assert (points["own_footprint"] != 0).sum() >= 8
    altitudes = oracle.receivers(scene, points)["altitude_m"] + rng.uniform(-3.0, 40.0, points.size)
    got, failed = gpu.horizons(square, points, altitudes)
    want, want_failed = oracle_records(scene, points, altitudes)
    assert np.array_equal(failed, want_failed)
    compare(got, want, f"fine={fine}")
    # Buildings are seen: most records hold building edges, and drop the walls of their own.
    assert (np.isfinite(got[:, abi.TERRAIN_HORIZON_FLOATS])).mean() > 0.8
    # At the receivers' own altitude (ground + 4 m in FP32).
    got0, _ = gpu.horizons(square, points)
    own = oracle.receivers(scene, points)["altitude_m"]
    want0, _ = oracle_records(scene, points, own)
    compare(got0, want0, f"own altitude fine={fine}", altitude_tol_m=1e-4)  # the card's altitude is FP32
    square.close()


def test_card_records_screen_like_the_oracles(gpu):
    """The aircraft grid screened by the card's records equals the grid screened by the oracle's."""
    print("case: aircraft screening from card horizons", flush=True)
    sc = scenes.tile_scene(scale=0.06)
    square = gpu.upload(sc.scene, sc.sources)
    air = scenes.aircraft_scene(scale=0.02)
    nx, ny = air.grid_shape
    n = nx * ny
    points = np.zeros(n, abi.POINT_DTYPE)
    xy = np.asarray(air.points_xy, np.float64).reshape(-1, 2)
    # The aircraft receivers moved onto the tile (its terrain and buildings give their horizons).
    span = sc.grid_shape[0] * 10.0
    points["x_m"] = np.mod(xy[:, 0] - xy[:, 0].min(), span * 0.8) + span * 0.1
    points["y_m"] = np.mod(xy[:, 1] - xy[:, 1].min(), span * 0.8) + span * 0.1
    alt = oracle.receivers(sc.scene, points)["altitude_m"]
    got, failed = gpu.horizons(square, points, alt)
    want, _ = oracle_records(sc.scene, points, alt)
    compare(got, want, "aircraft receivers")
    of = np.arange(n, dtype=np.uint32)
    a = gpu.aircraft_grid(air, screening=dict(horizons=got, horizon_of=of))
    b = gpu.aircraft_grid(air, screening=dict(horizons=want, horizon_of=of))
    with np.errstate(divide="ignore"):
        da = 10 * np.log10(a["period_energy"])
        dbb = 10 * np.log10(b["period_energy"])
    both = np.isfinite(da) & np.isfinite(dbb)
    assert np.array_equal(np.isfinite(da), np.isfinite(dbb))
    assert np.all(np.abs(da[both] - dbb[both]) < 0.01)
    square.close()
