"""The horizon builder of the oracle (Horizons::build, popup/src/aircraft/horizons.rs:138-201, and
write_f32) against what fixes it on a scene of known shape: ground on a tilted plane (the lattice's
bilinear read reproduces a plane, so each sample's tangent is (slope * east - altitude) / range)
and one square building, whose walls a ray meets at ranges found here by plain ray-segment
geometry. Checks each sector's steepest, each band's edge (the steepest offered in its range band,
bands closed at their limit plus half a metre), the receiver's own walls dropped, and the float
layout (steepest of all, then per sector its steepest and six (range, tangent) pairs, NaN where a
band is empty)."""
import math

import numpy as np

from oracle import oracle
from quietmap_amd import abi, scenes

SLOPE = 0.01  # metres up per metre east
SPACING = 100.0
HALF_SIDE = 9_000.0
BOX = (40.0, -6.0, 52.0, 6.0)  # x0, y0, x1, y1 of the building, metres (multiples of 0.2 m)
BOX_HEIGHT = 14.0
BOX_ID = 77
TERRAIN_LIMITS = [500.0, 1000.0, 2000.0, 3500.0, 5500.0, 8000.0]
BUILDING_LIMITS = [16.0, 32.0, 64.0, 128.0, 256.0, 512.0]
SECTOR_FLOATS = 1 + 2 * abi.HORIZON_BANDS


def plane_scene():
    n = int(round(2 * HALF_SIDE / SPACING)) + 1
    xs = -HALF_SIDE + SPACING * np.arange(n)
    heights = np.tile(SLOPE * xs, (n, 1))  # rows run north, columns east
    terrain = abi.Terrain(-HALF_SIDE, -HALF_SIDE, SPACING, abi.Terrain.pack(heights, np.zeros((n, n))))
    x0, y0, x1, y1 = BOX
    ring = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    obstacles = scenes._obstacles([ring], [BOX_HEIGHT], [abi.KIND_BUILDING], np.array([BOX_ID], np.uint64),
                                  (-600.0, -600.0), (1200.0, 1200.0))
    weather = scenes._weather((-1000.0, -1000.0), (2000.0, 2000.0), lambda r, c: np.full((3, 16), 50))
    return abi.Scene(terrain, obstacles, weather), ring


def band(r, limits):
    for k, limit in enumerate(limits):
        if r <= limit + 0.5:
            return k
    return len(limits) - 1


def direction(sector, sectors):
    a = (sector + 0.5) * 2.0 * math.pi / sectors
    return math.cos(a), math.sin(a)


def wall_ranges(ring, x, y, east, north, reach):
    """Ranges at which the ray from (x, y) meets the ring's edges within (0, reach)."""
    out = []
    for i in range(len(ring)):
        (ax, ay), (bx, by) = ring[i], ring[(i + 1) % len(ring)]
        ex, ey = bx - ax, by - ay
        den = east * ey - north * ex
        if den == 0.0:
            continue
        wx, wy = ax - x, ay - y
        t = (wx * ey - wy * ex) / den
        u = (wx * north - wy * east) / den
        if 0.0 < t < reach and 0.0 <= u <= 1.0:
            out.append(t)
    return sorted(out)


def expected(x, y, altitude, ring, own):
    terrain = np.full((abi.TERRAIN_SECTORS, SECTOR_FLOATS), np.nan)
    growth = (8000.0 / 30.0) ** (1.0 / 47.0)
    for s in range(abi.TERRAIN_SECTORS):
        east, north = direction(s, abi.TERRAIN_SECTORS)
        edges = {}
        for k in range(48):
            r = 30.0 * growth**k
            t = (SLOPE * (x + east * r) - altitude) / r
            b = band(r, TERRAIN_LIMITS)
            if b not in edges or t > edges[b][1]:
                edges[b] = (r, t)
        terrain[s, 0] = max(t for _, t in edges.values())
        for b, (r, t) in edges.items():
            terrain[s, 1 + 2 * b], terrain[s, 2 + 2 * b] = r, t
    buildings = np.full((abi.BUILDING_SECTORS, SECTOR_FLOATS), np.nan)
    buildings[:, 0] = -np.inf
    for s in range(abi.BUILDING_SECTORS):
        east, north = direction(s, abi.BUILDING_SECTORS)
        if own:
            continue
        for r in wall_ranges(ring, x, y, east, north, 512.0):
            roof = SLOPE * (x + east * r) + BOX_HEIGHT
            t = (roof - altitude) / r
            b = band(r, BUILDING_LIMITS)
            if np.isnan(buildings[s, 1 + 2 * b]) or t > buildings[s, 2 + 2 * b]:
                buildings[s, 1 + 2 * b], buildings[s, 2 + 2 * b] = r, t
            buildings[s, 0] = max(buildings[s, 0], t)
    return terrain, buildings


def split(out):
    t = out[: abi.TERRAIN_HORIZON_FLOATS]
    b = out[abi.TERRAIN_HORIZON_FLOATS :]
    return t[0], t[1:].reshape(abi.TERRAIN_SECTORS, SECTOR_FLOATS), b[0], b[1:].reshape(abi.BUILDING_SECTORS,
                                                                                         SECTOR_FLOATS)


def check(got, want, rtol):
    assert np.array_equal(np.isnan(got), np.isnan(want))
    finite = np.isfinite(want)
    assert np.array_equal(np.isinf(got), np.isinf(want))
    assert np.allclose(got[finite], want[finite], rtol=rtol, atol=1e-6)


def test_layout_size():
    assert abi.HORIZON_FLOATS == 1 + 32 * SECTOR_FLOATS + 1 + 256 * SECTOR_FLOATS


def test_plane_and_box_horizons():
    scene, ring = plane_scene()
    for x, y, altitude in ((0.0, 0.0, 4.0), (20.0, 30.0, 30.0), (-100.0, 2.0, 2.5)):
        out, failed = oracle.horizons(scene, x, y, altitude)
        assert failed == 0
        steepest_t, terrain, steepest_b, buildings = split(out.astype(np.float64))
        want_t, want_b = expected(x, y, altitude, ring, own=False)
        check(terrain, want_t, 1e-6)
        check(buildings, want_b, 1e-6)
        assert abs(steepest_t - want_t[:, 0].max()) < 1e-6
        assert steepest_b == want_b[:, 0].max() or abs(steepest_b - want_b[:, 0].max()) < 1e-6
        if abs(y) < BOX[3]:
            # The building stands due east: the sectors either side of east hold it, none facing west.
            assert np.isfinite(buildings[0, 0]) and np.isfinite(buildings[255, 0])
            assert np.all(np.isnan(buildings[64:192, 1:]))
        assert np.isfinite(buildings[:, 0]).sum() >= 3


def test_own_building_is_no_horizon():
    scene, ring = plane_scene()
    inside = (46.0, 0.0, 8.0)
    out, failed = oracle.horizons(scene, *inside, own_footprint=BOX_ID)
    assert failed == 0
    _, _, steepest_b, buildings = split(out)
    assert steepest_b == -np.inf and np.all(np.isnan(buildings[:, 1:])) and np.all(buildings[:, 0] == -np.inf)
    # Seen from inside as someone else's wall, every sector meets it once, within 12 m.
    out, failed = oracle.horizons(scene, *inside, own_footprint=0)
    _, _, _, buildings = split(out.astype(np.float64))
    want_t, want_b = expected(*inside, ring, own=False)
    check(buildings, want_b, 1e-6)
    assert np.all(np.isfinite(buildings[:, 1]) ^ np.isfinite(buildings[:, 3]))


def test_off_terrain_samples_are_left_out():
    """Samples beyond the lattice offer nothing: near the east edge the far bands are empty."""
    scene, _ = plane_scene()
    out, failed = oracle.horizons(scene, HALF_SIDE - 700.0, 0.0, 100.0)
    _, terrain, _, _ = split(out)
    east = terrain[0]
    assert np.isfinite(east[1]) and np.isnan(east[1 + 2 * 5])
    west = terrain[16]
    assert np.all(np.isfinite(west[1:]))
