"""The oracle's scene side: the lattice walk of the ground under a ray (popup/src/scene.rs:242-420)
and the cell walk of obstacle crossings (popup/src/obstacles/crossings.rs:163-230), pinned by
brute force and by the reference's crossing semantics (popup/src/obstacles/tests/crossings.rs)."""
import math

import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi, scenes


@pytest.fixture(scope="module")
def road():
    return scenes.road_scene()


@pytest.fixture(scope="module")
def tile():
    return scenes.tile_scene(scale=0.06)


def brute_force_profile(terrain, s, r):
    """Every lattice line crossing by direct enumeration, sampled between that line's two nodes."""
    t_list = []
    ox, oy, sp = terrain.origin_x_m, terrain.origin_y_m, terrain.spacing_m
    ca, cb = (s[0] - ox) / sp, (r[0] - ox) / sp
    ra, rb = (s[1] - oy) / sp, (r[1] - oy) / sp
    for c in range(int(math.floor(min(ca, cb))), int(math.ceil(max(ca, cb))) + 1):
        if cb != ca:
            t = (c - ca) / (cb - ca)
            if 0 < t < 1:
                t_list.append(t)
    for k in range(int(math.floor(min(ra, rb))), int(math.ceil(max(ra, rb))) + 1):
        if rb != ra:
            t = (k - ra) / (rb - ra)
            if 0 < t < 1:
                t_list.append(t)
    return sorted(set(t_list))


def test_profile_vertices_are_every_lattice_line(tile):
    rng = np.random.default_rng(5)
    T = tile.scene.terrain
    ext = tile.grid_shape[0] * 10.0
    for _ in range(50):
        s, r = rng.uniform(0, ext, 2), rng.uniform(0, ext, 2)
        failed, p = oracle.fill_profile(T, s, r)
        assert failed == 0
        expected = brute_force_profile(T, s, r)
        assert len(p["t"]) == len(expected) + 2
        np.testing.assert_allclose(p["t"][1:-1], expected, atol=1e-12)
        assert np.all(np.diff(p["t"]) > 0)
        # Each vertex reads the bilinear surface on its lattice line (linear there).
        for t, z in zip(p["t"], p["ground_m"]):
            x, y = s[0] + t * (r[0] - s[0]), s[1] + t * (r[1] - s[1])
            _, g = oracle.ground_at(T, x, y)
            assert abs(g[0] - z) < 1e-6


def test_crossings_match_brute_force(road, tile):
    rng = np.random.default_rng(7)
    for sc in (road, tile):
        O = sc.scene.obstacles
        ext = sc.grid_shape[0] * 10.0
        total = 0
        for _ in range(300):
            s, r = rng.uniform(0, ext, 2), rng.uniform(0, ext, 2)
            f1, walk = oracle.crossings(O, s, r)
            f2, brute = oracle.crossings(O, s, r, brute_force=True)
            assert f1 == 0 and f2 == 0
            assert walk.tobytes() == brute.tobytes()
            total += len(walk)
        assert total > 100


def square(o, x0, y0, side, height=10.0, fid=1):
    return [(x0, y0), (x0 + side, y0), (x0 + side, y0 + side), (x0, y0 + side)]


def test_crossing_semantics_square_and_tangent():
    """A ray through a square crosses it twice, both bounding its roof; a ray through its corner
    vertex alone is a tangent contact: one top that bounds no roof (crossings.rs:80-111, 34-56)."""
    mps = scenes.METRES_PER_STEP
    ring = [(100.0, 100.0), (120.0, 100.0), (120.0, 120.0), (100.0, 120.0)]
    O = scenes._obstacles([ring], [12.0], [abi.KIND_BUILDING], [7], (0.0, 0.0), (400.0, 400.0))
    _, c = oracle.crossings(O, (50.0, 110.0), (200.0, 110.0))
    assert len(c) == 2 and list(c["bounds_roof"]) == [1, 1] and list(c["footprint_id"]) == [7, 7]
    assert abs(c["height_m"][0] - 12.0) < 1e-12
    np.testing.assert_allclose(c["t"] * 150.0, [50.0, 70.0], atol=1e-9)
    # Through the corner (100, 100) from below-left, leaving away from the square: tangent.
    _, c = oracle.crossings(O, (80.0, 120.0), (120.0, 80.0))
    assert len(c) == 1 and c["bounds_roof"][0] == 0
    # Through the corner into the square: enters at the vertex, one crossing that bounds.
    _, c = oracle.crossings(O, (80.0, 80.0), (110.0, 110.0))
    assert len(c) == 1 and c["bounds_roof"][0] == 1
    # A ray sliding along a wall grazes it (collinear overlap is no crossing), the far walls count.
    _, c = oracle.crossings(O, (50.0, 100.0), (200.0, 100.0))
    assert len(c) == 2 and list(c["bounds_roof"]) == [0, 0]


def test_endpoints_excluded_and_symmetry():
    ring = [(100.0, 100.0), (120.0, 100.0), (120.0, 120.0), (100.0, 120.0)]
    O = scenes._obstacles([ring], [12.0], [abi.KIND_BUILDING], [7], (0.0, 0.0), (400.0, 400.0))
    _, c = oracle.crossings(O, (100.0, 110.0), (200.0, 110.0))  # starts on the west wall
    assert len(c) == 1
    _, a = oracle.crossings(O, (50.0, 107.0), (210.0, 113.0))
    _, b = oracle.crossings(O, (210.0, 113.0), (50.0, 107.0))
    np.testing.assert_allclose(np.sort(a["t"]), np.sort(1.0 - b["t"]), atol=1e-12)
