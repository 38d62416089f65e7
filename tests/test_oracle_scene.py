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


def _coarse_height(T, x, y):
    """Bilinear height of the coarse lattice at (x, y), numpy."""
    cp, rp = (x - T.origin_x_m) / T.spacing_m, (y - T.origin_y_m) / T.spacing_m
    c, r = int(math.floor(cp)), int(math.floor(rp))
    fc, fr = cp - c, rp - r
    h = -500.0 + (T.nodes & 0xFFFF).astype(np.float64) * 0.2
    c1, r1 = c + (fc > 0), r + (fr > 0)
    return ((1 - fr) * ((1 - fc) * h[r, c] + fc * h[r, c1]) + fr * ((1 - fc) * h[r1, c] + fc * h[r1, c1]))


def _detail(T, x, y):
    """The fine detail at (x, y) where its block has fine terrain, numpy (bilinear on the fine
    lattice, which is the line's linear reading on a fine line)."""
    cp, rp = (x - T.origin_x_m) / T.spacing_m, (y - T.origin_y_m) / T.spacing_m
    by = min(max(int(math.floor(rp / abi.FINE_BLOCK_CELLS)), 0), T.fine_blocks.shape[0] - 1)
    bx = min(max(int(math.floor(cp / abi.FINE_BLOCK_CELLS)), 0), T.fine_blocks.shape[1] - 1)
    if not T.fine_blocks[by, bx]:
        return 0.0
    fr, fc = max(rp * abi.FINE_PER_NODE, 0.0), max(cp * abi.FINE_PER_NODE, 0.0)
    R, Cn = T.fine_detail.shape
    r, c = min(int(math.floor(fr)), R - 2), min(int(math.floor(fc)), Cn - 2)
    a, b = fr - r, fc - c
    d = T.fine_detail.astype(np.float64)
    return ((1 - a) * ((1 - b) * d[r, c] + b * d[r, c + 1]) + a * ((1 - b) * d[r + 1, c] + b * d[r + 1, c + 1])) * 0.2


def _lines(lo_pos, hi_pos, t_from, t_to):
    """t of every integer line strictly between fractions t_from and t_to of a ray from lo_pos
    (t = 0) to hi_pos (t = 1), in lattice units."""
    out = []
    if hi_pos == lo_pos:
        return out
    for k in range(int(math.floor(min(lo_pos, hi_pos))) - 1, int(math.ceil(max(lo_pos, hi_pos))) + 2):
        t = (k - lo_pos) / (hi_pos - lo_pos)
        if t_from < t < t_to:
            out.append(t)
    return out


def _fine_near(T, a, b):
    for x in (min(a[0], b[0]), max(a[0], b[0])):
        for y in (min(a[1], b[1]), max(a[1], b[1])):
            cp, rp = (x - T.origin_x_m) / T.spacing_m, (y - T.origin_y_m) / T.spacing_m
            by = min(max(int(math.floor(rp / abi.FINE_BLOCK_CELLS)), 0), T.fine_blocks.shape[0] - 1)
            bx = min(max(int(math.floor(cp / abi.FINE_BLOCK_CELLS)), 0), T.fine_blocks.shape[1] - 1)
            if T.fine_blocks[by, bx]:
                return True
    return False


def test_fine_terrain_profile_is_every_line_of_each_stretch(tile):
    """With fine terrain (C3'): within 100 m of either end, where a block under that stretch has
    fine terrain, every line of the 1/6 lattice; every coarse line in between; a vertex where the
    lattice changes; each vertex's height the coarse bilinear ground, plus on the fine lattice and at
    the ends and switches (Ground::at) the fine detail of its block (numpy bilinear, brute force)."""
    T = scenes.with_fine_terrain(tile.scene.terrain)
    rng = np.random.default_rng(8)
    ext = tile.grid_shape[0] * 10.0
    fine_rays = 0
    for trial in range(60):
        s = rng.uniform(50, ext - 50, 2)
        length = 60.0 if trial % 5 == 0 else rng.uniform(150, 900)
        ang = rng.uniform(0, 2 * math.pi)
        r = s + length * np.array([math.cos(ang), math.sin(ang)])
        if not (0 < r[0] < ext and 0 < r[1] < ext):
            continue
        failed, p = oracle.fill_profile(T, s, r)
        assert failed == 0
        L = math.hypot(*(r - s))
        along = lambda t: s + t * (r - s)
        if L <= 200.0:
            stretches = [(0.0, 1.0, 6 if _fine_near(T, s, r) else 1)]
        else:
            e = 100.0 / L
            stretches = []
            for a, b, sc in ((0.0, e, 6 if _fine_near(T, s, along(e)) else 1), (e, 1 - e, 1),
                             (1 - e, 1.0, 6 if _fine_near(T, along(1 - e), r) else 1)):
                if stretches and stretches[-1][2] == sc:
                    stretches[-1] = (stretches[-1][0], b, sc)
                else:
                    stretches.append((a, b, sc))
        expected, fine_at = [], []
        for a, b, sc in stretches:
            unit = T.spacing_m / sc
            ts = _lines((s[0] - T.origin_x_m) / unit, (r[0] - T.origin_x_m) / unit, a, b)
            ts += _lines((s[1] - T.origin_y_m) / unit, (r[1] - T.origin_y_m) / unit, a, b)
            ts = sorted(set(np.round(ts, 12)))
            expected += ts
            fine_at += [sc == 6] * len(ts)
            if b < 1.0:
                expected.append(b)
                fine_at.append(True)  # Ground::at reads the fine height
            fine_rays += sc == 6
        assert len(p["t"]) == len(expected) + 2, trial
        np.testing.assert_allclose(p["t"][1:-1], expected, atol=1e-9)
        for t, z, fine in zip(p["t"], p["ground_m"], [True] + fine_at + [True]):
            x, y = along(t)
            want = _coarse_height(T, x, y) + (_detail(T, x, y) if fine else 0.0)
            assert abs(want - z) < 1e-6, (trial, t)
    assert fine_rays >= 20
