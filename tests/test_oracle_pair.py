"""The oracle's whole pair (evaluate_pair, evaluate_line_pair, paint_points) against closed forms.

The pieces of a pair are pinned elsewhere (ISO/TR 17534-4, the ray tests' transfer constants, the
line quadrature). Here the composition is pinned on flat scenes with no buildings, where CNOSSOS
leaves no diffraction and the ground term has a closed form:

- the point divergence 1 / (max(hypot(max(d, r_excl), dz), 1)^2 * 10^1.1)
  (popup/src/evaluate.rs:195-203), with the receiver at ground + 4 m (popup/src/answer.rs:47);
- hard ground (G = 0 everywhere, source G = 0): A_ground = -3 dB in homogeneous air, and the
  favourable floor -3 (1 + 2 (1 - 30 (zs + zr) / dp)) beyond dp = 30 (zs + zr), -3 dB within
  (physics/src/cnossos/ground.rs:92-101,116-128);
- air absorption 10^(-alpha * slant / 10,000) over the slant between the source and the receiver
  (physics/src/ray.rs:155-190), the exclusion radius acting on the divergence only;
- the period's p by the 16-sector lerp of the bearing of source -> receiver
  (physics/src/weather.rs:61-70), mixed in energy p * T_F + (1 - p) * T_H (physics/src/ray.rs:178-181);
- the reference's pinned transfer constants 0.9097167331685742 and 1.4125375446227544 for G = 0.5
  over 110.54 m (physics/src/ray_tests.rs:223-276), here through a whole pair on a lattice;
- Lden(60, 55, 50) = 60 dB through a pair (physics/src/bands.rs:59-84);
- a line piece over hard ground in calm, still air equals twice the 400,000-step point sum of its
  divergence (physics/src/line_tests.rs:49-110 through popup/src/evaluate.rs:171-267).

A plausible slip in the composition (a squared exclusion radius, the slant taken from the
exclusion radius, the sector lerp reversed or offset by half a sector, periods permuted, the
receiver height or the 10^1.1 dropped) fails one of these checks.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi, scenes

POINT_DIVERGENCE = 10 ** 1.1
# ISO 9613-1 at 15 C / 70 % per band, as the weather node stores it (0.01 dB/km).
ALPHA_CODED = np.round(scenes.ALPHA_15C_70 * 100.0) / 100.0


def flat_scene(height_m=0.0, impervious=100, percent=None, alpha=ALPHA_CODED, extent=3000.0):
    """A flat square from (0, 0) to (extent, extent): terrain at `height_m` on a 30 m lattice with
    imperviousness `impervious` percent at every node, no obstacles, one weather node pattern."""
    spacing = 30.0
    n = int(extent / spacing) + 3
    nodes = abi.Terrain.pack(np.full((n, n), height_m), np.full((n, n), impervious))
    terrain = abi.Terrain(-spacing, -spacing, spacing, nodes)
    obstacles = scenes._obstacles([], [], [], [], (0.0, 0.0), (extent, extent))
    if percent is None:
        percent = np.full((3, 16), 50)
    weather = scenes._weather((0.0, 0.0), (extent, extent), lambda r, c: percent, alpha=alpha)
    return abi.Scene(terrain, obstacles, weather)


def point(x, y):
    p = np.zeros(1, abi.POINT_DTYPE)
    p["x_m"], p["y_m"] = x, y
    return p


def source(x, y, height_m, energy, exclusion_m=0.0, ground_percent=0):
    s = np.zeros(1, abi.SOURCE_DTYPE)
    s["x_m"], s["y_m"], s["height_m"] = x, y, height_m
    s["exclusion_radius_m"] = exclusion_m
    s["ground_percent"] = ground_percent
    s["energy"] = energy
    return s


def sector_p(row_percent, sx, sy, rx, ry):
    """p of one period for the bearing of source -> receiver: clockwise from north, the sector
    centres at multiples of 22.5 degrees, linear between the two nearest."""
    bearing = math.degrees(math.atan2(rx - sx, ry - sy)) % 360.0
    k = bearing / 22.5
    lo = int(math.floor(k)) % 16
    f = k - math.floor(k)
    return (row_percent[lo] * (1.0 - f) + row_percent[(lo + 1) % 16] * f) / 100.0


def hard_ground_transfers(horizontal, source_h, alpha):
    """(T_H, T_F) per band over flat hard ground: 10^(-A/10) times the air's transfer."""
    zs, zr = source_h, abi.RECEIVER_HEIGHT_M
    slant = math.hypot(horizontal, zr - zs)
    air = 10.0 ** (-np.asarray(alpha) * slant / 10000.0)
    limit = 30.0 * (zs + zr)
    floor = -3.0 if horizontal <= limit else -3.0 * (1.0 + 2.0 * (1.0 - limit / horizontal))
    return 10.0 ** 0.3 * air, 10.0 ** (-floor / 10.0) * air


def divergence(horizontal, source_h, exclusion_m):
    dz = abi.RECEIVER_HEIGHT_M - source_h
    d = max(math.hypot(max(horizontal, exclusion_m), dz), 1.0)
    return 1.0 / (d * d * POINT_DIVERGENCE)


CASES = [  # (horizontal, bearing deg, source height, exclusion radius)
    (3.0, 10.0, 0.05, 0.0),
    (3.0, 100.0, 0.05, 12.0),
    (25.0, 33.75, 0.5, 5.0),
    (80.0, 200.0, 0.05, 0.0),
    (121.0, 271.0, 0.05, 30.0),
    (400.0, 359.0, 1.0, 0.0),
    (400.0, 45.0, 0.05, 800.0),
    (1350.0, 168.75, 0.5, 0.0),
]


@pytest.mark.parametrize("horizontal,bearing,h,excl", CASES)
def test_flat_hard_ground_pair_is_the_closed_form(horizontal, bearing, h, excl):
    rng = np.random.default_rng(int(horizontal * 10 + bearing))
    percent = np.stack([5 + 5 * np.arange(16), 100 - 6 * np.arange(16), (np.arange(16) * 37) % 101])
    sc = flat_scene(percent=percent)
    sx, sy = 1500.0, 1500.0
    b = math.radians(bearing)
    rx, ry = sx + horizontal * math.sin(b), sy + horizontal * math.cos(b)
    energy = 10.0 ** (rng.normal(80.0, 6.0, (3, 8)) / 10.0)
    src = source(sx, sy, h, energy, exclusion_m=excl)
    got = oracle.evaluate_pairs(sc, point(rx, ry), src, [0], [0])
    assert got["failed"][0] == 0
    e = src["energy"][0].astype(np.float64)  # the f32 the ABI carries
    th, tf = hard_ground_transfers(horizontal, h, ALPHA_CODED)
    w = divergence(horizontal, h, excl)
    for period in range(3):
        p = sector_p(percent[period], sx, sy, rx, ry)
        calm = w * float(e[period] @ th)
        downwind = w * float(e[period] @ tf)
        mixed = p * downwind + (1.0 - p) * calm
        assert got["sums"][0, 3 + period] == pytest.approx(calm, rel=1e-9)
        assert got["sums"][0, 6 + period] == pytest.approx(downwind, rel=1e-9)
        assert got["sums"][0, period] == pytest.approx(mixed, rel=1e-9)


def test_the_reference_transfer_constants_through_a_whole_pair():
    # G = 0.5 on flat ground at 200 m, source 1 m with G_s = 0.5, receiver 4 m up 110.54 m away:
    # per band states = divergence x constant x air (physics/src/ray_tests.rs:223-276). Period 0
    # calm, period 1 favourable, period 2 half and half; one band lit at a time.
    percent = np.stack([np.zeros(16), np.full(16, 100), np.full(16, 50)]).astype(int)
    sc = flat_scene(height_m=200.0, impervious=50, percent=percent)
    sx, sy, horizontal = 1000.0, 1000.0, 110.54
    slant = math.hypot(horizontal, 3.0)
    w = 1.0 / (slant * slant * POINT_DIVERGENCE)
    for band, calm_k, downwind_k in ((4, 0.9097167331685742, 1.4125375446227544), (0, 1.4125375446227544, None)):
        energy = np.zeros((3, 8))
        energy[:, band] = 1.0
        src = source(sx, sy, 1.0, energy, ground_percent=50)
        got = oracle.evaluate_pairs(sc, point(sx + horizontal, sy), src, [0], [0])
        air = 10.0 ** (-ALPHA_CODED[band] * slant / 10000.0)
        assert got["sums"][0, 3] / (w * air) == pytest.approx(calm_k, abs=1e-9)
        if downwind_k is not None:
            assert got["sums"][0, 6] / (w * air) == pytest.approx(downwind_k, abs=1e-9)
        # The periods' p: 0, 1 and 0.5.
        calm, downwind = got["sums"][0, 3], got["sums"][0, 6]
        assert got["sums"][0, 0] == pytest.approx(calm, rel=1e-12)
        assert got["sums"][0, 1] == pytest.approx(downwind, rel=1e-12)
        assert got["sums"][0, 2] == pytest.approx(0.5 * calm + 0.5 * downwind, rel=1e-12)


def test_lden_60_55_50_through_a_pair():
    sc = flat_scene()
    sx, sy, horizontal, h = 500.0, 700.0, 260.0, 0.05
    th, tf = hard_ground_transfers(horizontal, h, ALPHA_CODED)
    w = divergence(horizontal, h, 0.0)
    shape = 10.0 ** (np.array([-26.2, -16.1, -8.6, -3.2, 0.0, 1.2, 1.0, -1.1]) / 10.0)
    received = w * (0.5 * (shape @ tf) + 0.5 * (shape @ th))  # p = 0.5 in every sector
    target = 10.0 ** (np.array([60.0, 55.0, 50.0]) / 10.0)
    energy = (target / received)[:, None] * shape[None, :]
    src = source(sx, sy, h, energy)
    got = oracle.evaluate_pairs(sc, point(sx, sy + horizontal), src, [0], [0])
    mixed = got["sums"][0, :3]
    assert 10 * np.log10(mixed) == pytest.approx([60.0, 55.0, 50.0], abs=1e-6)
    assert 10 * math.log10(oracle.lden_energy(mixed)) == pytest.approx(60.0, abs=1e-6)


def test_paint_points_sums_the_sources_within_the_cutoff():
    sc = flat_scene()
    rng = np.random.default_rng(17)
    n = 40
    srcs = np.zeros(n, abi.SOURCE_DTYPE)
    srcs["x_m"], srcs["y_m"] = rng.uniform(200, 2800, n), rng.uniform(200, 2800, n)
    srcs["height_m"] = 0.05
    srcs["energy"] = 10.0 ** (rng.normal(80.0, 6.0, (n, 3, 8)) / 10.0)
    pts = np.concatenate([point(1500.0, 1500.0), point(900.0, 2100.0)])
    cutoff = 1000.0
    got = oracle.paint_points(sc, pts, srcs, cutoff)
    for i in range(pts.size):
        rx, ry = pts["x_m"][i], pts["y_m"][i]
        want = np.zeros(3)
        count = 0
        for s in srcs:
            horizontal = math.hypot(s["x_m"] - rx, s["y_m"] - ry)
            assert abs(horizontal - cutoff) > 1e-6
            if horizontal > cutoff:
                continue
            count += 1
            th, tf = hard_ground_transfers(horizontal, 0.05, ALPHA_CODED)
            w = divergence(horizontal, 0.05, 0.0)
            e = s["energy"].astype(np.float64)
            want += w * (0.5 * (e @ th) + 0.5 * (e @ tf))
        assert got["paths"][i] == count > 0
        assert got["period_energy"][i] == pytest.approx(want, rel=1e-9)


def point_sum(start, end, receiver, steps=400_000):
    f = (np.arange(steps) + 0.5) / steps
    p = np.asarray(start)[None, :] + f[:, None] * (np.asarray(end) - np.asarray(start))[None, :]
    p = p - np.asarray(receiver)[None, :]
    length = float(np.linalg.norm(np.asarray(end) - np.asarray(start)))
    return float(np.sum(1.0 / np.sum(p * p, axis=1))) * length / steps / POINT_DIVERGENCE


@pytest.mark.parametrize("a,b,r,h", [
    ((1250.0, 1500.0), (1750.0, 1500.0), (1500.0, 1560.0), 0.05),
    ((1000.0, 1000.0), (1250.0, 1000.0), (1500.0, 1030.0), 0.05),
    ((1400.0, 1200.0), (1650.0, 1450.0), (1450.0, 1500.0), 0.5),
    ((1495.0, 1500.0), (1505.0, 1500.0), (1500.0, 1501.0), 0.05),
])
def test_line_pair_in_calm_still_air_is_twice_the_point_sum(a, b, r, h):
    # Hard ground, p = 0 and no air absorption: every node's transfer is 10^0.3 in every band, so
    # the pair is 2 x energy per metre x the divergence integral along the piece.
    sc = flat_scene(percent=np.zeros((3, 16), int), alpha=np.zeros(8))
    piece = np.zeros(1, abi.LINE_PIECE_DTYPE)
    piece["ax_m"], piece["ay_m"], piece["bx_m"], piece["by_m"] = a[0], a[1], b[0], b[1]
    piece["height_m"] = h
    rng = np.random.default_rng(3)
    piece["energy_per_m"] = 10.0 ** (rng.normal(70.0, 6.0, (3, 8)) / 10.0)
    got = oracle.evaluate_line_pairs(sc, point(*r), piece, [0], [0])
    assert got["failed"][0] == 0
    integral = point_sum((a[0], a[1], h), (b[0], b[1], h), (r[0], r[1], abi.RECEIVER_HEIGHT_M))
    e = piece["energy_per_m"][0].astype(np.float64)
    for period in range(3):
        want = 10.0 ** 0.3 * integral * e[period].sum()
        assert 10 * math.log10(got["sums"][0, period] / want) == pytest.approx(0.0, abs=1e-5)
        assert got["sums"][0, 3 + period] == pytest.approx(got["sums"][0, period], rel=1e-12)
