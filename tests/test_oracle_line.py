"""The line-source quadrature of the oracle against what pins it: the reference's own checks
(physics/src/line_tests.rs:1-192: the free-field quadrature equals a 400,000-step incoherent point
sum within 1e-6 dB on six geometries, sits 1.9533 dB above the retired 2-pi chain, an infinite
line reads -10 lg d - 6.0285, the in-plane angle's 1.899 dB beside a short elevated line, the
half-metre floor, nodes uniform in the in-plane angle, one node under 3 degrees, and a wall's
blocked and clear nodes), plus the skyline against a brute force over every edge."""
import math

import numpy as np
import pytest

from oracle import oracle

POINT_DIVERGENCE = 10 ** 1.1


def db(ratio):
    return 10.0 * math.log10(ratio)


def rel(p, r):
    return [p[i] - r[i] for i in range(3)]


def free_field(start, end, receiver):
    n = oracle.line_nodes(rel(start, receiver), rel(end, receiver))
    return n["weight"].sum() * n["divergence"]


def point_sum(start, end, receiver, steps=400_000):
    f = (np.arange(steps) + 0.5) / steps
    p = np.array(start)[None, :] + f[:, None] * (np.array(end) - np.array(start))[None, :] - np.array(receiver)[None, :]
    length = math.hypot(end[0] - start[0], end[1] - start[1])
    return float(np.sum(1.0 / np.sum(p * p, axis=1))) * length / steps / POINT_DIVERGENCE


@pytest.mark.parametrize("start,end,receiver", [
    ([-125.0, 0.0, 0.05], [125.0, 0.0, 0.05], [0.0, 50.0, 4.0]),
    ([0.0, 0.0, 0.05], [250.0, 0.0, 0.05], [500.0, 30.0, 4.0]),
    ([-125.0, 0.0, 0.05], [125.0, 0.0, 0.05], [0.0, 1.0, 4.0]),
    ([0.0, 0.0, 0.5], [250.0, 0.0, 12.0], [40.0, 90.0, 4.0]),
    ([2000.0, 0.0, 0.05], [2250.0, 0.0, 0.05], [0.0, 45.0, 4.0]),
    ([-5.0, 0.0, 0.05], [5.0, 0.0, 0.05], [0.0, 1.0, 4.0]),
])
def test_free_field_quadrature_is_the_exact_point_sum(start, end, receiver):
    assert abs(db(free_field(start, end, receiver) / point_sum(start, end, receiver))) < 1e-6


def test_the_line_is_1_9533_db_above_the_retired_chain_and_infinite_lines():
    theta = 2.0 * math.atan(125.0 / 1000.0)
    retired = -db(2.0 * math.pi * 1000.0) + db(theta / math.pi)
    gap = db(free_field([-125.0, 0.0, 4.0], [125.0, 0.0, 4.0], [0.0, 1000.0, 4.0])) - retired
    assert abs(gap - 1.9533) < 0.0005
    infinite = db(free_field([-1e7, 0.0, 4.0], [1e7, 0.0, 4.0], [0.0, 100.0, 4.0]))
    assert abs(infinite + db(100.0) + 6.0285) < 1e-3


def test_in_plane_angle_beside_a_short_elevated_line():
    slant = math.sqrt(1.0 + 3.95 * 3.95)
    retired = -db(2.0 * math.pi * slant) + db(2.0 * math.atan(5.0) / math.pi)
    normalization = db(2.0 * math.pi * math.pi / POINT_DIVERGENCE)
    excess = retired + normalization - db(free_field([-5.0, 0.0, 0.05], [5.0, 0.0, 0.05], [0.0, 1.0, 4.0]))
    assert abs(excess - 1.899) < 0.002


def test_receiver_on_the_line_and_degenerate_piece():
    n = oracle.line_nodes([-125.0, 0.0, 0.0], [125.0, 0.0, 0.0])
    assert n["divergence"] == 1.0 / (12.589254117941673 * 0.5)
    assert oracle.line_nodes([1.0, 2.0, 3.0], [1.0, 2.0, 3.0]) is None


def test_nodes_uniform_in_the_in_plane_angle():
    n = oracle.line_nodes([-125.0, 20.0, -3.95], [125.0, 20.0, -3.95])
    assert n["along"].size == 5
    perp = math.hypot(20.0, 3.95)
    angles = np.arctan((n["along"] - 125.0) / perp)
    assert np.allclose(np.diff(angles), n["subtended"] / 5, atol=1e-12)
    assert abs(n["along"][2] - 125.0) < 1e-9


def test_small_angle_takes_one_node():
    n = oracle.line_nodes([-10.0, 1000.0, -3.95], [10.0, 1000.0, -3.95])
    assert n["subtended"] < 3 * math.pi / 180
    assert n["along"].size == 1
    assert abs(n["weight"][0] - n["subtended"]) < 1e-15 and abs(n["along"][0] - 10.0) < 1e-9


@pytest.mark.parametrize("nearest", [10.0, 0.5])
def test_wall_places_blocked_and_clear_nodes(nearest):
    wall = [math.radians(80.0), math.radians(100.0), nearest]
    n = oracle.line_nodes([-125.0, 20.0, 0.0], [125.0, 20.0, 0.0], arcs=[wall])
    assert n["along"].size > 5
    assert abs(n["weight"].sum() - n["subtended"]) < 1e-12
    az = np.degrees(np.arctan2(20.0, n["along"] - 125.0))
    clear = ~n["obstacles"]
    assert clear.any()
    assert not np.any((az[clear] >= 80.0) & (az[clear] <= 100.0))
    assert np.any(n["obstacles"] & (az >= 80.0) & (az <= 100.0))


def test_skyline_against_every_edge():
    """Every edge the skyline hands over is a real edge within the radius, and every edge within the
    radius whose short arc meets the wedge (under a half-turn, as the quadrature asks) is handed over."""
    from quietmap_amd import scenes

    sc = scenes.road_scene()
    O = sc.scene.obstacles
    rng = np.random.default_rng(3)
    verts = O.vertices.reshape(-1, 2).astype(np.float64)
    mps = O.metres_per_step
    for _ in range(6):
        x, y = rng.uniform(300, 2200, 2)
        lo = rng.uniform(-math.pi, math.pi)
        hi = lo + rng.uniform(0.2, 2.5)
        radius = 250.0
        got = {round(float(d), 9) for d in oracle.skyline_arcs(O, x, y, lo, hi, radius, 0.0)[:, 2]}
        ox = O.origin_steps[0] + x / mps
        oy = O.origin_steps[1] - y / mps
        within, meeting = set(), set()
        for ol in O.outlines:
            v = verts[ol["first_vertex"]:ol["first_vertex"] + ol["vertex_count"]]
            px, py = (v[:, 0] - ox) * mps, (oy - v[:, 1]) * mps
            for k in range(len(v) - 1):
                a, b = np.array([px[k], py[k]]), np.array([px[k + 1], py[k + 1]])
                e = b - a
                t = np.clip(-(a @ e) / (e @ e), 0, 1) if e @ e > 0 else 0.0
                d = round(float(np.hypot(*(a + t * e))), 9)
                if not 1e-6 <= d <= radius:
                    continue
                within.add(d)
                za, zb = math.atan2(a[1], a[0]), math.atan2(b[1], b[0])
                ub = za + (zb - za + math.pi) % (2 * math.pi) - math.pi
                alo, ahi = min(za, ub), max(za, ub)
                if any(alo + s_ <= hi and ahi + s_ >= lo for s_ in (0.0, 2 * math.pi, -2 * math.pi)):
                    meeting.add(d)
        assert got <= within
        assert meeting <= got
