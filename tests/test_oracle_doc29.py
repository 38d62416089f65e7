"""The oracle's Doc 29 segment path against the reference's own pins: the corrections
(physics/src/doc29/corrections_tests.rs), the closest points of an inclined path
(segment_tests.rs:84-105), collinear pieces carrying their segment's energy (segment_tests.rs:262-279),
and the NPD levels under a long overhead segment (segment_tests.rs:283-300)."""
import math

import numpy as np

from oracle import oracle
from quietmap_amd import abi, scenes


def test_lateral_attenuation_matches_doc29():
    for h, l, want in ((100.0, 1000.0, 5.326), (100.0, 8000.0, 9.901), (-50.0, 200.0, 4.988)):
        assert abs(oracle.lateral_attenuation(h, l) - want) < 0.001
    assert oracle.lateral_attenuation(1000.0, 0.0) == 0.0
    for l in (0.0, 5.0, 100.0, 913.0, 915.0, 16000.0):
        for h in (-3000.0, -1.0, 0.0, 1.0, 50.0, 800.0, 12000.0):
            assert oracle.lateral_attenuation(h, l) >= 0.0


def test_finite_segment_fraction():
    alongside = oracle.finite_fraction(500.0, 1000.0, 370.0)
    assert 0.1 < alongside < 1.0
    assert oracle.finite_fraction(-500.0, 1000.0, 370.0) < 10 ** -0.5
    assert abs(oracle.finite_fraction(500_000.0, 1_000_000.0, 370.0) - 1.0) < 1e-3
    assert abs(oracle.finite_fraction(0.0, 1_000_000.0, 370.0) - 0.5) < 1e-3


def test_installation_peak_and_overhead():
    """corrections.rs:8-10: wing-mounted peaks at 0.401395 dB near 50.43 deg; 0 dB overhead."""
    best = max(10 * math.log10(oracle.installation_gain(abi.INSTALL_WING, math.sin(math.radians(a / 100)), 1.0))
               for a in range(4500, 5600))
    assert abs(best - 0.401395) < 1e-4
    assert abs(oracle.installation_gain(abi.INSTALL_WING, 100.0, 100.0) - 1.0) < 1e-12
    assert oracle.installation_gain(abi.INSTALL_PROPELLER, 10.0, 100.0) == 1.0


def test_closest_point_is_the_perpendicular_to_the_inclined_path():
    climb = math.radians(27.0)
    lateral, up = 62.2, 90.7
    at = lambda x: [x, lateral, up + x * math.tan(climb)]
    c = oracle.closest_points(at(-500.0), at(500.0))
    expected = math.hypot(lateral, up * math.cos(climb))
    assert abs(np.linalg.norm(c["on_line"]) - expected) < 1e-6
    assert abs(c["lateral_m"] - lateral) < 1e-9
    assert abs(c["height_m"] - up * math.cos(climb)) < 1e-6
    assert c["lambda_height_m"] == c["height_m"]
    assert abs(c["length_m"] - 1000.0 / math.cos(climb)) < 1e-6
    b = oracle.closest_points(at(500.0), at(1500.0))
    assert b["along"] < 0.0
    assert abs(b["lambda_height_m"] - at(500.0)[2] / math.cos(climb)) < 1e-6


def flat_class():
    sc = scenes.aircraft_scene(scale=0.01)
    cl = sc.classes.copy()
    return cl, sc


def test_collinear_pieces_carry_the_energy_of_their_segment():
    cl, sc = flat_class()
    track = sc.tracks[0].copy()
    e = lambda a, b: 10 ** (oracle.segment_levels(cl, track, [a, 8000.0, 8000.0], [b, 8000.0, 8000.0])[0] / 10)
    whole = e(-10_000.0, 10_000.0)
    pieces = sum(e(-10_000.0 + 2000.0 * k, -8000.0 + 2000.0 * k) for k in range(10))
    assert abs(pieces / whole - 1.0) < 1e-9


def test_npd_levels_under_a_long_overhead_segment():
    """Overhead a long level segment Lambda, Delta_I and Delta_F vanish: the SEL is the NPD row's
    own level plus Delta_V at every NPD distance (segment_tests.rs:283-300)."""
    cl, sc = flat_class()
    track = sc.tracks[0].copy()
    track["power_weight"] = 0.0
    k = int(track["aircraft_class"])
    row = int(track["power_row"])
    dep = int(track["departure"])
    dv = 10 * math.log10(cl[k]["v_ref_kt"] / float(track["speed_kt"])) if track["speed_kt"] > 10 else 0.0
    for i, ft in enumerate([200.0, 400.0, 630.0, 1000.0, 2000.0, 4000.0, 6300.0, 10000.0, 16000.0, 25000.0]):
        h = ft * 0.3048
        sel, _ = oracle.segment_levels(cl, track, [-5e5, 0.0, h], [5e5, 0.0, h])
        assert abs(sel - (cl[k]["sel_db"][dep, row, i] + dv)) < 1e-3, i


def test_histogram_counts_flights():
    sc = scenes.aircraft_scene(scale=0.03)
    r = oracle.aircraft_points(sc, sc.points_xy[:50])
    flights = int(sc.tracks["flights"].sum())
    assert np.all(r["hist"].sum(axis=1) <= flights)
    assert np.all(r["n70"] <= r["n65"])
    assert r["segments"].sum() > 0 and np.all(r["period_energy"] >= 0)
