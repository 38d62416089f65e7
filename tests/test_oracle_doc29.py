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


SECONDS = np.array([12.0, 4.0, 8.0]) * 3600.0


def heard_lengths(sc, x, y):
    """Per track, the summed 3-D length of its segments within reach of (x, y), by brute force."""
    v = sc.vertices.reshape(-1, 3).astype(np.float64)
    out = np.zeros(sc.tracks.size)
    for t, tr in enumerate(sc.tracks):
        p = v[tr["first_vertex"]:tr["first_vertex"] + tr["vertex_count"]]
        a, b = p[:-1], p[1:]
        s = a[:, :2] - [x, y]
        d = b[:, :2] - a[:, :2]
        ll = np.sum(d * d, 1)
        u = np.clip(-np.sum(s * d, 1) / np.where(ll > 0, ll, 1.0), 0.0, 1.0) * (ll > 0)
        q = s + u[:, None] * d
        heard = np.sum(q * q, 1) <= sc.reach_m ** 2
        out[t] = np.sqrt(np.sum((b - a) ** 2, 1))[heard].sum()
    return out


def test_pass_classes_hold_the_energy_and_the_time_in_progress():
    """The classes' energies sum to the period's SEL energy over its seconds, and their passing
    shares to every heard segment's flights times its duration at the track's ground speed
    (popup/src/aircraft/boxes.rs:225-265 with the segment as the chord)."""
    sc = scenes.aircraft_scene(scale=0.02)
    pts = sc.points_xy[::37][:12]
    r = oracle.aircraft_points(sc, pts, passes=True)
    energy, passing = r["passes"][:, 0], r["passes"][:, 1]
    want = r["period_energy"] / SECONDS
    assert np.allclose(energy.sum(axis=2), want, rtol=1e-12, atol=0)
    v_ref = sc.classes["v_ref_kt"][sc.tracks["aircraft_class"]]
    speed = np.where(sc.tracks["speed_kt"] > 10.0, sc.tracks["speed_kt"].astype(np.float64), v_ref)
    spm = 3600.0 / (speed * 1852.0)
    flights = sc.tracks["flights"].astype(np.float64)
    for i, (x, y) in enumerate(pts):
        lengths = heard_lengths(sc, x, y)
        expected = (flights * (lengths * spm)[:, None]).sum(axis=0) / SECONDS
        assert np.allclose(passing[i].sum(axis=1), expected, rtol=1e-9, atol=0), i
    assert (passing > 0).sum() > 50


def test_a_pass_lands_in_the_class_of_its_level():
    """One level segment: its pass sits in the class of 10 lg(energy / passing) (3 dB classes from
    -30 dB, aircraft.rs:86-95), the same class in every period with flights, clamped at both ends."""
    sc = scenes.aircraft_scene(scale=0.01)
    tr = sc.tracks[:1].copy()
    tr["first_vertex"], tr["vertex_count"] = 0, 2
    tr["flights"] = [3, 0, 2]
    cases = ((300.0, 2000.0, 0.0), (3000.0, 800.0, 0.0), (60.0, 20000.0, 0.0), (9000.0, 5.0, 0.0),
             (30000.0, 1000.0, 0.0), (5.0, -999.9, -1000.0))
    hit = set()
    for height, x1, rx in cases:
        verts = np.array([[-1000.0, 50.0, height], [x1, 50.0, height]], np.float32)
        one = scenes.AircraftScene(np.array([[rx, 50.0]]), (1, 1), sc.classes, tr, verts, 0.0, 1e9)
        r = oracle.aircraft_points(one, one.points_xy, passes=True)
        rel = verts.astype(np.float64) - [rx, 50.0, 0.0]
        sel, _ = oracle.segment_levels(sc.classes, tr[0], rel[0], rel[1])
        v = float(sc.classes[tr[0]["aircraft_class"]]["v_ref_kt"])
        speed = float(tr[0]["speed_kt"]) if tr[0]["speed_kt"] > 10 else v
        duration = float(np.linalg.norm(verts[1].astype(np.float64) - verts[0])) * 3600.0 / (speed * 1852.0)
        level = sel - 10 * math.log10(duration)
        k = min(max(int(math.floor((level + 30.0) / 3.0)), 0), abi.PASS_CLASSES - 1)
        hit.add(k)
        for p, f in enumerate(tr[0]["flights"]):
            row = r["passes"][0, 1, p]
            if f == 0:
                assert not row.any()
                continue
            assert np.flatnonzero(row).tolist() == [k], (height, p)
            assert abs(r["passes"][0, 0, p, k] - f * 10 ** (sel / 10) / SECONDS[p]) <= 1e-9 * r["passes"][0, 0, p, k]
            assert abs(row[k] - f * duration / SECONDS[p]) <= 1e-12 * row[k]
    assert {0, abi.PASS_CLASSES - 1} <= hit


def test_screening_loss_and_its_composition():
    """screening_tests.rs: zero at grazing, capped at 18 dB; an edge screens only between receiver
    and source above the line of sight; smooth across the shadow boundary."""
    assert oracle.path_difference_loss(0.0) == 0.0 and oracle.path_difference_loss(-1.0) == 0.0
    six = 10.0 * math.log10((3.0 + 29.2 * 6.0) / 3.0)
    assert abs(oracle.path_difference_loss(6.0) - six) < 1e-12
    assert oracle.path_difference_loss(6.5) == 18.0
    assert oracle.edge_loss(500.0, 500.0, 3000.0, 0.0) == 18.0
    assert oracle.edge_loss(1000.0, 100.0, 2000.0, 150.0) > 5.0
    for args in ((1000.0, 100.0, 800.0, 20.0), (1000.0, 100.0, 1000.0, 20.0), (1000.0, 100.0, 2000.0, 250.0),
                 (1000.0, 100.0, 2000.0, 200.0), (1000.0, 100.0, 0.0, 500.0)):
        assert oracle.edge_loss(*args) == 0.0
    losses = [oracle.edge_loss(158.0, 20.0, 5000.0, 5000.0 * math.tan(math.radians(6.7 + 0.1 * k))) for k in range(11)]
    assert any(x > 0 for x in losses) and 0.0 in losses
    assert all(abs(b - a) < 0.5 for a, b in zip(losses, losses[1:]))


def test_horizons_screen_what_lies_behind_their_edges():
    """horizons_tests.rs: sectors by direction; a ridge 40 m high 1 km east (the steeper of two edges
    in its band) screens a low aircraft 5 km east, not one west, nor one before the ridge, nor one
    above its line of sight."""
    assert oracle.sector_of(1.0, 0.001, 32) == 0 and oracle.sector_of(0.0, 1.0, 32) == 8
    assert oracle.sector_of(-1.0, -0.001, 32) == 16
    assert oracle.sector_of(0.001, -1.0, 256) == 192 and oracle.sector_of(1.0, -0.001, 256) == 255
    rec = np.full(abi.TERRAIN_HORIZON_FLOATS, np.nan, np.float32)
    rec[0] = 0.04
    for sector in range(abi.TERRAIN_SECTORS):
        rec[1 + sector * 13] = -np.inf
    rec[1] = 0.04
    rec[1 + 1 + 2 * 1], rec[1 + 2 + 2 * 1] = 1000.0, 0.04
    behind = oracle.horizon_loss(rec, 32, 5000.0, 10.0, 150.0)
    expected = oracle.edge_loss(1000.0, float(np.float32(0.04)) * 1000.0, math.hypot(5000.0, 10.0), 150.0)
    assert behind > 0.0 and abs(behind - expected) < 1e-9
    assert oracle.horizon_loss(rec, 32, -5000.0, -10.0, 150.0) == 0.0
    assert oracle.horizon_loss(rec, 32, 800.0, 1.0, 10.0) == 0.0
    assert oracle.horizon_loss(rec, 32, 5000.0, 10.0, 400.0) == 0.0


def test_screening_lowers_only_the_sel():
    """With horizons, the SEL energy of each receiver drops or stays (the larger of the losses less
    the lateral attenuation already taken), LAmax, N65/N70 and the histogram do not change, and a
    receiver with no horizon reads the unscreened values."""
    sc = scenes.aircraft_scene(scale=0.02)
    scr = scenes.aircraft_horizons(sc)
    pts = np.arange(0, sc.points_xy.shape[0], 7)
    sub = dict(scr, horizon_of=scr["horizon_of"][pts])
    free = oracle.aircraft_points(sc, sc.points_xy[pts])
    screened = oracle.aircraft_points(sc, sc.points_xy[pts], screening=sub)
    assert np.all(screened["period_energy"] <= free["period_energy"] * (1 + 1e-12))
    assert np.any(screened["period_energy"] < free["period_energy"] * 0.999)
    for k in ("n65", "n70", "hist"):
        assert np.array_equal(screened[k], free[k])
    none = sub["horizon_of"] == abi.NO_HORIZON
    assert none.any()
    assert np.array_equal(screened["period_energy"][none], free["period_energy"][none])
