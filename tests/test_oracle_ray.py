"""The oracle's ray (physics/src/ray.rs) pinned by the reference's own ray tests
(physics/src/ray_tests.rs), the mean-plane closed forms (physics/src/cnossos/mean_plane.rs:135-190)
and invariants the physics fixes (hard ground in calm air, free field, energy mixing)."""
import math

import numpy as np
import pytest

from oracle import oracle

ALPHA = oracle.alpha_bands(15.0, 70.0)


def profile():
    """ray_tests.rs:15-25: 100 m, terrain rising to 105 m at x = 30 m."""
    return dict(horizontal_m=100.0, t=[0.0, 0.25, 0.3, 0.5, 0.75, 1.0],
                ground_m=[100.0, 102.0, 105.0, 101.0, 103.0, 100.0],
                ground_factor=[1.0, 0.5, 0.4, 0.0, 0.5, 1.0])


def building(t, h, f):
    return (t, h, True, True, f)


def barrier(t, h, f):
    return (t, h, False, False, f)


def test_platform_clamps_near_terrain():
    """ray_tests.rs:58-80."""
    path = oracle.vertical_path(profile(), [], 4.0, platform_half_width_m=30.0, own_footprint=1)
    np.testing.assert_allclose(path["d"], [0.0, 25.0, 30.0, 50.0, 75.0, 100.0], atol=1e-9)
    np.testing.assert_allclose(path["z"], [100.0, 100.0, 105.0, 101.0, 103.0, 100.0], atol=1e-9)
    np.testing.assert_allclose(path["g"], [1.0, 0.5, 0.4, 0.0, 0.5, 1.0], atol=1e-9)
    assert len(path["crests"]) == 6 and len(path["tops"]) == 0


def test_a_source_on_a_deck_stands_on_it_above_the_valley():
    """ray_tests.rs:82-104: a deck at 150 m leaves the valley in the profile (the clamp takes the
    lower of the two) and the ray runs from the deck plus the source's height."""
    path = oracle.vertical_path(profile(), [], 4.0, platform_half_width_m=30.0, own_footprint=1, source_ground_m=150.0)
    np.testing.assert_allclose(path["z"], [100.0, 102.0, 105.0, 101.0, 103.0, 100.0], atol=1e-9)
    ends = dict(source_height_m=4.0, receiver_altitude_m=104.0, source_ground_m=150.0, platform_half_width_m=30.0,
                own_footprint=1)
    terms = oracle.ray_terms(profile(), [], ends, [0.0] * 3, ALPHA)
    assert abs(terms["slant_m"] - math.hypot(100.0, 50.0)) < 1e-9
    # A deck below the terrain cuts the platform down to it.
    low = oracle.vertical_path(profile(), [], 4.0, platform_half_width_m=30.0, own_footprint=1, source_ground_m=99.0)
    np.testing.assert_allclose(low["z"], [99.0, 99.0, 105.0, 101.0, 103.0, 100.0], atol=1e-9)


def test_own_footprint_skipped_and_tops_on_interpolated_terrain():
    """ray_tests.rs:107-134."""
    crossings = [building(0.05, 6.0, 1), building(0.1, 5.0, 3), building(0.12, 5.0, 3), barrier(0.6, 3.0, 2),
                 building(0.9, 6.0, 1), barrier(1.0, 4.0, 4)]
    path = oracle.vertical_path(profile(), crossings, 4.0, platform_half_width_m=30.0, own_footprint=1)
    np.testing.assert_allclose(path["tops"], [(10.0, 105.0), (12.0, 105.0), (60.0, 104.8), (100.0, 104.0)],
                               atol=1e-9)


def test_footprints_containing_a_buildings_source_do_not_screen_it():
    """ray_tests.rs:139-165."""
    crossings = [building(0.02, 20.0, 5), building(0.05, 20.0, 1), building(0.4, 8.0, 6), building(0.45, 8.0, 6)]
    path = oracle.vertical_path(profile(), crossings, 4.0, platform_half_width_m=30.0, own_footprint=1)
    np.testing.assert_allclose(path["tops"][:, 0], [40.0, 45.0], atol=1e-9)
    road = oracle.vertical_path(profile(), crossings, 4.0, platform_half_width_m=30.0, own_footprint=0)
    assert len(road["tops"]) == 4


def test_overlapping_roofs_clip_in_closing_order():
    """ray_tests.rs:169-210."""
    p = dict(horizontal_m=100.0, t=[0.0, 0.5, 1.0], ground_m=[100.0, 100.0, 100.0], ground_factor=[1.0, 0.5, 0.0])
    crossings = [building(0.2, 6.0, 11), building(0.4, 6.0, 11), building(0.3, 8.0, 12), building(0.5, 10.0, 12),
                 building(0.25, 7.0, 13), building(0.4, 7.0, 13)]
    path = oracle.vertical_path(p, crossings, 4.0, platform_half_width_m=30.0, own_footprint=0)
    np.testing.assert_allclose(path["d"], [0.0, 20.0, 20.0, 40.0, 40.0, 40.0, 40.0, 50.0, 50.0, 100.0], atol=1e-9)
    np.testing.assert_allclose(path["z"], [100.0, 100.0, 106.0, 106.0, 100.0, 100.0, 109.0, 110.0, 100.0, 100.0],
                               atol=1e-9)
    np.testing.assert_allclose(path["g"], [1.0, 0.8, 0.0, 0.0, 0.6, 0.6, 0.0, 0.0, 0.5, 0.0], atol=1e-9)


def flat(horizontal_m):
    """ray_tests.rs:213-221."""
    return dict(horizontal_m=horizontal_m, t=[s / 10 for s in range(11)], ground_m=[200.0] * 11,
                ground_factor=[0.5] * 11)


ENDS = dict(source_height_m=1.0, receiver_altitude_m=204.0, source_ground_factor=0.5)


def test_flat_ground_transfer_constants():
    """ray_tests.rs:223-276: 1 kHz calm dips to 0.9097167331685742, favourable 1.4125375446227544."""
    p = flat(110.54)
    homogeneous = oracle.ray_terms(p, [], ENDS, [0.0] * 3, ALPHA)
    favourable = oracle.ray_terms(p, [], ENDS, [1.0] * 3, ALPHA)
    mixed = oracle.ray_terms(p, [], ENDS, [0.5, 0.0, 1.0], ALPHA)
    slant = math.hypot(110.54, 3.0)
    assert abs(homogeneous["slant_m"] - slant) < 1e-12
    air = lambda b: 10 ** (-ALPHA[b] * slant / 1000.0 / 10.0)
    assert abs(homogeneous["periods"][0][4] / air(4) - 0.9097167331685742) < 1e-9
    assert abs(favourable["periods"][0][4] / air(4) - 1.4125375446227544) < 1e-9
    assert abs(homogeneous["periods"][0][0] / air(0) - 1.4125375446227544) < 1e-9
    for b in range(8):
        expected = 0.5 * homogeneous["periods"][0][b] + 0.5 * favourable["periods"][0][b]
        assert abs(mixed["periods"][0][b] - expected) < 1e-12
        assert mixed["periods"][1][b] == homogeneous["periods"][1][b]
        assert mixed["periods"][2][b] == favourable["periods"][2][b]


def test_forest_takes_table_a1_off_each_state():
    """ray_tests.rs:280-313 and physics/src/foliage.rs Table A.1."""
    p = flat(110.54)
    open_ = oracle.ray_terms(p, [], ENDS, [0.5] * 3, ALPHA)["states"]
    wooded_profile = dict(p, forest_cover=[1.0] * 11, canopy_m=[20.0] * 11)
    wooded = oracle.ray_terms(wooded_profile, [], ENDS, [0.5] * 3, ALPHA)["states"]
    depth = math.hypot(110.54, 3.0)
    table = [0.02, 0.03, 0.04, 0.05, 0.06, 0.08, 0.09, 0.12]
    for state, tolerance in ((0, 1e-9), (1, 0.01)):
        for b in range(8):
            taken = 10 * math.log10(open_[state][b] / wooded[state][b])
            assert abs(taken - table[b] * depth) < tolerance


def test_a_wall_halfway_screens_over_10_db():
    """ray_tests.rs:317-352."""
    p = dict(flat(200.0), ground_factor=[0.0] * 11)
    ends = dict(source_height_m=0.05, receiver_altitude_m=204.0, source_ground_factor=0.0)
    open_ = oracle.ray_terms(p, [], ends, [0.5] * 3, ALPHA)
    walled = oracle.ray_terms(p, [barrier(0.5, 6.0, 9)], ends, [0.5] * 3, ALPHA)
    assert 10 * math.log10(open_["periods"][0][4] / walled["periods"][0][4]) > 10.0


def test_crests_thin_to_the_highest_of_each_150_m():
    """ray_tests.rs:356-389: 39 middle samples keep (350, 5), (510, 5), (670, 5)."""
    t = [k / 100 for k in range(101)]
    z = [5.0 if k % 8 == 3 else 0.0 for k in range(101)]
    p = dict(horizontal_m=1000.0, t=t, ground_m=z, ground_factor=[0.0] * 101)
    crests = oracle.vertical_path(p, [], 1.0)["crests"]
    middle = [tuple(c) for c in crests if 300.0 < c[0] < 700.0]
    np.testing.assert_allclose(middle, [(350.0, 5.0), (510.0, 5.0), (670.0, 5.0)], atol=1e-9)
    ends = [tuple(c) for c in crests if not (300.0 < c[0] < 700.0)]
    assert len(ends) == 101 - 39


def test_hard_ground_calm_air_is_minus_3_db():
    """physics/src/cnossos/ground.rs:116-119: G'path below 1e-3 gives -3 dB in every band."""
    d = [0.0, 50.0, 100.0]
    att, _ = oracle.state_boundaries(d, [0.0] * 3, [0.0] * 3, (0.0, 1.0), (100.0, 4.0), 0.0, [(50.0, 0.0)], [])
    np.testing.assert_allclose(att[0], [-3.0] * 8, atol=1e-12)


@pytest.mark.parametrize("start,end", [(0.0, 100.0), (1.5, 70.0), (40.0, 100.0)])
def test_a_ramp_under_a_high_path_has_no_diffraction(start, end):
    """A straight ramp is its own mean plane (mean_plane.rs:135-153): high ends over a ramp give the
    ground term alone, equal for any vertex split of the ramp."""
    x1 = [0.0, 3.0, 40.0, 100.0]
    z1 = [2.0 + 0.1 * x for x in x1]
    x2 = [0.0, 100.0]
    z2 = [2.0, 12.0]
    a1, _ = oracle.state_boundaries(x1, z1, [0.5] * 4, (0.0, 30.0), (100.0, 40.0), 0.5, [(3.0, 2.3), (40.0, 6.0)], [])
    a2, _ = oracle.state_boundaries(x2, z2, [0.5] * 2, (0.0, 30.0), (100.0, 40.0), 0.5, [(3.0, 2.3), (40.0, 6.0)], [])
    np.testing.assert_allclose(a1, a2, atol=1e-9)
