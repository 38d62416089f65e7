"""The oracle's geographic frame pinned independently of its formulas: the frame's metres per
second of arc against the WGS84 meridian and parallel arcs, every walk vertex on a row line of
the table or a column line (and every such line crossed once), a table of even rows equal to the
plain lattice, and the weather read on the world table by the place's latitude and longitude
(popup/src/scene.rs:303-420, tiles/src/geo.rs:224-251, physics/src/weather.rs:127-200)."""
import math

import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi, scenes


@pytest.fixture(scope="module")
def geo():
    return scenes.geo_scene()


def test_frame_against_the_ellipsoid(geo):
    """Near the origin a row is one second of the meridian arc, M(phi) pi / 648000, and a column one
    second of the parallel, N(phi) cos(phi) pi / 648000 (WGS84 radii)."""
    T = geo.scene.terrain
    ox, oy, east, north = geo.frame
    lat, _ = scenes.mercator_degrees(ox, oy)
    phi = math.radians(float(lat))
    w = 1.0 - scenes.WGS84_E2 * math.sin(phi) ** 2
    meridian = scenes.WGS84_A_M * (1.0 - scenes.WGS84_E2) / w ** 1.5 * math.pi / 648000.0
    parallel = scenes.WGS84_A_M / math.sqrt(w) * math.cos(phi) * math.pi / 648000.0
    assert abs(T.column_m / parallel - 1.0) < 1e-9
    k = int(np.argmin(np.abs(T.row_y_m)))  # the row at the origin
    assert abs((T.row_y_m[k + 1] - T.row_y_m[k]) / meridian - 1.0) < 2e-4
    # Rows are uneven: their spacing grows northward (Mercator stretches with latitude).
    d = np.diff(T.row_y_m)
    assert np.all(np.diff(d) > 0)


def lines_between(values, a, b):
    lo, hi = min(a, b), max(a, b)
    return np.flatnonzero((values > lo) & (values < hi))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_walk_vertices_lie_on_the_table(geo, seed):
    """Every vertex between the ends lies on a row line (north = row_y_m[k]) or a column line, and
    every line strictly between the ends gives one vertex (a node crossing once)."""
    T = geo.scene.terrain
    rng = np.random.default_rng(seed)
    columns = T.origin_x_m + np.arange(T.nodes.shape[1]) * T.column_m
    for _ in range(20):
        a = rng.uniform(0, 640, 2)
        b = rng.uniform(0, 640, 2)
        failed, prof = oracle.fill_profile(T, a, b)
        assert failed == 0
        t = prof["t"][1:-1]
        x = a[0] + t * (b[0] - a[0])
        y = a[1] + t * (b[1] - a[1])
        on_row = np.min(np.abs(y[:, None] - T.row_y_m[None, :]), axis=1) < 1e-6
        on_column = np.min(np.abs(x[:, None] - columns[None, :]), axis=1) < 1e-6
        assert np.all(on_row | on_column)
        rows_crossed = lines_between(T.row_y_m, a[1], b[1]).size
        cols_crossed = lines_between(columns, a[0], b[0]).size
        assert t.size == rows_crossed + cols_crossed - int(np.sum(on_row & on_column))
        assert np.all(np.diff(prof["t"]) >= 0)


def test_even_table_equals_the_plain_lattice():
    """A table of evenly spaced rows and columns at the spacing paints as the plain lattice does."""
    sc = scenes.road_scene(seed=3, receivers=32, lines=10, buildings=60)
    T = sc.scene.terrain
    rows = T.nodes.shape[0]
    table = abi.Terrain(T.origin_x_m, T.origin_y_m, T.spacing_m, T.nodes, T.canopy_m, column_m=T.spacing_m,
                        row_y_m=T.origin_y_m + np.arange(rows) * T.spacing_m)
    even = abi.Scene(table, sc.scene.obstacles, sc.scene.weather)
    rng = np.random.default_rng(5)
    pp = rng.integers(0, sc.points.size, 400).astype(np.uint32)
    ps = rng.integers(0, sc.sources.size, 400).astype(np.uint32)
    plain = oracle.evaluate_pairs(sc.scene, sc.points, sc.sources, pp, ps, point_bands=False)
    got = oracle.evaluate_pairs(even, sc.points, sc.sources, pp, ps, point_bands=False)
    assert np.array_equal(plain["failed"], got["failed"])
    np.testing.assert_allclose(got["sums"], plain["sums"], rtol=1e-9, atol=0)


def corners_by_hand(nodes, npd, lat, lon):
    """weather.rs corners and place_of restated for one place."""
    rows, cols = 180 * npd + 1, 360 * npd
    y = min(max((90.0 - lat) * npd, 0.0), rows - 1)
    x = (lon % 360.0) * npd
    row, col = min(int(math.floor(y)), rows - 2), int(math.floor(x))
    fy, fx = y - row, x - col
    idx = [(row, col % cols), (row, (col + 1) % cols), (row + 1, col % cols), (row + 1, (col + 1) % cols)]
    w = [(1 - fy) * (1 - fx), (1 - fy) * fx, fy * (1 - fx), fy * fx]
    fav = sum(wk * nodes[r, c, :48].astype(float) for wk, (r, c) in zip(w, idx)) / 100.0
    alpha = sum(wk * nodes[r, c, 48:].view("<u2").astype(float) for wk, (r, c) in zip(w, idx)) / 100.0
    return fav.reshape(3, 16), alpha


@pytest.mark.parametrize("lat,lon", [(50.0875, 14.4214), (51.4779, -0.0004), (-33.86, 151.21)])
def test_weather_by_latitude_and_longitude(lat, lon):
    sc = scenes.geo_scene(lat_deg=lat, lon_deg=lon, receivers=16, buildings=10, lines=2)
    W = sc.scene.weather
    ox, oy, east, north = sc.frame
    pts = sc.points[::37]
    got = oracle.receivers(sc.scene, pts)
    for i, p in enumerate(pts):
        plat, plon = scenes.mercator_degrees(ox + p["x_m"] / east, oy - p["y_m"] / north)
        fav, alpha = corners_by_hand(W.nodes, W.nodes_per_degree, float(plat), float(plon))
        np.testing.assert_allclose(got["favourable"][i], fav, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(got["alpha"][i], alpha, rtol=1e-12, atol=1e-12)
