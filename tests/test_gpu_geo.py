"""The geographic frame on the card against the FP64 oracle (popup/src/scene.rs's one arc-second
lattice through Mercator rows, physics/src/weather.rs's world table by latitude and longitude,
tiles/src/geo.rs's frame with its east and north metres per unit): receivers' ground and weather,
point pairs, the grid painter, line pairs and fine terrain over the table, Lden within 0.05 dB,
failures and paths alike."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import scenes

pytestmark = pytest.mark.gpu

RECEIVER_DTYPE = np.dtype([("x_m", "<f8"), ("y_m", "<f8"), ("own_footprint", "<u8"), ("altitude_m", "<f4"),
                           ("failed", "<u4"), ("alpha", "<f4", (8,)), ("favourable", "<f4", (3, 16))])


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


@pytest.fixture(scope="module")
def geo():
    return scenes.geo_scene()


def db_error(got, want):
    assert np.array_equal(want > 0, got > 0), "silent on one side only"
    ok = want > 0
    return np.abs(10 * np.log10(got[ok] / want[ok]))


def test_geo_receivers(gpu, geo):
    from quietmap_amd import hip
    from quietmap_amd.painter import RECEIVER_BYTES

    print("case: geo receivers", flush=True)
    assert RECEIVER_DTYPE.itemsize <= RECEIVER_BYTES
    square = gpu.upload(geo.scene, geo.sources)
    pts = geo.points
    d_pts = hip.DeviceBuffer.from_array(pts)
    d_rcv = hip.DeviceBuffer(pts.size * RECEIVER_BYTES)
    gpu.receivers_device(square, d_pts.ptr, pts.size, d_rcv.ptr)
    hip.synchronize()
    raw = d_rcv.to_array(np.uint8, pts.size * RECEIVER_BYTES).reshape(pts.size, RECEIVER_BYTES)
    got = np.frombuffer(raw[:, :RECEIVER_DTYPE.itemsize].tobytes(), RECEIVER_DTYPE)
    want = oracle.receivers(geo.scene, pts)
    assert np.array_equal(got["failed"], want["failed"])
    dz = np.abs(got["altitude_m"] - want["altitude_m"]).max()
    dp = np.abs(got["favourable"] - want["favourable"]).max()
    da = np.abs(got["alpha"] - want["alpha"]).max()
    print(f"receivers {pts.size}: altitude {dz:.2e} m, favourable {dp:.2e}, alpha {da:.2e} dB/km", flush=True)
    assert dz < 1e-3 and dp < 1e-5 and da < 1e-4
    square.close()


def test_geo_pairs_and_grid(gpu, geo):
    from quietmap_amd.painter import lden_db

    print("case: geo pairs and grid", flush=True)
    square = gpu.upload(geo.scene, geo.sources)
    rng = np.random.default_rng(31)
    n = 3000
    pp = rng.integers(0, geo.points.size, n).astype(np.uint32)
    ps = rng.integers(0, geo.sources.size, n).astype(np.uint32)
    got = gpu.evaluate_pairs(square, geo.points, pp, ps)
    want = oracle.evaluate_pairs(geo.scene, geo.points, geo.sources, pp, ps, point_bands=False)
    assert np.array_equal(got["failed"] != 0, want["failed"] != 0)
    ok = got["failed"] == 0
    assert ok.mean() > 0.99
    err = db_error(got["sums"][ok].astype(np.float64), want["sums"][ok])
    print(f"geo pairs: n={err.size} max {err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    nx, ny = geo.grid_shape
    g = gpu.paint_grid(square, geo.points, nx, ny, 1e9)
    sample = rng.choice(geo.points.size, 256, replace=False)
    ref = oracle.paint_points(geo.scene, geo.points[sample], geo.sources, 1e9)
    assert np.array_equal(g["paths"][sample].astype(np.int64), ref["paths"].astype(np.int64))
    got_l, want_l = lden_db(g["period_energy"][sample]), lden_db(ref["period_energy"])
    assert np.array_equal(np.isfinite(got_l), np.isfinite(want_l))
    e = np.abs(got_l - want_l)
    e = e[np.isfinite(e)]
    print(f"geo grid Lden: n={e.size} max {e.max():.5f} dB", flush=True)
    assert e.max() <= 0.05
    square.close()


def test_geo_lines_and_fine(gpu, geo):
    from quietmap_amd import abi

    print("case: geo lines and fine terrain", flush=True)
    rng = np.random.default_rng(32)
    pieces = scenes.line_pieces(rng, 10, (640.0, 640.0), (0.0, 0.0))
    square = gpu.upload(geo.scene, geo.sources)
    n = 1500
    pp = rng.integers(0, geo.points.size, n).astype(np.uint32)
    pl = rng.integers(0, pieces.size, n).astype(np.uint32)
    got = gpu.evaluate_line_pairs(square, geo.points, pieces, pp, pl)
    want = oracle.evaluate_line_pairs(geo.scene, geo.points, pieces, pp, pl)
    assert np.array_equal(got["failed"] != 0, want["failed"] != 0)
    ok = got["failed"] == 0
    assert np.array_equal(got["rays"][ok], want["rays"][ok])
    err = db_error(got["sums"][ok].astype(np.float64), want["sums"][ok])
    print(f"geo line pairs: n={err.size} max {err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    square.close()
    fine_t = scenes.with_fine_terrain(geo.scene.terrain, share=0.6, amplitude_m=5.0)
    fine = abi.Scene(fine_t, geo.scene.obstacles, geo.scene.weather)
    square = gpu.upload(fine, geo.sources)
    ps = rng.integers(0, geo.sources.size, n).astype(np.uint32)
    got = gpu.evaluate_pairs(square, geo.points, pp, ps)
    want = oracle.evaluate_pairs(fine, geo.points, geo.sources, pp, ps, point_bands=False)
    # The card may flag a path whose hull outgrows its buffer (capacity), never otherwise.
    bad = got["failed"] != 0
    assert not (want["failed"] != 0).any()
    assert np.all((got["failed"][bad] & ~np.uint32(4 | 64)) == 0) and bad.mean() < 1e-2
    ok = ~bad
    err = db_error(got["sums"][ok].astype(np.float64), want["sums"][ok])
    plain = oracle.evaluate_pairs(geo.scene, geo.points, geo.sources, pp, ps, point_bands=False)
    moved = np.abs(10 * np.log10(want["sums"][ok, 0] / plain["sums"][ok, 0]))
    print(f"geo fine pairs: n={err.size} max {err.max():.5f} dB; the detail moves pairs by up to "
          f"{np.nanmax(moved):.2f} dB", flush=True)
    assert err.max() <= 0.05
    square.close()


def test_geo_square_refused_where_unsupported(gpu, geo):
    from quietmap_amd.painter import PainterError
    from quietmap_amd.tiles import TilePlan
    from quietmap_amd import hip

    print("case: geo refusals", flush=True)
    square = gpu.upload(geo.scene, geo.sources)
    m = geo.points.size
    tiles = np.zeros(1, [("receiver_first", "<u4"), ("nx", "<u4"), ("ny", "<u4"), ("source_first", "<u4"),
                         ("source_count", "<u4")])
    tiles["nx"], tiles["ny"], tiles["source_count"] = 16, 16, 4
    d = hip.DeviceBuffer(m * 256)
    outs = [hip.DeviceBuffer(m * 24), hip.DeviceBuffer(m * 4), hip.DeviceBuffer(m * 4)]
    with pytest.raises(PainterError):
        TilePlan(gpu, square, tiles, d.ptr, m, 16, 2, *[o.ptr for o in outs])
    square.close()
