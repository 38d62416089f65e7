"""Aircraft boxes on the card (qm_flight_sums: the reference card's flight_sums, paint-gpu/kernels/
aircraft.cuh:260-311, over the host's box records) against the FP64 oracle, whose box physics is
pinned by the reference's boxes_tests.rs checks: each sum slot's energy per period within
0.05 dB (FP32 terms; measured differences are ~1e-5 dB), every pass in the oracle's class (the
non-empty classes equal, class energies within 1e-4 relative, passing shares 1e-9), unscreened
and screened by horizon records, with far and near weights adding up to the whole, segments of
one point sharing a slot, and records out of range refused."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi, flights, scenes

pytestmark = pytest.mark.gpu

EXTENT = 20000.0


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def point_grid(n, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.15 * EXTENT, 0.85 * EXTENT, (n, 2))


def db(x):
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.asarray(x, np.float64))


def compare(got, want, label):
    e_got, e_want = got[:, :3], want[:, :3]
    assert np.array_equal(e_got > 0, e_want > 0), label
    heard = e_want > 0
    err = np.abs(db(e_got[heard]) - db(e_want[heard]))
    assert err.max() < 0.05, (label, err.max())
    c_got, c_want = got[:, 3 : 3 + 144], want[:, 3 : 3 + 144]
    p_got, p_want = got[:, 3 + 144 :], want[:, 3 + 144 :]
    total = np.repeat(e_want, 48, axis=1)
    assert np.array_equal(p_got > 0, p_want > 0), label
    assert np.allclose(p_got, p_want, rtol=1e-9, atol=0.0), label
    assert np.all(np.abs(c_got - c_want) <= 1e-4 * np.abs(c_want) + 1e-9 * total), label
    return err.max()


@pytest.mark.parametrize("screened", [False, True])
def test_flight_sums_match_oracle(gpu, screened):
    print(f"case: flight sums screened={screened}", flush=True)
    boxes = flights.synthetic_boxes(seed=11, count=3000, extent_m=EXTENT)
    xy = point_grid(48)
    boxes, points, segments, n_sums = flights.flight_requests(boxes, xy, extent_m=EXTENT, horizons=screened)
    horizons = None
    if screened:
        rng = np.random.default_rng(4)
        records = []
        for _ in range(points.size):
            rec = scenes._sector_edges(rng, abi.TERRAIN_SECTORS, scenes.TERRAIN_BAND_LIMITS_M, 0.0, 0.08, 0.3)
            rec += scenes._sector_edges(rng, abi.BUILDING_SECTORS, scenes.BUILDING_BAND_LIMITS_M, 0.05, 0.6, 0.5)
            records.append(np.array(rec, np.float32))
        horizons = np.concatenate(records)
        points["altitude_m"] = 2.0
    box_handle = gpu.flight_boxes(boxes)
    got = gpu.flight_sums(box_handle, points, horizons, segments, n_sums)
    want = oracle.flight_sums(boxes, points, horizons, segments, n_sums)
    worst = compare(got, want, f"screened={screened}")
    print(f"max energy difference {worst:.2e} dB", flush=True)
    assert (want[:, 3 + 144 :] > 0).sum() > 100
    if screened:
        # The horizons screen: every slot loses energy against the same points unscreened.
        plain = points.copy()
        plain["horizons"] = abi.NO_HORIZONS
        free = oracle.flight_sums(boxes, plain, None, segments, n_sums)
        assert np.all(want[:, :3] <= free[:, :3] * (1 + 1e-12)) and (want[:, 0] < 0.99 * free[:, 0]).mean() > 0.5
    box_handle.close()


def test_weights_add_up_and_share_slots(gpu):
    print("case: flight weights", flush=True)
    boxes = flights.synthetic_boxes(seed=12, count=1500, extent_m=EXTENT)
    xy = point_grid(16, seed=9)
    weights = (abi.FLIGHT_WEIGHT_ONE, abi.FLIGHT_WEIGHT_FAR, abi.FLIGHT_WEIGHT_NEAR)
    boxes, points, segments, n_sums = flights.flight_requests(boxes, xy, extent_m=EXTENT, weights=weights,
                                                             coarse_m=2000.0, within_m=12000.0)
    box_handle = gpu.flight_boxes(boxes)
    got = gpu.flight_sums(box_handle, points, None, segments, n_sums).reshape(points.size, 3, abi.FLIGHT_SUMS)
    want = oracle.flight_sums(boxes, points, None, segments, n_sums).reshape(points.size, 3, abi.FLIGHT_SUMS)
    compare(got.reshape(-1, abi.FLIGHT_SUMS), want.reshape(-1, abi.FLIGHT_SUMS), "weights")
    one, far, near = got[:, 0, :3], got[:, 1, :3], got[:, 2, :3]
    assert np.allclose(far + near, one, rtol=1e-6)
    # All of a point's tiles add into its slot: one segment per tile here, 16 tiles.
    assert np.unique(segments["sum"]).size == n_sums < segments.size
    box_handle.close()


def test_records_out_of_range_fail(gpu):
    from quietmap_amd.painter import PainterError

    boxes = flights.synthetic_boxes(seed=13, count=64, extent_m=EXTENT)
    boxes, points, segments, n_sums = flights.flight_requests(boxes, point_grid(2), extent_m=EXTENT)
    handle = gpu.flight_boxes(boxes)
    for field, value in (("point", 5), ("sum", n_sums), ("count", 10_000)):
        bad = segments.copy()
        bad[field][0] = value
        with pytest.raises(PainterError):
            gpu.flight_sums(handle, points, None, bad, n_sums)
    bad_points = points.copy()
    bad_points["horizons"] = 0
    with pytest.raises(PainterError):
        gpu.flight_sums(handle, bad_points, np.zeros(10, np.float32), segments, n_sums)
    empty = gpu.flight_sums(handle, points, None, segments[:0], n_sums)
    assert not empty.any()
    handle.close()
