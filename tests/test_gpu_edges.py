"""Empty and out-of-range inputs on every pair, horizon and box entry point: empty calls return
nothing and touch no device memory, a receiver off the terrain reports its failure bits in every
output that carries them, and an index past its array is refused with a message (the reference's
Err(String)) rather than read."""
import numpy as np
import pytest

from quietmap_amd import abi, flights, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


@pytest.fixture(scope="module")
def tile(gpu):
    sc = scenes.tile_scene(scale=0.03)
    square = gpu.upload(sc.scene, sc.sources)
    yield sc, square
    square.close()


def test_empty_calls(gpu, tile):
    print("case: empty calls", flush=True)
    sc, square = tile
    none = np.zeros(0, np.uint32)
    out = gpu.evaluate_pairs(square, sc.points, none, none, pair_bands=True)
    assert out["sums"].shape == (0, abi.PAIR_SUMS) and not out["point_bands"].any()
    lines = scenes.line_scene()
    ls = gpu.upload(lines.scene, lines.sources)
    out = gpu.evaluate_line_pairs(ls, lines.points, lines.pieces, none, none, point_bands=True, pair_bands=True)
    assert out["sums"].shape == (0, abi.PAIR_SUMS) and not out["point_bands"].any()
    ls.close()
    rec, failed = gpu.horizons(square, sc.points[:0])
    assert rec.shape == (0, abi.HORIZON_FLOATS) and failed.size == 0
    boxes = gpu.flight_boxes(np.zeros(0, abi.FLIGHT_BOX_DTYPE))
    pts = np.zeros(1, abi.FLIGHT_POINT_DTYPE)
    pts["horizons"] = abi.NO_HORIZONS
    assert not gpu.flight_sums(boxes, pts, None, np.zeros(0, abi.FLIGHT_SEGMENT_DTYPE), 2).any()
    boxes.close()
    assert gpu.loudness(np.zeros((0, 28), np.float32)).size == 0


def test_a_receiver_off_the_terrain(gpu, tile):
    print("case: receiver off the terrain", flush=True)
    sc, square = tile
    pts = sc.points[:3].copy()
    pts["x_m"][1] = -1.0e6  # far off the lattice
    out = gpu.evaluate_pairs(square, pts, np.array([0, 1, 2], np.uint32), np.array([0, 0, 0], np.uint32),
                             pair_bands=True)
    assert out["failed"][1] != 0 and out["failed"][0] == 0 and out["failed"][2] == 0
    assert not out["sums"][1].any() and not out["pair_bands"][1].any()
    rec, failed = gpu.horizons(square, pts)
    assert failed[1] != 0 and failed[0] == 0


def test_indices_out_of_range_are_refused(gpu, tile):
    from quietmap_amd.painter import PainterError

    print("case: indices out of range", flush=True)
    sc, square = tile
    with pytest.raises(PainterError):
        gpu.evaluate_pairs(square, sc.points, np.array([sc.points.size], np.uint32), np.array([0], np.uint32))
    with pytest.raises(PainterError):
        gpu.evaluate_pairs(square, sc.points, np.array([0], np.uint32), np.array([sc.sources.size], np.uint32))
    lines = scenes.line_scene()
    ls = gpu.upload(lines.scene, lines.sources)
    with pytest.raises(PainterError):
        gpu.evaluate_line_pairs(ls, lines.points, lines.pieces, np.array([0], np.uint32),
                                np.array([lines.pieces.size], np.uint32))
    ls.close()
    bad = flights.synthetic_boxes(seed=3, count=4)
    bad["passing"][0, 0] = np.nan
    with pytest.raises(PainterError):
        gpu.flight_boxes(bad)
