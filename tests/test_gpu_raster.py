"""Raster-out on the card against the numpy restatement (oracle/raster.py), byte for byte: the
Nden and Lden cells of painted pixels (paint/src/levels.rs PaintedCells::from_painted) over levels
spanning silence to the top of the scale and pixels not assessed, the refusal of an invalid
energy naming its pixel and layer, and the HM3 predicted streams (paint/src/hm3.rs
encode_metric_tile before Brotli) of random and smooth tiles and of the reference's stored square
(tests/golden/square-v7.hm3), which the card re-encodes to the very bytes stored."""
import os

import numpy as np
import pytest

from oracle import raster

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "square-v7.hm3")


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def test_level_cells_bit_exact(gpu):
    print("case: level cells", flush=True)
    rng = np.random.default_rng(41)
    n = 200_000
    assessed = rng.uniform(size=n) < 0.9
    energy = 10.0 ** rng.uniform(-2.0, 14.0, (n, 6))
    energy[rng.uniform(size=(n, 6)) < 0.05] = 0.0
    ndens = 2.0 ** rng.uniform(-6.0, 9.0, (n, 7))
    ndens[rng.uniform(size=(n, 7)) < 0.02] = 0.0
    # Values a quarter step either side of the rounding edges of both scales.
    energy[:1000, 0] = 10.0 ** ((np.arange(1000) * 0.25 + 0.125) / 20.0)
    ndens[:1000, 0] = 2.0 ** ((np.arange(1000) * 0.25 - 80.0 + 0.125) / 20.0)
    got_n, got_l = gpu.level_cells(assessed, energy, ndens)
    want_n, want_l = raster.painted_cells(assessed, energy, ndens)
    assert np.array_equal(got_n, want_n)
    assert np.array_equal(got_l, want_l)
    assert (got_l == 254).any() and (got_l == 253).any() and (got_n == 255).any()


def test_invalid_energy_fails_naming_the_pixel(gpu):
    from quietmap_amd.painter import PainterError

    print("case: level cells invalid energy", flush=True)
    energy = np.ones((10, 6))
    energy[7, 2] = -1.0
    energy[8, 4] = np.nan
    with pytest.raises(PainterError, match="pixel 7, layer 2"):
        gpu.level_cells(np.ones(10, np.uint8), energy, np.ones((10, 7)))


def test_hm3_predicted_streams_bit_exact(gpu):
    print("case: hm3 predict", flush=True)
    rng = np.random.default_rng(43)
    at = np.arange(512 * 512)
    tiles = [rng.integers(0, 256, 512 * 512, dtype=np.uint8),
             ((at // 512 * 19 + at % 512 * 31) % 256).astype(np.uint8),
             np.clip(80 + (np.sin(at / 5000.0) * 60).astype(int), 0, 253).astype(np.uint8)]
    square = raster.read_square(open(GOLDEN, "rb").read())
    stored = []
    for (metric, layer, ty, tx), stream in sorted(square["streams"].items()):
        decoded = raster.brotli_decompress(stream, 6 + 512 * 512 + 16)
        stored.append((layer, decoded))
        tiles.append(raster.unpredict(decoded[6:]))
    layers = np.array([6, 0, 3] + [layer for layer, _ in stored], np.uint8)
    for metric in ("nden", "lden"):
        got = gpu.hm3_predict(np.stack(tiles), layers, metric)
        for k, cells in enumerate(tiles):
            assert got[k].tobytes() == raster.predicted_stream(cells, int(layers[k]), metric), (metric, k)
    got = gpu.hm3_predict(np.stack(tiles[3:]), layers[3:], "nden")
    for k, (layer, decoded) in enumerate(stored):
        assert got[k].tobytes() == decoded
