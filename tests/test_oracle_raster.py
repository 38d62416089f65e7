"""The raster-out restatement (oracle/raster.py) against what pins it: the reference's lden_cell
known answers (paint/src/levels.rs:114-136), the Nden cell's closed form (1 sone is 40 phon, byte
80; a doubling is 20 bytes; 1/16 sone is 0 phon), the reference's stored square bundle
(paint/tests/fixtures/square-v7.hm3, copied to tests/golden/: each tile's stream, Brotli-decoded
by the system library, is exactly the predicted stream of the cells it decodes to), and the
reference's every-byte-and-row-boundary pattern (hm3_tests.rs:23-39) round-tripping through the
predictor and the sequential inverse filter."""
import os

import numpy as np
import pytest

from oracle import raster

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "square-v7.hm3")


def energy(db):
    return 10.0 ** (np.asarray(db, np.float64) / 10.0)


def test_lden_cells_known_answers():
    assert raster.lden_cell(0.0) == 254
    assert raster.lden_cell(energy(-0.01)) == 254
    assert raster.lden_cell(1.0) == 0
    assert raster.lden_cell(energy(60.26)) == 121
    assert raster.lden_cell(energy(150.0)) == 253
    layers = np.zeros(6)
    layers[0] = energy(60.0)
    layers[5] = energy(60.0)
    nden, lden = raster.painted_cells([True], layers, np.ones(7))
    assert lden[0, 6] == 126 and lden[0, 1] == 254
    for bad in (-1.0, np.inf, np.nan):
        with pytest.raises(ValueError):
            raster.lden_cell(bad)


def test_nden_cells_closed_form():
    assert list(raster.cell([1.0, 2.0, 0.5, 1.0 / 16.0, 1.0 / 17.0, 0.0, np.nan, 1e9])) == [80, 100, 60, 0, 254,
                                                                                              254, 254, 253]
    nden, lden = raster.painted_cells([False, True], np.ones((2, 6)), np.ones((2, 7)))
    assert (nden[0] == 255).all() and (lden[0] == 255).all() and (nden[1] == 80).all()


def test_the_references_stored_square_is_the_predicted_stream():
    square = raster.read_square(open(GOLDEN, "rb").read())
    assert square["version"] == 7 and square["zoom"] == 12 and square["square"] == (9, 17)
    assert len(square["streams"]) == 7
    for (metric, layer, ty, tx), stream in square["streams"].items():
        decoded = raster.brotli_decompress(stream, 6 + 512 * 512 + 16)
        assert len(decoded) == 6 + 512 * 512
        cells = raster.unpredict(decoded[6:])
        assert raster.predicted_stream(cells, layer, metric) == decoded
        assert cells.min() < cells.max()


def test_every_byte_and_row_boundary_round_trips():
    at = np.arange(512 * 512)
    cells = ((at // 512 * 19 + at % 512 * 31) % 256).astype(np.uint8)
    stream = raster.predicted_stream(cells, 6, "lden")
    assert stream[:6] == b"HMD \x01\x06"
    assert np.array_equal(raster.unpredict(stream[6:]), cells)
