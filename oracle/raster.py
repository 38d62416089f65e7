"""TEST INFRASTRUCTURE — the raster-out step restated in numpy: the quantized cells of a painted pixel
and the HM3 tile's predicted stream before Brotli. Only the tests use it, as the checker of the
library's raster kernels.

    paint/src/levels.rs:26-34    cell (Nden: twice the loudness level in phon, 80 + 20 log2 sone)
    paint/src/levels.rs:87-98    lden_cell (twice Lden in dB; 254 below 0 dB; invalid energies refused)
    paint/src/levels.rs:100-109  lden_cells (six layers, then their sum before the logarithm)
    paint/src/levels.rs:58-66    PaintedCells::from_painted (255 for a pixel not assessed)
    paint/src/hm3.rs:68-93       paeth_prediction (PNG filter 4 with its tie order)
    paint/src/hm3.rs:95-108      encode_metric_tile up to Brotli (header, layer byte, residuals)
    paint/src/hm3_square.rs:22-30 the square bundle's header and directory (read for the fixture)
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

LAYERS = 6
NO_LEVEL = 255
SILENCE = 254
TILE_PX = 512
HEADERS = {"nden": b"HM3 \x07", "lden": b"HMD \x01"}


def cell(sone):
    """Nden cells of loudness values (sone), per element."""
    sone = np.asarray(sone, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        byte = 80.0 + 20.0 * np.log2(sone)
    out = np.full(sone.shape, SILENCE, np.uint8)
    ok = ~np.isnan(byte) & (byte >= 0.0)
    # Rust's f64::round: half away from zero (values here are >= 0).
    out[ok] = np.minimum(np.floor(byte[ok] + 0.5), SILENCE - 1).astype(np.uint8)
    return out


def lden_cell(energy):
    """Lden cells of linear energies; ValueError for a negative or non-finite one."""
    energy = np.asarray(energy, np.float64)
    if not np.all(np.isfinite(energy)) or np.any(energy < 0.0):
        raise ValueError("invalid Lden energy")
    out = np.full(energy.shape, SILENCE, np.uint8)
    loud = energy >= 1.0
    out[loud] = np.minimum(np.floor(2.0 * 10.0 * np.log10(energy[loud]) + 0.5), SILENCE - 1).astype(np.uint8)
    return out


def painted_cells(assessed, layer_energy, ndens):
    """Per pixel its Nden and Lden cells, layers then the total: (pixels, 7) each."""
    assessed = np.asarray(assessed, bool)
    e = np.asarray(layer_energy, np.float64).reshape(-1, LAYERS)
    nd = np.asarray(ndens, np.float64).reshape(-1, LAYERS + 1)
    nden = np.full((assessed.size, LAYERS + 1), NO_LEVEL, np.uint8)
    lden = np.full((assessed.size, LAYERS + 1), NO_LEVEL, np.uint8)
    nden[assessed] = cell(nd[assessed])
    lden[assessed, :LAYERS] = lden_cell(e[assessed])
    lden[assessed, LAYERS] = lden_cell(e[assessed].sum(axis=1))
    return nden, lden


def paeth_prediction(cells):
    """The predictor of every cell of one 512 x 512 tile (row-major), from the cells themselves."""
    c = np.asarray(cells, np.int16).reshape(TILE_PX, TILE_PX)
    left = np.zeros_like(c)
    above = np.zeros_like(c)
    corner = np.zeros_like(c)
    left[:, 1:] = c[:, :-1]
    above[1:, :] = c[:-1, :]
    corner[1:, 1:] = c[:-1, :-1]
    estimate = left + above - corner
    dl, da, dc = np.abs(estimate - left), np.abs(estimate - above), np.abs(estimate - corner)
    return np.where((dl <= da) & (dl <= dc), left, np.where(da <= dc, above, corner)).astype(np.uint8).reshape(-1)


def predicted_stream(cells, layer, metric):
    """encode_metric_tile's bytes before Brotli: the metric's header, the layer, the residuals."""
    cells = np.asarray(cells, np.uint8).reshape(-1)
    assert cells.size == TILE_PX * TILE_PX and 0 <= layer <= LAYERS
    residual = (cells.astype(np.int16) - paeth_prediction(cells).astype(np.int16)).astype(np.uint8)
    return HEADERS[metric] + bytes([layer]) + residual.tobytes()


def unpredict(residual):
    """The cells of a tile's residuals (decode_metric_tile's inverse filter), cell by cell in order."""
    r = list(bytes(residual))
    c = [0] * (TILE_PX * TILE_PX)
    for at in range(TILE_PX * TILE_PX):
        x = at % TILE_PX
        left = c[at - 1] if x else 0
        above = c[at - TILE_PX] if at >= TILE_PX else 0
        corner = c[at - TILE_PX - 1] if x and at >= TILE_PX else 0
        estimate = left + above - corner
        dl, da, dc = abs(estimate - left), abs(estimate - above), abs(estimate - corner)
        pred = left if dl <= da and dl <= dc else (above if da <= dc else corner)
        c[at] = (r[at] + pred) & 0xFF
    return np.array(c, np.uint8)


def brotli_decompress(data, capacity):
    """The system's Brotli decoder (libbrotlidec), for reading stored streams in tests."""
    lib = C.CDLL("libbrotlidec.so.1")
    lib.BrotliDecoderDecompress.argtypes = [C.c_size_t, C.c_char_p, C.POINTER(C.c_size_t), C.c_void_p]
    lib.BrotliDecoderDecompress.restype = C.c_int
    out = (C.c_uint8 * capacity)()
    size = C.c_size_t(capacity)
    if lib.BrotliDecoderDecompress(len(data), bytes(data), C.byref(size), out) != 1:
        raise ValueError("not a Brotli stream")
    return bytes(out[: size.value])


def read_square(data):
    """A square bundle's header fields and its tiles' stored streams by (metric, layer, ty, tx)."""
    assert data[:4] == b"HM3S"
    version, zoom = data[4], data[5]
    x, y = struct.unpack("<II", data[6:14])
    per_side = 1 << (zoom - 12)
    metrics = ["nden"] + (["lden"] if version == 8 else [])
    streams = {}
    at = 14
    for metric in metrics:
        for layer in range(LAYERS + 1):
            for ty in range(per_side):
                for tx in range(per_side):
                    offset, length = struct.unpack("<II", data[at:at + 8])
                    at += 8
                    streams[(metric, layer, ty, tx)] = data[offset:offset + length]
    return dict(version=version, zoom=zoom, square=(x, y), per_side=per_side, streams=streams)
