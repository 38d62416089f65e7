"""Tile-in: a square's stored tiles as the card's scene (the host half of DeviceSquare::upload,
paint-gpu/src/square_upload.rs:18-287, over the formats the reference writes).

- ``parse_terrain(bytes)``: a ``qmterr2`` tile (tiles/src/terrain.rs:1-31, 140-220): the header
  (window of one arc-second nodes: north node, west node, rows, columns) and one zstd frame of
  five planes, the heights as the zigzag residual of the west + north - north-west prediction.
  The prediction inverts to a two-dimensional prefix sum in the u16 ring.
- ``parse_obstacles(bytes)``: an ``qmobs2`` tile (tiles/src/obstacles.rs:1-23, parse.rs): the
  occupied-cell bitmap and counts, each occupied cell's runs, the footprint squares, the runs,
  the outline records and the tile-local int16 vertices.
- ``parse_weather(bytes)``: the ``qmwthr2`` world table (physics/src/weather.rs:15-25, 103-113).
- ``square_scene(...)``: the tiles of a square stitched into ``abi.Scene`` in the square's
  ``LocalFrame`` (tiles/src/geo.rs:224-251): the terrain windows into one lattice of the tiles'
  own nodes (neighbours share seam nodes) with the rows' north metres through Mercator, the
  obstacle tiles' outlines in global steps, one outline per ring however many tiles store it, in
  one cell grid over the tiles' box (a tile missing inside the box reads as empty, one outside
  as not read), and the world weather table read by latitude and longitude.

zstd is the system's libzstd (the reference links the same library through its zstd crate).
"""
from __future__ import annotations

import ctypes as C
import math
import struct
from dataclasses import dataclass

import numpy as np

from . import abi
from .scenes import STEPS_PER_TILE, TILES_PER_AXIS, local_frame, mercator

TERRAIN_MAGIC = b"qmterr2\n"
TERRAIN_HEADER = 24
TERRAIN_PLANES = 5
NODES_PER_DEGREE = 3600
OBSTACLES_MAGIC = b"qmobs2\n\0"
OBSTACLES_HEADER = 32
CELLS_PER_SIDE = 128
CELLS = CELLS_PER_SIDE * CELLS_PER_SIDE
CELL_WORDS = CELLS // 64
CELL_STEPS = 256
HALF_TILE_CELLS = CELLS_PER_SIDE // 2
WEATHER_MAGIC = b"qmwthr2\n"
WEATHER_NODES_PER_DEGREE = 2
# A node of a tile without terrain (the reference's OCEAN sample): 0 m, hard (G 0), no forest.
OCEAN_NODE = 2500 | (100 << 16)

_zstd = None


def zstd():
    """The system's libzstd (ZSTD_decompress, ZSTD_compress)."""
    global _zstd
    if _zstd is None:
        lib = C.CDLL("libzstd.so.1")
        lib.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
        lib.ZSTD_decompress.restype = C.c_size_t
        lib.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
        lib.ZSTD_compress.restype = C.c_size_t
        lib.ZSTD_compressBound.argtypes = [C.c_size_t]
        lib.ZSTD_compressBound.restype = C.c_size_t
        lib.ZSTD_isError.argtypes = [C.c_size_t]
        lib.ZSTD_isError.restype = C.c_uint
        _zstd = lib
    return _zstd


@dataclass
class TerrainTile:
    """A parsed terrain tile: its window and its nodes (rows x columns, row 0 northernmost)."""

    north_node: int
    west_node: int
    height_code: np.ndarray  # u16
    impervious: np.ndarray  # u8
    forest: np.ndarray  # u8
    canopy_m: np.ndarray  # u8

    @property
    def rows(self):
        return self.height_code.shape[0]

    @property
    def columns(self):
        return self.height_code.shape[1]


def unpredict(low, high, rows, columns):
    """terrain.rs unpredict: the codes of a window from its residuals' zigzag byte planes. The
    prediction west + north - north-west inverts to the 2D prefix sum of the residuals (u16 ring)."""
    zigzag = low.astype(np.uint16) | (high.astype(np.uint16) << np.uint16(8))
    residual = (zigzag >> np.uint16(1)) ^ (np.uint16(0) - (zigzag & np.uint16(1)))
    r = residual.reshape(rows, columns)
    return np.cumsum(np.cumsum(r, axis=0, dtype=np.uint16), axis=1, dtype=np.uint16)


def parse_terrain(data) -> TerrainTile:
    data = bytes(data)
    if len(data) < TERRAIN_HEADER or data[:8] != TERRAIN_MAGIC:
        raise ValueError("terrain: bad magic")
    north, west, rows, columns = struct.unpack("<iiII", data[8:24])
    if not (0 < rows <= 1024 and 0 < columns <= 1024):
        raise ValueError("terrain: bad window")
    n = rows * columns
    out = np.empty(TERRAIN_PLANES * n, np.uint8)
    frame = np.frombuffer(data, np.uint8, offset=TERRAIN_HEADER)
    got = zstd().ZSTD_decompress(out.ctypes.data, out.size, frame.ctypes.data, frame.size)
    if zstd().ZSTD_isError(got) or got != out.size:
        raise ValueError("terrain: the planes do not decompress to the window")
    codes = unpredict(out[:n], out[n:2 * n], rows, columns)
    return TerrainTile(north, west, codes, out[2 * n:3 * n].reshape(rows, columns),
                       out[3 * n:4 * n].reshape(rows, columns), out[4 * n:].reshape(rows, columns).copy())


@dataclass
class ObstacleTile:
    """A parsed obstacles tile: per occupied cell (row * 128 + column, row 0 northernmost) its runs
    (outline, first edge, edge count), and per outline its footprint id, kind (0 exterior, 1 hole,
    2 wall), envelope, height (dm) and tile-local vertices."""

    cells: np.ndarray  # occupied cell indices
    first_run: np.ndarray  # per occupied cell and one more
    runs: np.ndarray  # (runs, 3) u32: outline, first edge, edge count
    footprint_id: np.ndarray  # u64 per outline
    kind: np.ndarray
    envelope: np.ndarray
    height_dm: np.ndarray
    vertex_first: np.ndarray  # per outline and one more
    vertices: np.ndarray  # (vertices, 2) i16, y south


def parse_obstacles(data) -> ObstacleTile:
    data = bytes(data)
    if len(data) < OBSTACLES_HEADER or data[:8] != OBSTACLES_MAGIC:
        raise ValueError("obstacles: bad magic")
    outlines, vertices, runs, occupied, squares, zero = struct.unpack("<6I", data[8:32])
    if zero != 0 or occupied > CELLS or squares > 256:
        raise ValueError("obstacles: bad header")
    lengths = [8 * CELL_WORDS, 2 * CELL_WORDS, 4 * (occupied + 1), 4 * squares, 8 * runs, 12 * outlines, 4 * vertices]
    at = [OBSTACLES_HEADER]
    for length in lengths:
        at.append(at[-1] + length)
    if at[-1] != len(data):
        raise ValueError("obstacles: length does not match the counts")
    words = np.frombuffer(data, "<u8", CELL_WORDS, at[0])
    before = np.frombuffer(data, "<u2", CELL_WORDS, at[1])
    bits = ((words[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool).reshape(-1)
    counts = bits.reshape(CELL_WORDS, 64).sum(axis=1)
    if not np.array_equal(before.astype(np.int64), np.concatenate([[0], np.cumsum(counts)[:-1]])) or bits.sum() != occupied:
        raise ValueError("obstacles: occupied cells miscounted")
    first_run = np.frombuffer(data, "<u4", occupied + 1, at[2]).astype(np.int64)
    if first_run[0] != 0 or first_run[-1] != runs or np.any(np.diff(first_run) <= 0):
        raise ValueError("obstacles: bad run index")
    square_ids = np.frombuffer(data, "<u4", squares, at[3]).astype(np.uint64)
    run_rec = np.frombuffer(data, np.dtype([("outline", "<u4"), ("first", "<u2"), ("count", "<u2")]), runs, at[4])
    rec = np.frombuffer(data, np.dtype([("first_vertex", "<u4"), ("ordinal", "<u4"), ("height", "<u2"), ("square", "u1"),
                                        ("kind", "u1")]), outlines, at[5])
    verts = np.frombuffer(data, "<i2", 2 * vertices, at[6]).reshape(vertices, 2)
    if np.any(rec["square"] >= max(squares, 1)) or (outlines and squares == 0):
        raise ValueError("obstacles: outline square out of range")
    vfirst = np.append(rec["first_vertex"].astype(np.int64), vertices)
    if np.any(np.diff(vfirst) < 2) or (outlines and vfirst[0] != 0):
        raise ValueError("obstacles: outline vertices out of range")
    footprint = (square_ids[rec["square"]] << np.uint64(32)) | rec["ordinal"].astype(np.uint64) if outlines else \
        np.zeros(0, np.uint64)
    runs_arr = np.stack([run_rec["outline"], run_rec["first"], run_rec["count"]], 1).astype(np.int64) if runs else \
        np.zeros((0, 3), np.int64)
    if runs:
        if np.any(runs_arr[:, 0] >= outlines) or np.any(runs_arr[:, 2] == 0):
            raise ValueError("obstacles: a run's outline out of range")
        edges = (vfirst[1:] - vfirst[:-1] - 1)[runs_arr[:, 0]]
        if np.any(runs_arr[:, 1] + runs_arr[:, 2] > edges):
            raise ValueError("obstacles: a run beyond its outline's edges")
    return ObstacleTile(np.flatnonzero(bits), first_run, runs_arr, footprint, rec["kind"] & 3, rec["kind"] >> 2,
                        rec["height"].astype(np.int64), vfirst, verts.astype(np.int64))


def parse_weather(data):
    """The world table's nodes: (361, 720, 64) bytes."""
    data = bytes(data)
    rows, columns = 180 * WEATHER_NODES_PER_DEGREE + 1, 360 * WEATHER_NODES_PER_DEGREE
    if len(data) != 8 + rows * columns * abi.WEATHER_NODE_BYTES or data[:8] != WEATHER_MAGIC:
        raise ValueError("weather table: bad magic or length")
    nodes = np.frombuffer(data, np.uint8, offset=8).reshape(rows, columns, abi.WEATHER_NODE_BYTES)
    if np.any(nodes[:, :, :48] > 100):
        raise ValueError("weather table: a percentage above 100")
    return nodes


def square_scene(origin_lat_lon, terrain_tiles, obstacle_tiles, weather_nodes):
    """The square's scene in the frame at `origin_lat_lon` (its LocalFrame origin):
    `terrain_tiles` {(x, y) z12: bytes or None for a tile without terrain}, `obstacle_tiles`
    {(x, y): bytes}, `weather_nodes` the world table (parse_weather). Returns (abi.Scene, frame)
    with frame = (Mercator x, y of the origin, east and north metres per unit)."""
    ox, oy = (float(v) for v in mercator(*origin_lat_lon))
    east_m, north_m = local_frame(ox, oy)
    # Terrain: the windows' global node box; the ocean's nodes where no tile has one.
    parsed = {k: parse_terrain(v) for k, v in terrain_tiles.items() if v is not None}
    if not parsed:
        raise ValueError("square_scene: no terrain tile")
    north = max(t.north_node for t in parsed.values())
    south = min(t.north_node - t.rows + 1 for t in parsed.values())
    west = min(t.west_node for t in parsed.values())
    east = max(t.west_node + t.columns - 1 for t in parsed.values())
    rows, columns = north - south + 1, east - west + 1
    nodes = np.full((rows, columns), OCEAN_NODE, np.uint32)
    canopy = np.zeros((rows, columns), np.uint8)
    for t in parsed.values():
        r0 = t.north_node - south  # our row of the tile's northernmost row (rows ascend north)
        c0 = t.west_node - west
        block = (t.height_code.astype(np.uint32) | (t.impervious.astype(np.uint32) << 16) |
                 (t.forest.astype(np.uint32) << 24))[::-1]
        nodes[r0 - t.rows + 1:r0 + 1, c0:c0 + t.columns] = block
        canopy[r0 - t.rows + 1:r0 + 1, c0:c0 + t.columns] = t.canopy_m[::-1]
    lat = np.arange(south, north + 1) / NODES_PER_DEGREE
    _, row_y = mercator(lat, np.zeros_like(lat))
    row_y_m = (oy - row_y) * north_m
    lon = np.array([west, west + 1]) / NODES_PER_DEGREE
    col_x, _ = mercator(np.zeros(2), lon)
    column_m = float((col_x[1] - col_x[0]) * east_m)
    terrain = abi.Terrain(float((col_x[0] - ox) * east_m), float(row_y_m[0]), float(np.mean(np.diff(row_y_m))), nodes,
                          canopy, column_m=column_m, row_y_m=row_y_m)
    # Obstacles: every tile's outlines in global steps, each ring once; cells of all tiles.
    keys, outlines, vertices, cell_runs = {}, [], [], {}
    for (tx, ty), blob in obstacle_tiles.items():
        tile = parse_obstacles(blob)
        centre = (tx * int(STEPS_PER_TILE) + int(STEPS_PER_TILE) // 2, ty * int(STEPS_PER_TILE) + int(STEPS_PER_TILE) // 2)
        mapping = np.zeros(tile.footprint_id.size, np.int64)
        for o in range(tile.footprint_id.size):
            v = tile.vertices[tile.vertex_first[o]:tile.vertex_first[o + 1]] + np.array(centre)
            key = (int(tile.footprint_id[o]), int(tile.kind[o]), v.tobytes())
            if key not in keys:
                keys[key] = len(outlines)
                outlines.append((int(tile.footprint_id[o]), int(tile.kind[o]), int(tile.height_dm[o]), len(v)))
                vertices.append(v)
            mapping[o] = keys[key]
        for k, cell in enumerate(tile.cells):
            row, column = divmod(int(cell), CELLS_PER_SIDE)
            global_cell = (tx * CELLS_PER_SIDE + column, ty * CELLS_PER_SIDE + row)
            for run in tile.runs[tile.first_run[k]:tile.first_run[k + 1]]:
                cell_runs.setdefault(global_cell, []).append((int(mapping[run[0]]), int(run[1]), int(run[2])))
    # The cell grid: every cell of the obstacle tiles' box (a ray leaving it is not read).
    if obstacle_tiles:
        xs = [k[0] for k in obstacle_tiles]
        ys = [k[1] for k in obstacle_tiles]
        x0, y0 = min(xs) * CELLS_PER_SIDE, min(ys) * CELLS_PER_SIDE
        nx, ny = (max(xs) + 1) * CELLS_PER_SIDE - x0, (max(ys) + 1) * CELLS_PER_SIDE - y0
    else:
        x0, y0, nx, ny = int(ox * CELLS_PER_SIDE), int(oy * CELLS_PER_SIDE), 1, 1
    first = np.zeros(nx * ny + 1, np.uint32)
    runs = []
    for index in range(nx * ny):
        first[index] = len(runs)
        y, x = divmod(index, nx)
        runs.extend(cell_runs.get((x0 + x, y0 + y), []))
    first[-1] = len(runs)
    run_arr = np.zeros(len(runs), abi.RUN_DTYPE)
    if runs:
        r = np.array(runs, np.int64)
        run_arr["outline"], run_arr["first_edge"], run_arr["edge_count"] = r[:, 0], r[:, 1], r[:, 2]
    out = np.zeros(len(outlines), abi.OUTLINE_DTYPE)
    at = 0
    for i, (fid, kind, height, count) in enumerate(outlines):
        out[i]["footprint_id"] = fid
        out[i]["first_vertex"] = at
        out[i]["vertex_count"] = count
        out[i]["height_dm"] = height
        out[i]["kind"] = abi.KIND_BARRIER if kind == 2 else abi.KIND_BUILDING
        at += count
    verts = np.concatenate(vertices).astype(np.int32) if vertices else np.zeros((0, 2), np.int32)
    obstacles = abi.Obstacles((ox * STEPS_PER_TILE, oy * STEPS_PER_TILE), east_m / STEPS_PER_TILE, (x0, y0), (nx, ny),
                              first, run_arr, out, verts, metres_per_step_y=north_m / STEPS_PER_TILE)
    weather = abi.Weather(0.0, 0.0, 1.0, weather_nodes, frame=(ox, oy, east_m, north_m),
                          nodes_per_degree=WEATHER_NODES_PER_DEGREE)
    return abi.Scene(terrain, obstacles, weather), (ox, oy, east_m, north_m)
