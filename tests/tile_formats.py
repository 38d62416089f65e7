"""Writers of the reference's tile formats from their documented layouts (tiles/src/terrain.rs:1-31
and encode, tiles/src/obstacles.rs:1-23 and encode), for the tile-in tests, and a square of four
z12 tiles about Prague written from one synthetic world."""
import math
import struct

import numpy as np

from quietmap_amd import scenes, tiles_in

STEPS = int(scenes.STEPS_PER_TILE)


def encode_terrain(north, west, codes, impervious, forest, canopy):
    """qmterr2: header, then one zstd frame of the five planes (height residual of the west + north
    - north-west prediction, zigzag, low and high bytes; imperviousness; forest; canopy)."""
    rows, columns = codes.shape
    c = codes.astype(np.int64)
    pred = np.zeros_like(c)
    pred[:, 1:] += c[:, :-1]
    pred[1:, :] += c[:-1, :]
    pred[1:, 1:] -= c[:-1, :-1]
    residual = ((c - pred) % 65536).astype(np.uint16).view(np.int16).astype(np.int32)
    zigzag = ((residual << 1) ^ (residual >> 15)).astype(np.uint32) & 0xFFFF
    planes = np.concatenate([(zigzag & 0xFF).astype(np.uint8).ravel(), (zigzag >> 8).astype(np.uint8).ravel(),
                             impervious.astype(np.uint8).ravel(), forest.astype(np.uint8).ravel(),
                             canopy.astype(np.uint8).ravel()])
    z = tiles_in.zstd()
    cap = z.ZSTD_compressBound(planes.size)
    out = np.empty(cap, np.uint8)
    n = z.ZSTD_compress(out.ctypes.data, cap, planes.ctypes.data, planes.size, 3)
    assert not z.ZSTD_isError(n)
    return tiles_in.TERRAIN_MAGIC + struct.pack("<iiII", north, west, rows, columns) + out[:n].tobytes()


def cells_of_edge(a, b):
    """The tile cells (column, row from the north-west corner) an edge's step box touches, padded a
    step: a superset of the cells it passes through."""
    lo = np.minimum(a, b) - 1 + tiles_in.HALF_TILE_CELLS * tiles_in.CELL_STEPS
    hi = np.maximum(a, b) + 1 + tiles_in.HALF_TILE_CELLS * tiles_in.CELL_STEPS
    c0, c1 = lo // tiles_in.CELL_STEPS, hi // tiles_in.CELL_STEPS
    for row in range(max(c0[1], 0), min(c1[1], tiles_in.CELLS_PER_SIDE - 1) + 1):
        for column in range(max(c0[0], 0), min(c1[0], tiles_in.CELLS_PER_SIDE - 1) + 1):
            yield column, row


def encode_obstacles(outlines):
    """qmobs2 from (footprint id, kind, envelope, height dm, tile-local vertices), sorted by id."""
    cells = [[] for _ in range(tiles_in.CELLS)]
    for o, (_, _, _, _, v) in enumerate(outlines):
        for e in range(len(v) - 1):
            for column, row in cells_of_edge(np.asarray(v[e]), np.asarray(v[e + 1])):
                runs = cells[row * tiles_in.CELLS_PER_SIDE + column]
                if runs and runs[-1][0] == o and runs[-1][1] + runs[-1][2] == e:
                    runs[-1][2] += 1
                else:
                    runs.append([o, e, 1])
    squares = sorted({fid >> 32 for fid, *_ in outlines})
    occupied = [i for i, r in enumerate(cells) if r]
    words = np.zeros(tiles_in.CELL_WORDS, np.uint64)
    for i in occupied:
        words[i // 64] |= np.uint64(1) << np.uint64(i % 64)
    counts = np.array([bin(int(w)).count("1") for w in words])
    before = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype("<u2")
    first, total = [], 0
    for i in occupied:
        first.append(total)
        total += len(cells[i])
    first.append(total)
    nverts = sum(len(o[4]) for o in outlines)
    out = bytearray(tiles_in.OBSTACLES_MAGIC)
    out += struct.pack("<6I", len(outlines), nverts, total, len(occupied), len(squares), 0)
    out += words.astype("<u8").tobytes() + before.tobytes() + np.array(first, "<u4").tobytes()
    out += np.array(squares, "<u4").tobytes()
    for i in occupied:
        for o, e, n in cells[i]:
            out += struct.pack("<IHH", o, e, n)
    at = 0
    for fid, kind, envelope, height, v in outlines:
        out += struct.pack("<IIHBB", at, fid & 0xFFFFFFFF, height, squares.index(fid >> 32), kind | (envelope << 2))
        at += len(v)
    for *_, v in outlines:
        out += np.asarray(v, "<i2").tobytes()
    return bytes(out)


def world_height_code(lat, lon):
    """The synthetic world's height codes at (latitude, longitude)."""
    return (2500 + 5 * (200 + 30 * np.sin(lat * 900.0) * np.cos(lon * 700.0))).round().astype(np.uint16)


def four_tile_square(lat0=50.0875, lon0=14.4214, buildings=40, seed=3):
    """Four z12 tiles (the one holding (lat0, lon0), its east, south and south-east neighbours)
    written from one world, and the square read from them: (scene, frame, world buildings,
    (tx, ty)). Buildings are rings in global steps within 4,000 steps of the tiles' shared corner,
    stored in every tile (their int16 frames reach half a tile beyond the tile)."""
    mx, my = scenes.mercator(lat0, lon0)
    tx, ty = int(mx), int(my)
    height_code = world_height_code
    terrain, obstacles = {}, {}
    rng = np.random.default_rng(seed)
    corner = np.array([(tx + 1) * STEPS, (ty + 1) * STEPS])
    world = []
    for k in range(buildings):
        c = corner + rng.integers(-4000, 4000, 2)
        half = rng.integers(30, 200, 2)
        ring = [c + [-half[0], -half[1]], c + [half[0], -half[1]], c + [half[0], half[1]], c + [-half[0], half[1]]]
        ring.append(ring[0])
        world.append(((7 << 32) | k, int(rng.integers(30, 400)), np.array(ring)))
    for dy in (0, 1):
        for dx in (0, 1):
            x, y = tx + dx, ty + dy
            # The tile's window: floor/ceil of its edges (terrain.rs Window::of_tile).
            west = (x * 360 * 3600) // 4096 - 180 * 3600
            east = -((-(x + 1) * 360 * 3600) // 4096) - 180 * 3600
            lat_n, _ = scenes.mercator_degrees(0.0, y)
            lat_s, _ = scenes.mercator_degrees(0.0, y + 1)
            north, south = int(math.ceil(lat_n * 3600)), int(math.floor(lat_s * 3600))
            rows, cols = north - south + 1, east - west + 1
            lat = (north - np.arange(rows)) / 3600.0
            lon = (west + np.arange(cols)) / 3600.0
            codes = height_code(lat[:, None], lon[None, :])
            imp = ((np.arange(rows)[:, None] * 7 + np.arange(cols)[None, :] * 3 + x + y) % 101)
            terrain[(x, y)] = encode_terrain(north, west, codes, imp, np.zeros_like(imp), np.zeros_like(imp))
            centre = np.array([x * STEPS + STEPS // 2, y * STEPS + STEPS // 2])
            mine = []
            for fid, h, ring in world:
                local = ring - centre
                if np.all(np.abs(local) < 32000):
                    mine.append((fid, 0, 1, h, [tuple(p) for p in local]))
            obstacles[(x, y)] = encode_obstacles(sorted(mine, key=lambda o: o[0]))
    nodes = np.zeros((361, 720, 64), np.uint8)
    nodes[:, :, :48] = 50
    scene, frame = tiles_in.square_scene((lat0, lon0), terrain, obstacles,
                                         tiles_in.parse_weather(tiles_in.WEATHER_MAGIC + nodes.tobytes()))
    return scene, frame, world, (tx, ty)
