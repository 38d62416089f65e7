"""Tile-in (quietmap_amd/tiles_in.py): the reference's stored formats read into the card's scene.
There are no stored tiles in the reference to read, so each format is written here from its
documented layout (tiles/src/terrain.rs:1-31 and encode, tiles/src/obstacles.rs:1-23 and encode,
physics/src/weather.rs:15-25) and read back: the heights' planar prediction against a node-by-node
restatement, every field of every outline and run, corrupt files refused, and a square of four z12
tiles stitched into one lattice whose nodes, rows and outlines are the world's they were written
from."""
import math
import struct

import numpy as np
import pytest

from quietmap_amd import abi, scenes, tiles_in
from tests.tile_formats import encode_obstacles, encode_terrain, cells_of_edge, four_tile_square, world_height_code



def test_unpredict_against_the_node_by_node_rule():
    rng = np.random.default_rng(1)
    codes = rng.integers(0, 65535, (37, 53)).astype(np.uint16)
    codes[5:9, 5:9] = 65535  # nodes without data survive
    blob = encode_terrain(180000, 51000, codes, rng.integers(0, 101, codes.shape), rng.integers(0, 101, codes.shape),
                          rng.integers(0, 40, codes.shape))
    t = tiles_in.parse_terrain(blob)
    assert (t.north_node, t.west_node, t.rows, t.columns) == (180000, 51000, 37, 53)
    assert np.array_equal(t.height_code, codes)
    # terrain.rs unpredict restated node by node over the stored residuals.
    n = codes.size
    planes = np.empty(5 * n, np.uint8)
    frame = np.frombuffer(blob, np.uint8, offset=24)
    tiles_in.zstd().ZSTD_decompress(planes.ctypes.data, planes.size, frame.ctypes.data, frame.size)
    got = np.zeros((37, 53), np.int64)
    for r in range(37):
        for c in range(53):
            z = int(planes[r * 53 + c]) | int(planes[n + r * 53 + c]) << 8
            residual = ((z >> 1) ^ -(z & 1)) & 0xFFFF
            w = got[r, c - 1] if c else 0
            no = got[r - 1, c] if r else 0
            nw = got[r - 1, c - 1] if r and c else 0
            got[r, c] = (residual + w + no - nw) % 65536
    assert np.array_equal(got, codes)
    with pytest.raises(ValueError):
        tiles_in.parse_terrain(blob[:-3])
    with pytest.raises(ValueError):
        tiles_in.parse_terrain(b"qmterr1\n" + blob[8:])


def test_obstacles_round_trip_and_refusals():
    rng = np.random.default_rng(2)
    outlines = []
    for k in range(60):
        fid = (int(rng.integers(1, 4)) << 32) | k
        centre = rng.integers(-15000, 15000, 2)
        if k % 7 == 3:
            v = [tuple(centre), tuple(centre + rng.integers(-900, 900, 2))]
            outlines.append((fid, 2, 0, int(rng.integers(10, 60)), v))
            continue
        n = int(rng.integers(3, 9))
        ang = np.sort(rng.uniform(0, 2 * math.pi, n))
        rad = rng.uniform(20, 300, n)
        v = [tuple((centre + rad[i] * np.array([math.cos(ang[i]), math.sin(ang[i])])).astype(int)) for i in range(n)]
        v.append(v[0])
        outlines.append((fid, k % 2 if k % 5 == 0 else 0, int(rng.integers(0, 6)), int(rng.integers(30, 600)), v))
    outlines.sort(key=lambda o: o[0])
    blob = encode_obstacles(outlines)
    t = tiles_in.parse_obstacles(blob)
    assert np.array_equal(t.footprint_id, np.array([o[0] for o in outlines], np.uint64))
    assert np.array_equal(t.kind, [o[1] for o in outlines]) and np.array_equal(t.envelope, [o[2] for o in outlines])
    assert np.array_equal(t.height_dm, [o[3] for o in outlines])
    for i, o in enumerate(outlines):
        assert np.array_equal(t.vertices[t.vertex_first[i]:t.vertex_first[i + 1]], np.array(o[4]))
    # Every edge listed in its cells, once.
    seen = {}
    for k, cell in enumerate(t.cells):
        for o, e, n in t.runs[t.first_run[k]:t.first_run[k + 1]]:
            for edge in range(e, e + n):
                seen.setdefault((int(o), edge), []).append(int(cell))
    for o, (*_, v) in enumerate(outlines):
        for e in range(len(v) - 1):
            want = sorted(row * 128 + column for column, row in cells_of_edge(np.asarray(v[e]), np.asarray(v[e + 1])))
            assert sorted(seen[(o, e)]) == want
    for bad in (blob[:-4], blob[:8] + struct.pack("<I", 9999) + blob[12:], b"qmobs1\n\0" + blob[8:]):
        with pytest.raises(ValueError):
            tiles_in.parse_obstacles(bad)


def test_square_of_four_tiles():
    """Four z12 tiles about Prague written from one world (heights a smooth function of latitude and
    longitude, buildings at global steps) and stitched: every node the world's, the rows' north
    metres through Mercator, each building once in global steps however many tiles store it."""
    scene, frame, world, _ = four_tile_square()
    height_code = world_height_code
    T = scene.terrain
    ox, oy, east_m, north_m = frame
    # Every node is the world's at its global node.
    rows, cols = T.nodes.shape
    lat_rows = np.array([scenes.mercator_degrees(ox, oy - y / north_m)[0] for y in T.row_y_m])
    lat_index = np.round(lat_rows * 3600).astype(int)
    assert np.allclose(lat_rows * 3600, lat_index, atol=1e-6) and np.all(np.diff(lat_index) == 1)
    lon_first = scenes.mercator_degrees(ox + T.origin_x_m / east_m, oy)[1] * 3600
    lon_index = round(lon_first) + np.arange(cols)
    assert abs(lon_first - round(lon_first)) < 1e-6
    want = height_code(lat_index[:, None] / 3600.0, lon_index[None, :] / 3600.0)
    assert np.array_equal(T.nodes & 0xFFFF, want.astype(np.uint32))
    # Each building once, at its global steps, with every edge in the cell index.
    O = scene.obstacles
    assert O.outlines.size == len(world)
    by_id = {int(o["footprint_id"]): o for o in O.outlines}
    for fid, h, ring in world:
        o = by_id[fid]
        v = O.vertices[o["first_vertex"]:o["first_vertex"] + o["vertex_count"]]
        assert np.array_equal(v, ring) and o["height_dm"] == h
    assert O.metres_per_step_y != O.metres_per_step
