"""The batched painter of many uneven tiles and its autotuner.

A plan (include/qm_tiles.h) cuts the tiles into work items of `block_sources` sources against a
16x16 receiver block, queues them heaviest tile first over the XCDs' stripes, and captures the
clearing kernels and one persistent work-stealing launch in a hipGraph; `run` replays it.
`autotune` times candidate (block_sources, groups_per_cu) pairs on a sample of the tiles and
returns the fastest with every timing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi, hip
from .painter import _check, library

CANDIDATES = ((32, 2), (64, 2), (128, 2), (256, 2), (512, 2), (1024, 2), (128, 1), (256, 1), (128, 3), (256, 3))


class TilePlan:
    """A captured plan over `tiles` (abi.TILE_DTYPE) with device outputs per receiver."""

    def __init__(self, gpu, square, tiles, d_receivers, n_receivers, block_sources, groups_per_cu, d_energy, d_paths,
                 d_failures):
        self.tiles = np.ascontiguousarray(tiles, abi.TILE_DTYPE)
        self._lib = library()
        err = C.create_string_buffer(512)
        handle = C.c_void_p()
        _check(self._lib.qm_tiles_plan(gpu.handle, square.handle, self.tiles.ctypes.data, self.tiles.size, d_receivers,
                                       n_receivers, block_sources, groups_per_cu, d_energy, d_paths, d_failures,
                                       C.byref(handle), err, 512), err)
        self.handle = handle
        self.block_sources, self.groups_per_cu = block_sources, groups_per_cu

    @property
    def paths(self) -> int:
        return int(self._lib.qm_tiles_paths(self.handle))

    @property
    def items(self) -> int:
        return int(self._lib.qm_tiles_items(self.handle))

    def run(self, stream=None):
        err = C.create_string_buffer(512)
        _check(self._lib.qm_tiles_run(self.handle, stream, err, 512), err)

    def close(self):
        if self.handle:
            self._lib.qm_tiles_plan_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sample_tiles(tiles, count):
    """`count` tiles spanning the cost range (paths per tile), evenly spaced in cost order."""
    tiles = np.asarray(tiles, abi.TILE_DTYPE)
    cost = tiles["source_count"].astype(np.int64) * tiles["nx"] * tiles["ny"]
    order = np.argsort(cost)
    pick = np.unique(np.linspace(0, tiles.size - 1, min(count, tiles.size)).round().astype(int))
    return tiles[order[pick]]


def autotune(gpu, square, tiles, d_receivers, n_receivers, candidates=CANDIDATES, sample=8, repeats=2):
    """The fastest (block_sources, groups_per_cu) on a sample of the tiles, with every timing
    (paths/s, the median of `repeats` replays after one warm replay)."""
    chosen = sample_tiles(tiles, sample)
    d_e, d_p, d_f = hip.DeviceBuffer(n_receivers * 24), hip.DeviceBuffer(n_receivers * 4), hip.DeviceBuffer(n_receivers * 4)
    stream = hip.Stream()
    timings = []
    for block, groups in candidates:
        plan = TilePlan(gpu, square, chosen, d_receivers, n_receivers, block, groups, d_e.ptr, d_p.ptr, d_f.ptr)
        plan.run(stream.handle)
        stream.synchronize()
        ts = []
        for _ in range(repeats):
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            plan.run(stream.handle)
            e1.record(stream)
            e1.synchronize()
            ts.append(e1.seconds_since(e0))
        timings.append({"block_sources": block, "groups_per_cu": groups, "paths_per_s": plan.paths / float(np.median(ts)),
                        "items": plan.items})
        plan.close()
    best = max(timings, key=lambda t: t["paths_per_s"])
    return {"block_sources": best["block_sources"], "groups_per_cu": best["groups_per_cu"], "sample_tiles": int(chosen.size),
            "timings": timings}
