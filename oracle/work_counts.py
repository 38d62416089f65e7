"""Algorithmic work per path of a profiled workload, counted by the oracle on a sample of its paths
(CPU only; test and measurement infrastructure like the rest of this folder).

    python -m oracle.work_counts --workload mosaic

`mosaic`: the batched uneven tiles at full size (quietmap_amd.scenes.tile_mosaic_fast, 64 x 64
tiles of 256 x 256 receivers, each tile's receivers against its own sources): `--tiles` random
tiles, `--receivers` random receivers of each, all their tile's sources. Prints the counts per path
(vertices, cells, occupied cells, edges, crossings, candidates), the FLOP and compulsory bytes per
path by quietmap_amd/roofline.json's fixed weights, and the share of sampled paths the oracle
reports as failed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from quietmap_amd import scenes  # noqa: E402


def per_path(work: dict, paths: int, weights: dict) -> dict:
    w = {k: v / max(paths, 1) for k, v in work.items()}
    f, b = weights["flop_per_path"], weights["bytes_per_path"]
    flop = (f["V"] * w["vertices"] + f["C"] * w["cells"] + f["E"] * w["edges"] + f["X"] * w["crossings"] +
            f["K"] * w["candidates"] + f["const"])
    compulsory = (b["V"] * w["vertices"] + b["C"] * w["cells"] + b["C_occ"] * w["cells_occupied"] +
                  b["E"] * w["edges"] + b["X"] * w["crossings"] + b["const"])
    return {"work_per_path": w, "flop_per_path": flop, "compulsory_bytes_per_path": compulsory}


def mosaic(args, weights):
    m = scenes.tile_mosaic_fast(tiles=(args.side, args.side))
    table = m.tile_table()
    rng = np.random.default_rng(args.seed)
    tiles = rng.choice(m.n_tiles, args.tiles, replace=False)
    points, pp, ps = [], [], []
    for t in tiles:
        first, nx, ny = int(table["receiver_first"][t]), int(table["nx"][t]), int(table["ny"][t])
        chosen = first + rng.choice(nx * ny, args.receivers, replace=False)
        s0, sn = int(table["source_first"][t]), int(table["source_count"][t])
        for i in chosen:
            pp.append(np.full(sn, len(points), np.uint32))
            ps.append(np.arange(s0, s0 + sn, dtype=np.uint32))
            points.append(m.points[i])
    points = np.array(points, m.points.dtype)
    pp, ps = np.concatenate(pp), np.concatenate(ps)
    res = oracle.evaluate_pairs(m.scene, points, m.sources, pp, ps, threads=args.threads, point_bands=False)
    out = per_path(res["work"], pp.size, weights)
    out.update({"workload": f"mosaic {args.side}x{args.side} tiles of 256x256 receivers",
                "sample": f"{args.tiles} random tiles x {args.receivers} receivers, all their tile's sources",
                "paths": int(pp.size), "failed_share": float((res["failed"] != 0).mean())})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="mosaic", choices=["mosaic"])
    ap.add_argument("--side", type=int, default=64)
    ap.add_argument("--tiles", type=int, default=16)
    ap.add_argument("--receivers", type=int, default=8)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    args.threads = args.threads or os.cpu_count() or 1
    weights = json.load(open(os.path.join(ROOT, "quietmap_amd", "roofline.json")))
    print(json.dumps(mosaic(args, weights)), flush=True)


if __name__ == "__main__":
    main()
