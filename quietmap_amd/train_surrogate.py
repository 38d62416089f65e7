"""Trains the excess-attenuation surrogate on the device and writes its parameters.

    python -m quietmap_amd.train_surrogate --out quietmap_amd/surrogate_params.npy

Samples: paths of a tile-type scene (seed 11, a different scene from the benchmark's seed 2) by
the exact kernel: `--sources` random sources, each with every `--stride`-th receiver row and column
between 500 m and the 2 km cutoff, kept where the path lies in the network's envelope; two rows a
path (calm, favourable), the exact A_ground + A_dif per band as targets. Training: Adam in
shuffled batches, the output biases starting at the per-band target means. Prints one line per
100 steps and a JSON summary (rows, steps, losses, seconds).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import scenes, surrogate
from .painter import Gpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_params.npy"))
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--scale", type=float, default=0.5)
    ap.add_argument("--sources", type=int, default=27000)
    ap.add_argument("--stride", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-end", type=float, default=2e-5)
    args = ap.parse_args()
    t0 = time.time()
    sc = scenes.tile_scene(seed=args.seed, scale=args.scale)
    print(f"scene: {sc.grid_shape} receivers, {sc.sources.size} sources ({time.time() - t0:.1f} s)", flush=True)
    gpu = Gpu(0)
    t1 = time.time()
    samples, (pr, _) = surrogate.generate_samples(gpu, sc, args.sources, seed=args.seed, stride=args.stride)
    t_gen = time.time() - t1
    print(f"samples: {pr.size} paths in the envelope, {samples.rows} rows ({t_gen:.1f} s)", flush=True)
    means = surrogate.band_means(samples)
    params0 = surrogate.with_output_bias(surrogate.init_params(args.seed), means)
    t2 = time.time()
    params, history = surrogate.train(gpu, samples, params0, epochs=args.epochs, batch=args.batch, lr=args.lr,
                                      lr_end=args.lr_end, seed=args.seed,
                                      log=lambda s: print(s, flush=True))
    t_train = time.time() - t2
    np.save(args.out, params.astype(np.float32))
    summary = {"paths": int(pr.size), "rows": int(samples.rows), "epochs": args.epochs, "batch": args.batch,
               "steps": int(history[-1][0]) + 1, "loss_first_db2": history[0][1], "loss_last_db2": history[-1][1],
               "generation_s": t_gen, "training_s": t_train, "band_means_db": [float(m) for m in means],
               "scene": {"seed": args.seed, "scale": args.scale, "sources_sampled": args.sources,
                         "stride": args.stride}}
    print(json.dumps(summary), flush=True)
    gpu.close()


if __name__ == "__main__":
    sys.exit(main())
