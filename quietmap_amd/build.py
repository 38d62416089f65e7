"""Builds libqmpaint.so (the HIP kernels and the C ABI) in-tree for gfx950 with hipcc: each
source to an object in parallel (the sources are independent translation units), then one link."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libqmpaint.so")
SOURCES = [os.path.join(CSRC, f) for f in ("painter.hip", "loudness.hip", "aircraft.hip", "surrogate.hip", "surrogate_paint.hip", "tiles.hip", "lines.hip", "pixels.hip", "horizons.hip", "flights.hip", "raster.hip", "candidates.hip")]
HEADERS = [os.path.join(CSRC, f) for f in ("qm_scene.h", "qm_aircraft.h", "qm_device.hpp", "qm_path.hpp", "qm_eval.hpp", "qm_queue.hpp", "qm_internal.hpp", "qm_features.hpp", "qm_line.hpp", "qm_zwicker.hpp", "qm_doc29.hpp")] + [
    os.path.join(os.path.dirname(HERE), "include", "qm_paint.h"),
    os.path.join(os.path.dirname(HERE), "include", "qm_surrogate.h"),
    os.path.join(os.path.dirname(HERE), "include", "qm_tiles.h"),
]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Packed FP32 instructions stay off (hand-written or compiler-made).
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    objects = [os.path.join(OBJDIR, os.path.basename(s) + ".o") for s in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, min(os.cpu_count() or 1, 16)))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(lambda so: _run([HIPCC, *FLAGS, "-c", "-o", so[1], so[0]], verbose), zip(SOURCES, objects)))
    failed = [r for r in results if r.returncode != 0]
    if failed:
        for r in failed:
            sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed compiling libqmpaint's sources")
    tmp = LIBRARY + ".tmp"
    link = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objects], verbose)
    if link.returncode != 0:
        sys.stderr.write(link.stdout + link.stderr)
        raise RuntimeError("hipcc failed linking libqmpaint.so")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
