"""Builds libqmpaint.so (the HIP kernels and the C ABI) in-tree for gfx950 with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBRARY = os.path.join(HERE, "libqmpaint.so")
SOURCES = [os.path.join(CSRC, f) for f in ("painter.hip", "loudness.hip", "aircraft.hip", "surrogate.hip", "surrogate_paint.hip", "tiles.hip", "lines.hip", "distribution.hip", "horizons.hip", "flights.hip")]
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
    "-shared",
    # Packed FP32 instructions stay off (hand-written or compiler-made).
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    tmp = LIBRARY + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError("hipcc failed building libqmpaint.so")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
