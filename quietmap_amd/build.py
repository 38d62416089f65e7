"""Builds libqmpaint.so (the HIP kernels and the C ABI) in-tree for gfx950 with hipcc: each
source to an object in parallel (the sources are independent translation units), then one link.

`build(phase_clocks=True)` also builds libqmpaint_phases.so: the same library with painter.hip
compiled with QM_PHASE_CLOCKS, whose path kernels split each wave's time by phase of the path
(crossings, walk, candidates, boundary terms ...) for the profiling workload's `phases` leg."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libqmpaint.so")
PHASES_LIBRARY = os.path.join(HERE, "libqmpaint_phases.so")
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


def _link(objects, library, verbose):
    tmp = library + ".tmp"
    link = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objects], verbose)
    if link.returncode != 0:
        sys.stderr.write(link.stdout + link.stderr)
        raise RuntimeError(f"hipcc failed linking {os.path.basename(library)}")
    os.replace(tmp, library)


def build(verbose: bool = False, jobs: int | None = None, phase_clocks: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    objects = [os.path.join(OBJDIR, os.path.basename(s) + ".o") for s in SOURCES]
    painter = SOURCES[0]
    phases_object = os.path.join(OBJDIR, "painter_phases.hip.o")
    jobs_list = list(zip(SOURCES, objects, [[]] * len(SOURCES)))
    if phase_clocks:
        jobs_list.append((painter, phases_object, ["-DQM_PHASE_CLOCKS"]))
    jobs = jobs or min(len(jobs_list), max(1, min(os.cpu_count() or 1, 16)))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(lambda j: _run([HIPCC, *FLAGS, *j[2], "-c", "-o", j[1], j[0]], verbose), jobs_list))
    failed = [r for r in results if r.returncode != 0]
    if failed:
        for r in failed:
            sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed compiling libqmpaint's sources")
    _link(objects, LIBRARY, verbose)
    if phase_clocks:
        _link([phases_object] + objects[1:], PHASES_LIBRARY, verbose)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True, phase_clocks="--phase-clocks" in sys.argv))
