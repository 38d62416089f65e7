"""Packed pixels for the loudness kernels, as the reference's card reads them.

The reference packs a painted pixel's sound in paint/src/sone.rs (`PixelSound::pack`, `pack_line`,
lines 28-52) and its card reads it in paint-gpu/kernels/distribution.cuh (`LoudLine`,
`loud_layers`) and loudness.cuh (`pixel_ndens`): per layer (six, tiles::sources::Layer::ALL) its
steady energy per period and its line count, then each line's 43 floats (energy per period, its
mean in each of the 10 weather states per period, lambda per period, hourly profile or -1, duty
per period or NaN, passes in progress per period or NaN); then per curve (the six layers', then
the place's) and period the spectrum's 28 third-octave levels at 0 dB(A), the first NaN where the
curve has no spectrum.

`synthetic_pixels` makes seeded pixels of every kind of line the card distinguishes (Kurze lines
from sparse to dense, hourly profiles, events on a duty, streams of flight passes from pass classes)
for the parity tests and the bench.
"""
from __future__ import annotations

import numpy as np

PERIODS = 3
STATES = 10
LAYERS = 6
CURVES = LAYERS + 1
THIRDS = 28
LINE_FLOATS = PERIODS + PERIODS * STATES + PERIODS + 1 + 2 * PERIODS
assert LINE_FLOATS == 43
BIN_DB = 2.0
BINS = 90


def pack_line(energy, means, lam, profile=None, duty=None, passing=None):
    """One line's 43 floats (sone.rs pack_line)."""
    out = np.empty(LINE_FLOATS, np.float32)
    out[0:3] = energy
    out[3:33] = np.asarray(means, np.float64).reshape(PERIODS, STATES).reshape(-1)
    out[33:36] = lam
    out[36] = -1.0 if profile is None else float(profile)
    out[37:40] = np.nan if duty is None else duty
    out[40:43] = np.nan if passing is None else passing
    return out


def pack_pixel(layers, curves):
    """`layers`: six (steady[3], [packed lines]); `curves`: (7, 3, 28) third-octave levels at
    0 dB(A) (-inf silent), a curve's first level NaN for none."""
    assert len(layers) == LAYERS
    parts = []
    for steady, lines in layers:
        parts.append(np.asarray(steady, np.float32).reshape(PERIODS))
        parts.append(np.array([len(lines)], np.float32))
        parts.extend(np.asarray(line, np.float32) for line in lines)
    parts.append(np.asarray(curves, np.float32).reshape(-1))
    return np.concatenate(parts)


def concat(pixels):
    """Packed pixels one after another and their offsets (floats)."""
    offsets = np.zeros(len(pixels), np.uint64)
    at = 0
    for i, p in enumerate(pixels):
        offsets[i] = at
        at += p.size
    return np.concatenate(pixels).astype(np.float32), offsets


def _curve(rng):
    """A received spectrum's shape: third-octave levels at 0 dB(A) (a road-like tilt with noise)."""
    k = np.arange(THIRDS)
    a_weight = np.array([-44.7, -39.4, -34.6, -30.2, -26.2, -22.5, -19.1, -16.1, -13.4, -10.9, -8.6, -6.6, -4.8,
                         -3.2, -1.9, -0.8, 0.0, 0.6, 1.0, 1.2, 1.3, 1.2, 1.0, 0.5, -0.1, -1.1, -2.5, -4.3])
    raw = -0.4 * np.abs(k - rng.uniform(10, 18)) + rng.normal(0, 1.5, THIRDS)
    a = np.sum(10 ** ((raw + a_weight) / 10))
    return raw - 10 * np.log10(a)


def synthetic_pixels(seed=7, count=64, lines_max=6, kurze=True):
    """Seeded pixels covering the card's kinds of line (without Kurze lines when `kurze` is false:
    events and passes alone, whose values keep their mean exactly); the list of packed pixels."""
    rng = np.random.default_rng(seed)
    pixels = []
    for _ in range(count):
        layers = []
        any_profile = rng.uniform() < 0.5
        for layer in range(LAYERS):
            if rng.uniform() < 0.3:
                layers.append((np.zeros(3), []))
                continue
            level = rng.uniform(30, 70)
            steady = 10 ** ((level - rng.uniform(0, 20)) / 10) * np.array([1.0, 0.6, 0.3])
            lines = []
            for _ in range(rng.integers(0, lines_max + 1)):
                e = 10 ** ((level + rng.normal(0, 6)) / 10) * np.array([1.0, rng.uniform(0.3, 1), rng.uniform(0.05, 1)])
                spread = 10 ** (rng.normal(0, 1.0, (PERIODS, STATES)) / 10)
                means = e[:, None] * spread / spread.mean(axis=1, keepdims=True)
                if rng.uniform() < 0.3:
                    means[:, 5:] = means[:, 4:5]  # equal neighbouring states
                kind = rng.uniform()
                if not kurze:
                    kind = 0.55 + 0.45 * kind
                if kind < 0.55:
                    lam = 10 ** rng.uniform(-5.5, 1.7, 3)
                    profile = int(rng.integers(0, 3)) if any_profile and rng.uniform() < 0.6 else None
                    lines.append(pack_line(e, means, lam, profile=profile))
                elif kind < 0.75:
                    duty = rng.uniform(0.001, 0.6, 3)
                    duty[rng.uniform(size=3) < 0.2] = 0.0
                    lines.append(pack_line(e, means, [np.inf] * 3, duty=duty))
                else:
                    passing = 10 ** rng.uniform(-4, 2.2, 3)
                    lines.append(pack_line(e, np.repeat(e[:, None], STATES, 1), [np.inf] * 3, passing=passing))
            layers.append((steady, lines))
        curves = np.stack([np.stack([_curve(rng) for _ in range(PERIODS)]) for _ in range(CURVES)])
        for c in range(CURVES):
            if rng.uniform() < 0.1:
                curves[c, :, 0] = np.nan
        pixels.append(pack_pixel(layers, curves))
    return pixels


def pack_line_f64(energy, means, lam, profile=None, duty=None, passing=None):
    """One unfolded line's 43 values in f64 (the packed layout before the map's folding)."""
    out = np.empty(LINE_FLOATS, np.float64)
    out[0:3] = energy
    out[3:33] = np.asarray(means, np.float64).reshape(-1)
    out[33:36] = lam
    out[36] = -1.0 if profile is None else float(profile)
    out[37:40] = np.nan if duty is None else duty
    out[40:43] = np.nan if passing is None else passing
    return out


def pack_pixel_f64(layers, curves):
    """One unfolded pixel (f64): as pack_pixel, the lines as placed (pack_line_f64)."""
    assert len(layers) == LAYERS
    parts = []
    for steady, lines in layers:
        parts.append(np.asarray(steady, np.float64).reshape(PERIODS))
        parts.append(np.array([len(lines)], np.float64))
        parts.extend(np.asarray(line, np.float64) for line in lines)
    parts.append(np.asarray(curves, np.float64).reshape(-1))
    return np.concatenate(parts)


def concat_f64(pixels):
    """Unfolded pixels one after another (f64) and their offsets (values)."""
    offsets = np.zeros(len(pixels), np.uint64)
    at = 0
    for i, p in enumerate(pixels):
        offsets[i] = at
        at += p.size
    return np.concatenate(pixels).astype(np.float64), offsets


def synthetic_unfolded(seed=7, count=64, lines_max=8):
    """Seeded unfolded pixels (f64) with every kind of line the map's folding sorts: weak and loud
    events on a duty, sparse Kurze lines (held, moving, following hourly profiles), steady lines
    whose weather states differ by less than the merge's spread, streams of passes that hold and
    that move. Zero energy in a period goes with zero means there. Returns (values, offsets)."""
    rng = np.random.default_rng(seed)
    pixels = []
    for _ in range(count):
        parts = []
        any_profile = rng.uniform() < 0.5
        for layer in range(LAYERS):
            if rng.uniform() < 0.25:
                parts.append(np.array([0.0, 0.0, 0.0, 0.0]))
                continue
            level = rng.uniform(30, 70)
            steady = 10 ** ((level - rng.uniform(-5, 20)) / 10) * np.array([1.0, 0.6, 0.3])
            lines = []
            for _ in range(rng.integers(0, lines_max + 1)):
                e = 10 ** ((level + rng.normal(-6, 8)) / 10) * np.array([1.0, rng.uniform(0.3, 1), rng.uniform(0.05, 1)])
                if rng.uniform() < 0.15:
                    e[rng.integers(0, 3)] = 0.0
                if rng.uniform() < 0.4:
                    spread = 10 ** (rng.uniform(-0.05, 0.05, (PERIODS, STATES)))  # within the merge spread
                else:
                    spread = 10 ** (rng.normal(0, 0.25, (PERIODS, STATES)))
                means = e[:, None] * spread / spread.mean(axis=1, keepdims=True)
                if rng.uniform() < 0.3:
                    means[:, 5:] = means[:, 4:5]
                kind = rng.uniform()
                if kind < 0.3:
                    lam = 10 ** rng.uniform(-5.5, 1.7, 3)
                    profile = int(rng.integers(0, 3)) if any_profile and rng.uniform() < 0.6 else None
                    lines.append(pack_line_f64(e, means, lam, profile=profile))
                elif kind < 0.45:
                    lines.append(pack_line_f64(e, means, [np.inf] * 3))
                elif kind < 0.7:
                    duty = rng.uniform(0.001, 0.6, 3)
                    duty[rng.uniform(size=3) < 0.2] = 0.0
                    lines.append(pack_line_f64(e * rng.uniform(0.001, 1.0), means * rng.uniform(0.001, 0.05), [np.inf] * 3,
                                               duty=duty))
                else:
                    passing = 10 ** rng.uniform(-4, 2.2, 3)
                    lines.append(pack_line_f64(e, np.repeat(e[:, None], STATES, 1), [np.inf] * 3, passing=passing))
            parts.append(np.concatenate([steady, [float(len(lines))]] + lines))
        curves = np.stack([np.stack([_curve(rng) for _ in range(PERIODS)]) for _ in range(CURVES)])
        for c in range(CURVES):
            if rng.uniform() < 0.1:
                curves[c, :, 0] = np.nan
        parts.append(curves.reshape(-1))
        pixels.append(np.concatenate(parts))
    return concat_f64(pixels)
