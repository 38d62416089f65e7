"""The painter's batch (the reference's `paint::exact::Batch`, paint/src/exact.rs:163-183, as
`GpuBatch` implements it in paint-gpu/src/batch.rs) over libqmpaint: what each candidate delivers
at each point of a pass, on the card.

- ``evaluate(points, pairs)``: every pair's energy (mixed per period, calm and downwind alone) and
  every point's bands per layer, period and band, the pairs' band weights applied to the bands
  (exact.rs:208-216 add_bands). Candidates are the square's point sources and line pieces alike.
  Pairs go to the card in order of candidate (batch.rs:73-90 by_candidate), at most
  ``PAIRS_PER_LAUNCH`` a launch, each launch's points once (batch.rs:111-160
  LaunchInputs::gather), the bands summed on the card at each reached point's slot. A pair the card refuses (failure bits) is listed in
  ``Evaluated.refused`` with its bits: the reference evaluates those on its CPU path
  (batch.rs:292-317), which stays the caller's.
- ``pair_spectra(points, pairs)``: each pair's unweighted energy and bands (hum_pairs.rs).
- ``loudness(input, offsets)``: Nden per layer and in total of at most ``PIXELS_PER_LOUDNESS_BATCH``
  packed pixels (paint-gpu/src/loudness.rs).
- ``flight_sums(boxes, points, horizons, segments, n_sums)``: the aircraft boxes' sums
  (paint-gpu/src/aircraft.rs:222-350), energies and pass classes per slot.

Errors raise ``PainterError`` (the reference's ``Err(String)``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import abi
from .painter import PainterError

PAIRS_PER_LAUNCH = 2_000_000
PIXELS_PER_LOUDNESS_BATCH = 8192

# paint/src/exact.rs:18-34 Point (position, the building it stands in) and :143-149 Pair.
PAIR_DTYPE = np.dtype([("point", "<u4"), ("candidate", "<u4"), ("band_weight", "<f8")])


def pairs_of(point, candidate, band_weight=None):
    """Pairs from parallel arrays (band weight 1 when None)."""
    p = np.zeros(np.size(point), PAIR_DTYPE)
    p["point"], p["candidate"] = point, candidate
    p["band_weight"] = 1.0 if band_weight is None else band_weight
    return p


def by_candidate(pairs, candidates):
    """batch.rs:73-90: the pairs' indices by candidate, each candidate's in their given order."""
    c = np.asarray(pairs["candidate"], np.int64)
    if c.size and (c.min() < 0 or c.max() >= candidates):
        raise PainterError("pair candidate out of range")
    return np.argsort(c, kind="stable").astype(np.uint32)


@dataclass
class Evaluated:
    """exact.rs:138-141 Evaluated: per pair mixed[3] and states[2][3] (f64), per point its bands
    [layers][periods][bands]; and the pairs the card refused with their failure bits."""

    mixed: np.ndarray
    states: np.ndarray
    bands: np.ndarray
    refused: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    refused_bits: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))


class GpuBatch:
    """batch.rs GpuBatch over one uploaded square (``painter.Square``) and its candidates: its
    point sources, then its line pieces (``Square.set_lines``)."""

    def __init__(self, gpu, square, pairs_per_launch=PAIRS_PER_LAUNCH):
        self.gpu, self.square = gpu, square
        self.candidates = int(square.candidates)
        self.pairs_per_launch = int(pairs_per_launch)

    def _launches(self, points, pairs, bands):
        """Yields (launch pair indices, the card's results) for the pairs in candidate order, each
        launch's receivers its points once in the order its pairs first reach them
        (LaunchInputs::gather) and each receiver's bands added at its point's slot in `bands`
        (batch.rs:251-261: every point a pair reaches, once, in that order)."""
        order = by_candidate(pairs, self.candidates)
        slot_of = np.full(points.size, np.iinfo(np.uint32).max, np.uint32)
        if bands is not None:
            unique, first = np.unique(pairs["point"], return_index=True)
            reached = unique[np.argsort(first, kind="stable")]
            slot_of[reached] = np.arange(reached.size, dtype=np.uint32)
        for at in range(0, order.size, self.pairs_per_launch):
            launch = order[at : at + self.pairs_per_launch]
            pts = pairs["point"][launch]
            unique, first = np.unique(pts, return_index=True)
            receivers = unique[np.argsort(first, kind="stable")]
            slot = np.empty(points.size, np.uint32)
            slot[receivers] = np.arange(receivers.size, dtype=np.uint32)
            w = pairs["band_weight"][launch]
            weights = None if np.all(w == 1.0) else w.astype(np.float32)
            out = self.gpu.evaluate_candidates(self.square, points[receivers], pairs["candidate"][launch], slot[pts],
                                               point_of_receiver=slot_of[receivers] if bands is not None else None,
                                               band_weight=weights, bands=bands, pair_spectra=bands is None)
            yield launch, out

    def _checked(self, points, pairs):
        points = np.ascontiguousarray(points, abi.POINT_DTYPE)
        pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
        if pairs.size and int(pairs["point"].max()) >= points.size:
            raise PainterError("pair point out of range")
        return points, pairs

    def evaluate(self, points, pairs) -> Evaluated:
        points, pairs = self._checked(points, pairs)
        n = pairs.size
        mixed, states = np.zeros((n, abi.PERIODS)), np.zeros((n, 2, abi.PERIODS))
        unique, first = np.unique(pairs["point"], return_index=True)
        reached = unique[np.argsort(first, kind="stable")]
        flat = np.zeros((reached.size, abi.LAYERS, abi.PERIODS, abi.BANDS), np.float32)
        refused, bits = [], []
        for launch, out in self._launches(points, pairs, flat):
            sums = out["sums"].astype(np.float64)
            mixed[launch] = sums[:, 0:3]
            states[launch, 0] = sums[:, 3:6]
            states[launch, 1] = sums[:, 6:9]
            bad = out["failed"] != 0
            refused.append(launch[bad].astype(np.int64))
            bits.append(out["failed"][bad])
        bands = np.zeros((points.size, abi.LAYERS, abi.PERIODS, abi.BANDS))
        bands[reached] = flat.astype(np.float64)
        refused = np.concatenate(refused) if refused else np.zeros(0, np.int64)
        bits = np.concatenate(bits) if bits else np.zeros(0, np.uint32)
        order = np.argsort(refused, kind="stable")
        return Evaluated(mixed, states, bands, refused[order], bits[order])

    def pair_spectra(self, points, pairs):
        """Each pair's unweighted energy (mixed[3], states[2][3]) and its bands [3][8]; refused
        pairs (failure bits) zero, listed as in ``evaluate``."""
        points, pairs = self._checked(points, pairs)
        n = pairs.size
        mixed, states = np.zeros((n, abi.PERIODS)), np.zeros((n, 2, abi.PERIODS))
        spectra = np.zeros((n, abi.PERIODS, abi.BANDS))
        refused, bits = [], []
        for launch, out in self._launches(points, pairs, None):
            sums = out["sums"].astype(np.float64)
            mixed[launch] = sums[:, 0:3]
            states[launch, 0] = sums[:, 3:6]
            states[launch, 1] = sums[:, 6:9]
            spectra[launch] = out["pair_spectra"].astype(np.float64)
            bad = out["failed"] != 0
            refused.append(launch[bad].astype(np.int64))
            bits.append(out["failed"][bad])
        refused = np.concatenate(refused) if refused else np.zeros(0, np.int64)
        bits = np.concatenate(bits) if bits else np.zeros(0, np.uint32)
        order = np.argsort(refused, kind="stable")
        return Evaluated(mixed, states, spectra, refused[order], bits[order])

    def loudness(self, input, offsets):
        """Nden per layer and in total [pixels, 7] of at most PIXELS_PER_LOUDNESS_BATCH pixels."""
        if np.size(offsets) > PIXELS_PER_LOUDNESS_BATCH:
            raise PainterError(f"{np.size(offsets)} pixels in one loudness batch (at most {PIXELS_PER_LOUDNESS_BATCH})")
        return self.gpu.pixel_loudness(input, offsets)

    def flight_sums(self, boxes, points, horizons, segments, n_sums):
        """Per sum slot its energies per period and its pass classes (energies [3][48], passing
        shares [3][48]): a dict of arrays."""
        sums = self.gpu.flight_sums(boxes, points, horizons, segments, n_sums)
        return dict(energy=sums[:, :3], class_energy=sums[:, 3 : 3 + 144].reshape(-1, 3, abi.PASS_CLASSES),
                    class_passing=sums[:, 3 + 144 :].reshape(-1, 3, abi.PASS_CLASSES))
