"""The hum on the card: each hum point's quiet ground sources importance-sampled and evaluated
through the batch's pair spectra (``GpuBatch.pair_spectra``), the loud ones exactly.

The reference's host drives this (paint/src/hum.rs:61-119 ``Hum::new``, paint/src/hum_samples.rs:53-219
``sampled_hum``) over ``popup::sampling`` (popup/src/sampling.rs: ``BoundSample``, ``SampleRound``)
and ``physics::bound`` (physics/src/bound.rs: ``received_energy_bound``); this is its mirror over
the card's batch, with the same rules:

- a source is loud at a point where its bound (free-field divergence at its nearest distance, the
  least air absorption, the largest CNOSSOS gain at the point's largest favourable share, the
  reflection) reaches 20 dB Lden (popup/src/candidates.rs:234 ``loud``); the loud ones are evaluated
  exactly (``GpuBatch.evaluate``), their energies summed per point and layer;
- per point and layer, the quiet sources within 12 km are left out when their bounds stay under
  the omitted fraction of the loud answer in every period, else sampled: sources ordered by their
  bound's Lden energy (largest first, ties by the larger index), those whose bound reaches the
  rest's sum over the sample size taken with certainty, the rest drawn with replacement in
  proportion to their bounds (SplitMix64 seeded by the point and the layer), the sample doubling
  from 64 draws until two standard errors fall under 0.05 dB of the answer, or under 15 dB Lden
  together (``unheard``), or the population is exhausted, or 32,768 draws;
- each round's new pairs go to the card in one ``pair_spectra`` call (at most 24,000,000 pairs a
  call); the estimate's importance weights apply to the pairs' bands, which are then scaled per
  period to the estimated energy.

The sampling arithmetic keeps the reference's order of operations (sequential sums, the same
SplitMix64 stream and inverse-CDF draw), so a sample draws the same sources; refused pairs (the
card's failure bits) go to ``on_refused`` — the reference's CPU path, which stays the caller's.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import abi
from .painter import PainterError

PERIODS, BANDS, LAYERS = abi.PERIODS, abi.BANDS, abi.LAYERS

# physics/src/bands.rs:16-18, 59-66
PERIOD_HOURS = (12.0, 4.0, 8.0)
PERIOD_PENALTY_DB = (0.0, 5.0, 10.0)
# physics/src/bound.rs:12-25; physics/src/line.rs:34-36
GAIN_BOUND_DB = 18.0
POINT_DIVERGENCE_OFFSET_DB = 11.0
POINT_DIVERGENCE_LINEAR = 12.589_254_117_941_673
POINT_DISTANCE_FLOOR_M = 1.0
LINE_PERPENDICULAR_FLOOR_M = 0.5
REACH_EDGE_LDEN_DB = 20.0
# popup/src/candidates.rs:23, popup/src/selection.rs:26, popup/src/sampling.rs:9
GROUND_REACH_M = 12_000.0
OMITTED_ENERGY_FRACTION = 0.023_292_992_280_754_13
SAMPLING_TOLERANCE = 0.011_579_454_376_269_092
# paint/src/hum_samples.rs:22, 31; paint/src/paint.rs:85
FIRST_DRAWS = 64
UNHEARD_LDEN_ENERGY = 31.622_776_601_683_793
PASS_PAIRS = 24_000_000
MAX_DRAWS = 32_768
GOLDEN = 0x9E37_79B9_7F4A_7C15
MASK = (1 << 64) - 1


def energy(level_db):
    """10^(L/10), physics/src/bands.rs energy."""
    return 10.0 ** (np.asarray(level_db, np.float64) / 10.0)


def lden_energy(period_energy) -> float:
    """The Lden energy of per-period energies (physics/src/bands.rs:59-66), summed in period order."""
    total = 0.0
    for p in range(PERIODS):
        total += PERIOD_HOURS[p] * float(period_energy[p]) * (10.0 ** (PERIOD_PENALTY_DB[p] / 10.0))
    return total / 24.0


def mixed_gain_bound_db(p: float) -> float:
    """physics/src/bound.rs:29-32: the largest mixed gain at favourable share p (both states 18 dB)."""
    g = 10.0 ** (GAIN_BOUND_DB / 10.0)
    return 10.0 * np.log10(p * g + (1.0 - p) * g)


@dataclass
class ReceiverBound:
    """physics/src/bound.rs:85-101: per period the largest gain, per band the place's air."""

    gain: np.ndarray  # [PERIODS]
    alpha_db_per_km: np.ndarray  # [BANDS]


def receiver_bound(favourable_max, reflection_db, alpha_db_per_km) -> ReceiverBound:
    """physics/src/bound.rs:104-114."""
    gain = np.array([10.0 ** ((mixed_gain_bound_db(float(p)) + reflection_db) / 10.0) for p in favourable_max])
    return ReceiverBound(gain, np.asarray(alpha_db_per_km, np.float64))


def receiver_bounds(receivers, reflection_db=None):
    """The bound of each receiver record (abi.RECEIVER_DTYPE, the card's receiver properties):
    its largest favourable share per period over the 16 sectors and its air absorption."""
    reflection = np.zeros(receivers.size) if reflection_db is None else np.asarray(reflection_db, np.float64)
    fav = receivers["favourable"].astype(np.float64).max(axis=2)
    alpha = receivers["alpha_db_per_km"].astype(np.float64)
    return [receiver_bound(fav[i], float(reflection[i]), alpha[i]) for i in range(receivers.size)]


def divergence_db(length_m: float, distance_m: float) -> float:
    """physics/src/bound.rs:45-58 Spread::divergence_db: a point (length 0) or a straight piece."""
    if length_m > 0.0:
        d = max(distance_m, LINE_PERPENDICULAR_FLOOR_M)
        infinite_line = np.pi / d
        whole_power = length_m / (d * d)
        return -10.0 * np.log10(min(infinite_line, whole_power) / POINT_DIVERGENCE_LINEAR)
    return 20.0 * np.log10(max(distance_m, POINT_DISTANCE_FLOOR_M)) + POINT_DIVERGENCE_OFFSET_DB


def received_energy_bound(emission_energy, length_m: float, distance_m: float, rb: ReceiverBound) -> np.ndarray:
    """physics/src/bound.rs:118-132: the upper bound of the received A-weighted energy per period."""
    divergence = 10.0 ** (-divergence_db(length_m, distance_m) / 10.0)
    air = [10.0 ** ((-float(rb.alpha_db_per_km[b]) * distance_m / 1000.0) / 10.0) for b in range(BANDS)]
    out = np.empty(PERIODS)
    for p in range(PERIODS):
        received = 0.0
        for b in range(BANDS):
            received += float(emission_energy[p][b]) * air[b]
        out[p] = received * divergence * float(rb.gain[p])
    return out


def loud(bound) -> bool:
    """popup/src/candidates.rs:234-236."""
    return lden_energy(bound) >= 10.0 ** (REACH_EDGE_LDEN_DB / 10.0)


def unheard(variance) -> bool:
    """paint/src/hum_samples.rs:33-35: two standard errors of every period together under 15 dB Lden."""
    return lden_energy([2.0 * np.sqrt(v) for v in variance]) <= UNHEARD_LDEN_ENERGY


# ---------------------------------------------------------------------------------------------
# popup/src/selection.rs:296-313 SplitMix, popup/src/sampling.rs BoundSample and SampleRound.


def splitmix_units(seed: int, count: int) -> np.ndarray:
    """The first `count` uniforms in [0, 1) of SplitMix64 from `seed` (the state advancing by the
    golden gamma before each draw)."""
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK) + k * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58_476D_1CE4_E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D0_49BB_1331_11EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def sequential_sum(values) -> float:
    """A left-to-right sum (the reference's iterator sums), not numpy's pairwise one."""
    v = np.asarray(values, np.float64)
    return float(np.cumsum(v)[-1]) if v.size else 0.0


@dataclass
class SampleRound:
    certain: range
    draws: np.ndarray
    exhaustive: bool
    population: int
    rest_order: float

    def estimate(self, orders, energies):
        """sampling.rs:111-146: the mean and variance per period of the draws' energies over their
        probabilities; `energies[i]` is source i's energy per period."""
        count = len(self.draws)
        if count == 0:
            return np.zeros(PERIODS), np.zeros(PERIODS)
        values = np.array([energies[int(i)] for i in self.draws], np.float64)  # [count, PERIODS]
        if self.exhaustive:
            probability = np.ones(count)
            scale = 1.0
        else:
            probability = np.asarray(orders, np.float64)[self.draws] / self.rest_order
            scale = 1.0 / count
        v = values / probability[:, None]
        mean = np.cumsum(v * scale, axis=0)[-1]
        square = np.cumsum(v * v * scale, axis=0)[-1]
        if self.exhaustive:
            variance = np.zeros(PERIODS)
        else:
            variance = np.maximum(square - mean * mean, 0.0) * count / (count - 1.0) / count
        return mean, variance

    def weights(self, orders) -> dict:
        """sampling.rs:148-160: each drawn source's weight (its draws over their expectation)."""
        unique, counts = np.unique(self.draws, return_counts=True)
        out = {}
        for i, c in zip(unique.tolist(), counts.tolist()):
            w = float(c)
            if not self.exhaustive:
                w /= len(self.draws) * float(orders[i]) / self.rest_order
            out[i] = w
        return out


class BoundSample:
    """sampling.rs:11-109: rounds of bound-proportional draws, doubling until confident."""

    def __init__(self, seed: int, size: int = 512):
        if size < 2:
            raise ValueError("a sample's variance takes two draws")
        self.first, self.size, self.seed = 0, int(size), int(seed) & MASK

    def round(self, orders) -> SampleRound:
        orders = np.asarray(orders, np.float64)
        rest_order = sequential_sum(orders[self.first:])
        certain = self.first
        while certain < orders.size and orders[certain] * self.size >= rest_order:
            rest_order -= orders[certain]
            certain += 1
        population = orders.size - certain
        exhaustive = population <= self.size
        if exhaustive:
            draws = np.arange(certain, orders.size, dtype=np.int64)
        else:
            cumulative = np.cumsum(orders[certain:])
            total = float(cumulative[-1])
            targets = splitmix_units(self.seed ^ self.size, self.size) * total
            at = np.searchsorted(cumulative, targets, side="right")
            draws = certain + np.minimum(at, population - 1).astype(np.int64)
        return SampleRound(range(self.first, certain), draws, exhaustive, population, rest_order)

    def converged(self, rnd: SampleRound, variance, answer) -> bool:
        return (rnd.exhaustive or self.size >= MAX_DRAWS or
                all(2.0 * np.sqrt(variance[p]) <= SAMPLING_TOLERANCE * answer[p] for p in range(PERIODS)))

    def advance(self, rnd: SampleRound):
        self.first = rnd.certain.stop
        self.size *= 2


# ---------------------------------------------------------------------------------------------
# The square's candidates as the hum reads them (popup/src/candidates.rs:138-193).


@dataclass
class HumCandidates:
    """Per candidate (the batch's numbering: point sources, then line pieces): its ends (a point's
    twice), whether it is a line, its length, layer and A-weighted emission energy per period and
    band (per metre for a line)."""

    ends: np.ndarray  # [n, 2, 2]
    line: np.ndarray  # [n] bool
    length_m: np.ndarray  # [n]
    layer: np.ndarray  # [n]
    emission: np.ndarray  # [n, PERIODS, BANDS]

    @classmethod
    def of_square(cls, square):
        s, l = square.sources, square.lines
        ends = np.concatenate([np.stack([np.stack([s["x_m"], s["y_m"]], 1)] * 2, 1),
                               np.stack([np.stack([l["ax_m"], l["ay_m"]], 1), np.stack([l["bx_m"], l["by_m"]], 1)], 1)])
        line = np.concatenate([np.zeros(s.size, bool), np.ones(l.size, bool)])
        length = np.where(line, np.hypot(ends[:, 1, 0] - ends[:, 0, 0], ends[:, 1, 1] - ends[:, 0, 1]), 0.0)
        layer = np.concatenate([s["layer"], l["layer"]]).astype(np.int64)
        emission = np.concatenate([s["energy"], l["energy_per_m"]]).astype(np.float64)
        return cls(ends, line, length, layer, emission)

    def offsets(self, position):
        """Offset of each candidate's nearest point from `position` (dx, dy), candidates.rs:188-193
        and 120-134 (a piece's nearest point, its fraction clamped to the piece)."""
        a = self.ends[:, 0] - np.asarray(position, np.float64)
        b = self.ends[:, 1] - np.asarray(position, np.float64)
        d = b - a
        length_sq = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
        with np.errstate(invalid="ignore", divide="ignore"):
            t = np.where(length_sq > 0.0, np.clip(-(a[:, 0] * d[:, 0] + a[:, 1] * d[:, 1]) / length_sq, 0.0, 1.0), 0.0)
        closest = np.stack([a[:, 0] + t * d[:, 0], a[:, 1] + t * d[:, 1]], 1)
        return np.where(self.line[:, None], closest, a)


def reached(candidates: HumCandidates, position):
    """The candidates within GROUND_REACH_M of a point and their distances (hum_samples.rs:66-80:
    the squared test with its rounding allowance, then the distance)."""
    off = candidates.offsets(position)
    sq = off[:, 0] * off[:, 0] + off[:, 1] * off[:, 1]
    reach_sq = GROUND_REACH_M * GROUND_REACH_M * (1.0 + 16.0 * np.finfo(np.float64).eps)
    near = np.flatnonzero(sq <= reach_sq)
    dist = np.hypot(off[near, 0], off[near, 1])
    keep = dist <= GROUND_REACH_M
    return near[keep], dist[keep]


# ---------------------------------------------------------------------------------------------


@dataclass
class _Sample:
    point: int
    layer: int
    indices: np.ndarray  # candidate per position in the order
    orders: np.ndarray
    sample: BoundSample
    evaluated: dict  # order position -> (mixed[3], bands[3][8])
    certain_energy: np.ndarray
    finished: SampleRound | None = None


def _spectra(batch, points, pairs, on_refused):
    out = batch.pair_spectra(points, pairs)
    mixed, bands = out.mixed, out.bands
    if out.refused.size:
        if on_refused is None:
            raise PainterError(f"{out.refused.size} pairs refused by the card (failure bits "
                               f"{sorted(set(int(b) for b in out.refused_bits))}): pass on_refused, the CPU path")
        m, b = on_refused(points, pairs[out.refused])
        mixed[out.refused], bands[out.refused] = m, b
    return mixed, bands


def sampled_hum(batch, points, candidates: HumCandidates, bounds, loud_energy, seeds, on_refused=None):
    """paint/src/hum_samples.rs:53-219 over the batch: per point and layer the quiet energy per
    period [n, LAYERS, PERIODS] and its bands [n, LAYERS, PERIODS, BANDS]."""
    from .batch import pairs_of

    n = points.size
    samples = []
    for at in range(n):
        near, dist = reached(candidates, (points["x_m"][at], points["y_m"][at]))
        pending = [[] for _ in range(LAYERS)]
        layer_bounds = np.zeros((LAYERS, PERIODS))
        for index, d in zip(near.tolist(), dist.tolist()):
            bound = received_energy_bound(candidates.emission[index], float(candidates.length_m[index]), d, bounds[at])
            if not loud(bound):
                layer = int(candidates.layer[index])
                pending[layer].append((index, lden_energy(bound)))
                layer_bounds[layer] += bound
        for layer in range(LAYERS):
            if not pending[layer] or all(layer_bounds[layer][p] <= OMITTED_ENERGY_FRACTION * loud_energy[at][layer][p]
                                         for p in range(PERIODS)):
                continue
            # Largest order first, ties by the larger index.
            pending[layer].sort(key=lambda e: (-e[1], -e[0]))
            indices = np.array([e[0] for e in pending[layer]], np.int64)
            orders = np.array([e[1] for e in pending[layer]], np.float64)
            seed = (int(seeds[at]) ^ ((layer * GOLDEN) & MASK)) & MASK
            samples.append(_Sample(at, layer, indices, orders, BoundSample(seed, FIRST_DRAWS), {}, np.zeros(PERIODS)))
    quiet = np.zeros((n, LAYERS, PERIODS))
    quiet_bands = np.zeros((n, LAYERS, PERIODS, BANDS))
    while True:
        rounds = [(k, s.sample.round(s.orders)) for k, s in enumerate(samples) if s.finished is None]
        if not rounds:
            break
        needed = []
        for k, rnd in rounds:
            s = samples[k]
            want = set(rnd.certain) | set(rnd.draws.tolist())
            needed.extend((k, i) for i in sorted(want) if i not in s.evaluated)
        for c in range(0, len(needed), PASS_PAIRS):
            chunk = needed[c:c + PASS_PAIRS]
            pairs = pairs_of(np.array([samples[k].point for k, _ in chunk], np.uint32),
                             np.array([samples[k].indices[i] for k, i in chunk], np.uint32), np.zeros(len(chunk)))
            mixed, bands = _spectra(batch, points, pairs, on_refused)
            for j, (k, i) in enumerate(chunk):
                samples[k].evaluated[i] = (mixed[j], bands[j])
        for k, rnd in rounds:
            s = samples[k]
            for i in rnd.certain:
                s.certain_energy = s.certain_energy + s.evaluated[i][0]
            est, var = rnd.estimate(s.orders, {i: s.evaluated[i][0] for i in set(rnd.draws.tolist())})
            answer = loud_energy[s.point][s.layer] + s.certain_energy + est
            if unheard(var) or s.sample.converged(rnd, var, answer):
                quiet[s.point, s.layer] = s.certain_energy + est
                s.finished = rnd
            else:
                s.sample.advance(rnd)
    for s in samples:
        rnd = s.finished
        weighted = [(i, 1.0) for i in range(rnd.certain.stop)] + list(rnd.weights(s.orders).items())
        for i, w in weighted:
            quiet_bands[s.point, s.layer] += w * s.evaluated[i][1]
    # Bands scaled per period to the energy (hum_samples.rs:204-216).
    sums = np.cumsum(quiet_bands, axis=3)[..., -1]
    with np.errstate(invalid="ignore", divide="ignore"):
        scale = np.where(sums > 0.0, quiet / np.where(sums > 0.0, sums, 1.0), 0.0)
    return quiet, quiet_bands * scale[..., None]


def loud_pairs(points, candidates: HumCandidates, bounds):
    """paint/src/exact.rs:256-274 loud_pairs: per point the candidates loud there, point-major."""
    from .batch import pairs_of

    pp, pc = [], []
    for at in range(points.size):
        near, dist = reached(candidates, (points["x_m"][at], points["y_m"][at]))
        for index, d in zip(near.tolist(), dist.tolist()):
            if loud(received_energy_bound(candidates.emission[index], float(candidates.length_m[index]), d, bounds[at])):
                pp.append(at)
                pc.append(index)
    return pairs_of(np.array(pp, np.uint32), np.array(pc, np.uint32))


@dataclass
class HumPoints:
    """The hum at its points: the loud sources' energy, the quiet sources' sampled energy and
    bands, per point, layer and period (paint/src/hum.rs:53-59)."""

    loud: np.ndarray  # [n, LAYERS, PERIODS]
    quiet: np.ndarray  # [n, LAYERS, PERIODS]
    quiet_bands: np.ndarray  # [n, LAYERS, PERIODS, BANDS]


def hum_points(batch, points, candidates: HumCandidates, bounds, seeds, on_refused=None) -> HumPoints:
    """paint/src/hum.rs:61-119 Hum::new at the given points: the loud pairs exactly on the card,
    then the quiet sources sampled (``sampled_hum``)."""
    points = np.ascontiguousarray(points, abi.POINT_DTYPE)
    pairs = loud_pairs(points, candidates, bounds)
    loud_energy = np.zeros((points.size, LAYERS, PERIODS))
    if pairs.size:
        out = batch.evaluate(points, pairs)
        mixed = out.mixed
        if out.refused.size:
            if on_refused is None:
                raise PainterError(f"{out.refused.size} loud pairs refused by the card: pass on_refused, the CPU path")
            mixed[out.refused] = on_refused(points, pairs[out.refused])[0]
        layers = candidates.layer[pairs["candidate"].astype(np.int64)]
        for j in range(pairs.size):
            loud_energy[pairs["point"][j], layers[j]] += mixed[j]
    quiet, bands = sampled_hum(batch, points, candidates, bounds, loud_energy, seeds, on_refused)
    return HumPoints(loud_energy, quiet, bands)
