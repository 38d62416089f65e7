"""The hum's host sampler (quietmap_amd/hum.py) on the CPU: the reference's own checks of its
sampling rounds, bounds and budget (popup/src/sampling.rs tests, physics/src/bound.rs tests,
paint/src/hum_samples.rs tests), SplitMix64's published first outputs, and whole samples of a
stand-in batch against the exact sums they estimate."""
import math

import numpy as np
import pytest

from quietmap_amd import abi, hum
from quietmap_amd.batch import Evaluated


def test_certain_and_drawn_pieces_keep_energy_and_silent_periods():
    # popup/src/sampling.rs certainty_and_random_draws_preserve_energy_and_silent_periods
    orders = np.ones(1025)
    orders[0] = 20.0
    sample = hum.BoundSample(3)
    rnd = sample.round(orders)
    assert rnd.certain == range(0, 1)
    assert not rnd.exhaustive
    energies = {i: np.array([2.0, 0.0, 4.0]) for i in range(1025)}
    est, var = rnd.estimate(orders, energies)
    assert est.tolist() == [2048.0, 0.0, 4096.0]
    assert var.tolist() == [0.0, 0.0, 0.0]
    assert sample.converged(rnd, var, est)
    assert sum(w * 2.0 for w in rnd.weights(orders).values()) == est[0]


def test_a_sample_starts_with_its_callers_draws():
    # popup/src/sampling.rs a_sample_starts_with_its_callers_draws
    orders = np.ones(1025)
    assert len(hum.BoundSample(3, 64).round(orders).draws) == 64
    assert len(hum.BoundSample(3).round(orders).draws) == 512
    with pytest.raises(ValueError):
        hum.BoundSample(3, 1)


def test_a_hums_errors_are_unheard_together_in_lden():
    # paint/src/hum_samples.rs a_hums_errors_are_unheard_together_in_lden
    def of(two_errors):
        return [(e / 2.0) ** 2 for e in two_errors]

    assert hum.unheard(of([20.0, 20.0, 3.0]))
    assert not hum.unheard(of([20.0, 20.0, 4.0]))


def splitmix_scalar(seed, count):
    state, out = seed, []
    for _ in range(count):
        state = (state + hum.GOLDEN) & hum.MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & hum.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & hum.MASK
        out.append(z ^ (z >> 31))
    return out


def test_splitmix_stream():
    # SplitMix64's published first outputs from state 0 (Steele, Lea and Flood 2014; the
    # reference generator of xoshiro's seeding).
    first = splitmix_scalar(0, 3)
    assert first == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    for seed in (0, 3 ^ 512, 0xFFFFFFFFFFFFFFFF, 123456789):
        want = np.array([(z >> 11) / float(1 << 53) for z in splitmix_scalar(seed, 200)])
        assert np.array_equal(hum.splitmix_units(seed, 200), want)


def test_bounds():
    # physics/src/bound.rs mixed_gain_spans_the_two_state_bounds
    for p in (0.0, 0.5, 1.0):
        assert abs(hum.mixed_gain_bound_db(p) - 18.0) < 1e-12
    # a_long_line_bounds_as_its_infinite_line_and_a_far_piece_as_its_power
    assert abs(hum.divergence_db(250.0, 10.0) - (10.0 + 6.0285)) < 1e-3
    far = hum.divergence_db(40.0, 5000.0)
    assert abs(far - (20 * math.log10(5000.0) + 11.0 - 10 * math.log10(40.0))) < 0.01
    # a point: 20 lg d + 11, floored at 1 m
    assert abs(hum.divergence_db(0.0, 100.0) - 51.0) < 1e-12
    assert hum.divergence_db(0.0, 0.3) == 11.0
    # every_period_keeps_its_own_energy: a night-only source keeps its night energy
    emission = np.zeros((3, 8))
    emission[2] = 10.0 ** ((80.0 + np.array([-26.2, -16.1, -8.6, -3.2, 0.0, 1.2, 1.0, -1.1])) / 10.0)
    rb = hum.receiver_bound([0.5, 0.5, 0.5], 0.0, [0.1, 0.4, 1.0, 1.9, 3.7, 9.7, 32.8, 117.0])
    b = hum.received_energy_bound(emission, 0.0, 100.0, rb)
    assert b[0] == 0.0 and b[1] == 0.0 and b[2] > 0.0
    # the bound is the closed form: sum of bands x air, divergence, the 18 dB gain and reflection
    rb3 = hum.receiver_bound([0.2, 0.9, 0.4], 3.0, np.zeros(8))
    e = np.full((3, 8), 2.0)
    got = hum.received_energy_bound(e, 0.0, 200.0, rb3)
    want = 16.0 / (200.0 ** 2 * 10 ** 1.1) * 10 ** (21.0 / 10.0)
    assert np.allclose(got, want, rtol=1e-12)
    assert hum.loud([0.0, 0.0, 100.0 * 24 / 8 / 10.0])  # night only, exactly 20 dB Lden
    assert not hum.loud([0.0, 0.0, 100.0 * 24 / 8 / 10.0 * 0.999])


class StandIn:
    """A batch whose pair energies are a fixed fraction (between `low` and 1) of each pair's bound,
    with bands rising across the octaves."""

    def __init__(self, candidates, bounds, points, low=0.05):
        self.c, self.b, self.points, self.low = candidates, bounds, points, low
        self.calls, self.pairs = 0, 0
        self.cache = {}

    def energy(self, at, index):
        if (at, index) not in self.cache:
            off = self.c.offsets((self.points["x_m"][at], self.points["y_m"][at]))
            shape = np.linspace(1.0, 2.0, 8)
            shape /= shape.sum()
            for i in range(self.c.layer.size):
                d = math.hypot(off[i, 0], off[i, 1])
                bound = hum.received_energy_bound(self.c.emission[i], float(self.c.length_m[i]), d, self.b[at])
                f = self.low + (1.0 - self.low) * ((i * 2654435761 + at * 40503) % 1000) / 999.0
                self.cache[(at, i)] = (bound * f, np.outer(bound * f, shape))
        return self.cache[(at, index)]

    def pair_spectra(self, points, pairs):
        self.calls += 1
        self.pairs += pairs.size
        mixed = np.zeros((pairs.size, 3))
        bands = np.zeros((pairs.size, 3, 8))
        for j, (p, c) in enumerate(zip(pairs["point"], pairs["candidate"])):
            mixed[j], bands[j] = self.energy(int(p), int(c))
        return Evaluated(mixed, np.zeros((pairs.size, 2, 3)), bands)

    evaluate = pair_spectra


def quiet_field(n_sources=3000, n_points=4, seed=5, spread_m=6000.0, level=62.0):
    rng = np.random.default_rng(seed)
    n = n_sources
    ends = np.repeat(rng.uniform(-spread_m, spread_m, (n, 1, 2)), 2, axis=1)
    layer = rng.integers(0, 2, n)
    # each source's own level, one spectrum shape and the periods' offsets (a road's profile)
    shape = np.array([-8.0, -5.0, -3.0, -1.0, 0.0, -1.0, -3.0, -7.0])
    db = rng.normal(level, 4.0, n)[:, None, None] + np.array([0.0, -3.0, -8.0])[None, :, None] + shape[None, None, :]
    emission = 10.0 ** (db / 10.0)
    cands = hum.HumCandidates(ends, np.zeros(n, bool), np.zeros(n), layer, emission)
    pts = np.zeros(n_points, abi.POINT_DTYPE)
    pts["x_m"] = rng.uniform(-500, 500, n_points)
    pts["y_m"] = rng.uniform(-500, 500, n_points)
    alpha = np.array([0.1, 0.4, 1.0, 1.9, 3.7, 9.7, 32.8, 117.0])
    bounds = [hum.receiver_bound([0.5, 0.6, 0.7], 0.0, alpha) for _ in range(n_points)]
    return cands, pts, bounds


def exact_quiet(batch, cands, pts, bounds):
    """Every quiet source of every point and layer evaluated and summed."""
    out = np.zeros((pts.size, abi.LAYERS, 3))
    for at in range(pts.size):
        near, dist = hum.reached(cands, (pts["x_m"][at], pts["y_m"][at]))
        for i, d in zip(near.tolist(), dist.tolist()):
            b = hum.received_energy_bound(cands.emission[i], 0.0, d, bounds[at])
            if not hum.loud(b):
                out[at, cands.layer[i]] += batch.energy(at, i)[0]
    return out


def test_sampled_hum_estimates_the_exact_quiet_sum():
    cands, pts, bounds = quiet_field()
    batch = StandIn(cands, bounds, pts)
    truth = exact_quiet(batch, cands, pts, bounds)
    loud = np.zeros((pts.size, abi.LAYERS, 3))
    quiet, bands = hum.sampled_hum(batch, pts, cands, bounds, loud, seeds=np.arange(pts.size) * 977 + 11)
    ok = truth > 0
    assert ok.sum() >= pts.size * 2 * 3
    rel = np.abs(quiet[ok] / truth[ok] - 1.0)
    # two standard errors within 1.16 % of the answer: the estimate lies well within 3 %
    assert rel.max() < 0.03, rel.max()
    # fewer pairs evaluated than the population
    population = sum(hum.reached(cands, (pts["x_m"][a], pts["y_m"][a]))[0].size for a in range(pts.size))
    assert batch.pairs < population
    # the bands carry the energy per period
    assert np.allclose(bands.sum(axis=3), quiet, rtol=1e-12, atol=0.0)
    assert not quiet[~ok].any()


def test_sampled_hum_is_unbiased_over_seeds():
    cands, pts, bounds = quiet_field(n_sources=12000, n_points=1, seed=8)
    batch = StandIn(cands, bounds, pts, low=0.6)
    truth = exact_quiet(batch, cands, pts, bounds)
    loud = np.zeros((1, abi.LAYERS, 3))
    runs = []
    for s in range(24):
        batch.pairs = 0
        runs.append(hum.sampled_hum(batch, pts, cands, bounds, loud, seeds=[s])[0])
        assert batch.pairs < 12000 // 3  # sampled, not exhausted
    runs = np.array(runs)
    ok = truth > 0
    rel = runs[:, ok] / truth[ok] - 1.0
    # each within its budget, the mean within a fifth of it (24 independent samples)
    assert np.abs(rel).max() < 0.03
    assert np.abs(rel.mean(axis=0)).max() < 0.006
    assert not np.array_equal(runs[0], runs[1])


def test_an_exhausted_population_is_the_exact_sum():
    cands, pts, bounds = quiet_field(n_sources=40, n_points=2, seed=9)
    batch = StandIn(cands, bounds, pts)
    truth = exact_quiet(batch, cands, pts, bounds)
    quiet, _ = hum.sampled_hum(batch, pts, cands, bounds, np.zeros((2, abi.LAYERS, 3)), seeds=[1, 2])
    assert np.allclose(quiet, truth, rtol=1e-12, atol=0.0)


def test_a_loud_answer_omits_its_quiet_layer():
    cands, pts, bounds = quiet_field(n_sources=300, n_points=1, seed=10)
    batch = StandIn(cands, bounds, pts)
    loud = np.full((1, abi.LAYERS, 3), 1e12)
    quiet, bands = hum.sampled_hum(batch, pts, cands, bounds, loud, seeds=[4])
    assert not quiet.any() and not bands.any() and batch.calls == 0


def test_loud_pairs_are_the_loud_bounds():
    cands, pts, bounds = quiet_field(n_sources=400, n_points=3, seed=12, spread_m=800.0, level=80.0)
    pairs = hum.loud_pairs(pts, cands, bounds)
    assert pairs.size > 0
    for p, c in zip(pairs["point"], pairs["candidate"]):
        off = cands.offsets((pts["x_m"][p], pts["y_m"][p]))[c]
        assert hum.loud(hum.received_energy_bound(cands.emission[c], 0.0, math.hypot(*off), bounds[p]))
    assert np.all(np.diff(pairs["point"].astype(np.int64)) >= 0)


def test_a_piece_is_bounded_at_its_nearest_point():
    ends = np.array([[[0.0, -50.0], [0.0, 50.0]], [[30.0, 40.0], [30.0, 40.0]]])
    c = hum.HumCandidates(ends, np.array([True, False]), np.array([100.0, 0.0]), np.zeros(2, np.int64),
                          np.ones((2, 3, 8)))
    off = c.offsets((10.0, 20.0))
    assert np.allclose(off[0], [-10.0, 0.0]) and np.allclose(off[1], [20.0, 20.0])
    off = c.offsets((10.0, 80.0))
    assert np.allclose(off[0], [-10.0, -30.0])
