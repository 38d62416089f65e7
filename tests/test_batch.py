"""The batch's host logic (quietmap_amd/batch.py, the reference's GpuBatch, paint-gpu/src/batch.rs)
over a stand-in card that answers each pair with a known function of its point and candidate:
pairs of point and line candidates launched in candidate order at most `pairs_per_launch` at a time,
each launch's points once and each reached point's slot in the band sums,
energies back in the pairs' own order, weighted bands summed per point and layer across launches,
refused pairs listed with their bits; too many pixels in a loudness batch and indices out of range
refused. The card itself is checked in tests/test_gpu_batch.py."""
import numpy as np
import pytest

from quietmap_amd import abi
from quietmap_amd.batch import PAIR_DTYPE, GpuBatch, by_candidate, pairs_of
from quietmap_amd.painter import PainterError


class FakeSquare:
    """Candidates: n point sources, then `lines` line pieces (layers cycling over both)."""

    def __init__(self, n, lines=0):
        self.sources = np.zeros(n, abi.SOURCE_DTYPE)
        self.sources["layer"] = np.arange(n) % abi.LAYERS
        self.lines = np.zeros(lines, abi.LINE_PIECE_DTYPE)
        self.lines["layer"] = (np.arange(lines) + 3) % abi.LAYERS
        self.candidates = n + lines

    def layer_of(self, candidate):
        return np.concatenate([self.sources["layer"], self.lines["layer"]])[np.asarray(candidate, np.int64)]


class FakeGpu:
    """Answers pair (receiver at x, candidate c): mixed = x + c per period (+ period), calm 2x,
    downwind 3x; bands all in band 0, added at the receiver's point with the pair's weight;
    refuses candidates divisible by 17."""

    def __init__(self, square):
        self.square = square
        self.launches = []

    def evaluate_candidates(self, square, receivers, candidate, receiver_of_pair, point_of_receiver=None,
                            n_points=None, band_weight=None, bands=None, pair_spectra=False):
        self.launches.append((receivers.copy(), receiver_of_pair.copy(), candidate.copy()))
        assert np.unique(receivers["x_m"]).size == receivers.size  # each point once
        assert np.all(np.diff(candidate.astype(np.int64)) >= 0)  # candidate order
        assert candidate.max() < square.candidates
        n = receiver_of_pair.size
        base = receivers["x_m"][receiver_of_pair] + candidate
        sums = np.zeros((n, 9), np.float32)
        for q in range(3):
            sums[:, q] = base + q
            sums[:, 3 + q] = 2 * (base + q)
            sums[:, 6 + q] = 3 * (base + q)
        failed = np.where(candidate % 17 == 0, 4, 0).astype(np.uint32)
        sums[failed != 0] = 0
        out = dict(sums=sums, failed=failed)
        w = np.ones(n) if band_weight is None else band_weight
        if bands is not None:
            por = np.arange(receivers.size) if point_of_receiver is None else point_of_receiver
            ok = failed == 0
            for q in range(3):
                np.add.at(bands[:, :, q, 0], (por[receiver_of_pair[ok]], square.layer_of(candidate[ok])),
                          (w[ok] * sums[ok, q]).astype(np.float32))
        if pair_spectra:
            sp = np.zeros((n, 3, 8), np.float32)
            sp[:, :, 0] = sums[:, 0:3]
            out["pair_spectra"] = sp
        return out


def setup(n_points=50, n_candidates=60, n_pairs=700, seed=1, lines=20):
    rng = np.random.default_rng(seed)
    points = np.zeros(n_points, abi.POINT_DTYPE)
    points["x_m"] = np.arange(n_points) * 1000.0
    pairs = pairs_of(rng.integers(0, n_points, n_pairs), rng.integers(0, n_candidates + lines, n_pairs),
                     np.where(rng.uniform(size=n_pairs) < 0.3, rng.uniform(0.1, 5.0, n_pairs), 1.0))
    square = FakeSquare(n_candidates, lines)
    return points, pairs, square


@pytest.mark.parametrize("per_launch", [64, 700, 5000])
def test_evaluate_orders_gathers_and_sums(per_launch):
    points, pairs, square = setup()
    gpu = FakeGpu(square)
    out = GpuBatch(gpu, square, pairs_per_launch=per_launch).evaluate(points, pairs)
    assert len(gpu.launches) == -(-pairs.size // per_launch)
    base = points["x_m"][pairs["point"]] + pairs["candidate"]
    refused = pairs["candidate"] % 17 == 0
    for q in range(3):
        want = np.where(refused, 0.0, np.float32(base + q))
        np.testing.assert_allclose(out.mixed[:, q], want, rtol=1e-6)
        np.testing.assert_allclose(out.states[:, 0, q], 2 * want, rtol=1e-6)
        np.testing.assert_allclose(out.states[:, 1, q], 3 * want, rtol=1e-6)
    assert np.array_equal(out.refused, np.flatnonzero(refused)) and np.all(out.refused_bits == 4)
    bands = np.zeros((points.size, abi.LAYERS, 3))
    for k in np.flatnonzero(~refused):
        layer = square.layer_of(pairs["candidate"][k])
        bands[pairs["point"][k], layer] += pairs["band_weight"][k] * (base[k] + np.arange(3))
    np.testing.assert_allclose(out.bands[:, :, :, 0], bands, rtol=1e-5)
    assert not out.bands[:, :, :, 1:].any()


def test_pair_spectra_are_unweighted():
    points, pairs, square = setup(seed=2)
    out = GpuBatch(FakeGpu(square), square, pairs_per_launch=100).pair_spectra(points, pairs)
    np.testing.assert_allclose(out.bands[:, :, 0], out.mixed, rtol=1e-6)


def test_by_candidate_is_a_stable_counting_order():
    pairs = pairs_of([0, 1, 2, 3, 4, 5], [3, 1, 3, 0, 1, 3])
    assert by_candidate(pairs, 4).tolist() == [3, 1, 4, 0, 2, 5]
    with pytest.raises(PainterError):
        by_candidate(pairs, 3)


def test_refusals():
    points, pairs, square = setup(seed=3)
    batch = GpuBatch(FakeGpu(square), square)
    bad = pairs.copy()
    bad["point"][5] = points.size
    with pytest.raises(PainterError):
        batch.evaluate(points, bad)
    with pytest.raises(PainterError):
        batch.loudness(np.zeros(10, np.float32), np.zeros(9000, np.uint64))
    empty = batch.evaluate(points, np.zeros(0, PAIR_DTYPE))
    assert empty.mixed.shape == (0, 3) and not empty.bands.any()
