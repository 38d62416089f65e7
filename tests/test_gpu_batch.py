"""The batch (quietmap_amd/batch.py GpuBatch, paint-gpu/src/batch.rs) on the card against the FP64
oracle: a pass's pairs in random order with importance weights, split over many launches, give
every pair's energy (mixed, calm, downwind) and every point's weighted bands within 0.05 dB, and
list exactly the pairs the oracle fails; pair_spectra gives each pair's bands summing to its
energy."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import scenes
from quietmap_amd.batch import GpuBatch, pairs_of

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def db_error(got, want):
    assert np.array_equal(want > 0, got > 0), "silent on one side only"
    ok = want > 0
    return np.abs(10 * np.log10(got[ok] / want[ok]))


def test_batch_evaluate_matches_oracle(gpu):
    print("case: batch evaluate", flush=True)
    sc = scenes.tile_scene(scale=0.06)
    rng = np.random.default_rng(61)
    n = 5000
    pairs = pairs_of(rng.integers(0, sc.points.size, n), rng.integers(0, sc.sources.size, n),
                     np.where(rng.uniform(size=n) < 0.4, rng.uniform(0.5, 8.0, n), 1.0))
    square = gpu.upload(sc.scene, sc.sources)
    out = GpuBatch(gpu, square, pairs_per_launch=700).evaluate(sc.points, pairs)
    want = oracle.evaluate_pairs(sc.scene, sc.points, sc.sources, pairs["point"], pairs["candidate"],
                                 pair_weight=pairs["band_weight"].astype(np.float32))
    assert np.array_equal(out.refused, np.flatnonzero(want["failed"] != 0))
    ok = want["failed"] == 0
    got = np.concatenate([out.mixed, out.states.reshape(n, 6)], axis=1)
    err = db_error(got[ok], want["sums"][ok])
    err_b = db_error(out.bands, want["point_bands"])
    print(f"batch: pairs {int(ok.sum())} max {err.max():.5f} dB; bands max {err_b.max():.5f} dB", flush=True)
    assert err.max() <= 0.05 and err_b.max() <= 0.05
    spectra = GpuBatch(gpu, square, pairs_per_launch=1500).pair_spectra(sc.points, pairs)
    assert np.allclose(spectra.bands.sum(axis=2)[ok], out.mixed[ok], rtol=1e-5)
    square.close()
