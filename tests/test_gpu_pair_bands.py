"""Each pair's own bands (the `flat` output that Batch::pair_spectra reads, paint-gpu/src/hum_pairs.rs)
from qm_evaluate_pairs and qm_evaluate_line_pairs: their bands sum to the pair's mixed energy per
period, their sums per point and layer equal the FP64 oracle's point bands within 0.05 dB, and a
failed pair's are zero."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import abi, scenes

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


def test_point_pair_bands(gpu):
    print("case: point pair bands", flush=True)
    sc = scenes.tile_scene(scale=0.06)
    rng = np.random.default_rng(51)
    pp = rng.integers(0, sc.points.size, 3000).astype(np.uint32)
    ps = rng.integers(0, sc.sources.size, 3000).astype(np.uint32)
    square = gpu.upload(sc.scene, sc.sources)
    got = gpu.evaluate_pairs(square, sc.points, pp, ps, pair_bands=True)
    spectra = got["pair_bands"].astype(np.float64)
    ok = got["failed"] == 0
    assert np.all(spectra[~ok] == 0.0)
    mixed = spectra.sum(axis=2)
    assert np.allclose(mixed[ok], got["sums"][ok, :3], rtol=1e-5, atol=0.0)
    want = oracle.evaluate_pairs(sc.scene, sc.points, sc.sources, pp, ps)
    assert np.array_equal(ok, want["failed"] == 0)
    per_point = np.zeros((sc.points.size, abi.LAYERS, 3, 8))
    np.add.at(per_point, (pp[ok], sc.sources["layer"][ps[ok]]), spectra[ok])
    err = db_error(per_point, want["point_bands"])
    print(f"point bands from pair bands: n={err.size} max={err.max():.5f} dB", flush=True)
    assert err.size > 1000 and err.max() <= 0.05
    square.close()


def test_line_pair_bands(gpu):
    print("case: line pair bands", flush=True)
    sc = scenes.line_scene()
    rng = np.random.default_rng(52)
    pp = rng.integers(0, sc.points.size, 1500).astype(np.uint32)
    pl = rng.integers(0, sc.pieces.size, 1500).astype(np.uint32)
    square = gpu.upload(sc.scene, sc.sources)
    got = gpu.evaluate_line_pairs(square, sc.points, sc.pieces, pp, pl, point_bands=True, pair_bands=True)
    spectra = got["pair_bands"].astype(np.float64)
    ok = got["failed"] == 0
    assert np.all(spectra[~ok] == 0.0)
    assert np.allclose(spectra.sum(axis=2)[ok], got["sums"][ok, :3], rtol=1e-5, atol=1e-30)
    per_point = np.zeros((sc.points.size, abi.LAYERS, 3, 8))
    np.add.at(per_point, (pp[ok], sc.pieces["layer"][pl[ok]]), spectra[ok])
    assert np.allclose(per_point, got["point_bands"], rtol=1e-5, atol=1e-30)
    square.close()
