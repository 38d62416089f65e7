"""Gpu::evaluate_pairs takes its pairs in a coherent order of its own (by source, then the
receiver's Morton code) and returns each pair's results at the pair's index: the same pairs given
in any order give the same answers, bit for bit (point bands, summed by FP32 atomics in whatever
order the waves reach them, within 1e-5 relative). Then the reordered launch against the FP64
oracle on the same pairs (0.05 dB)."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def test_pair_order_does_not_change_answers(gpu):
    print("case: pair order, tile scene scale 0.06, 30000 pairs", flush=True)
    sc = scenes.tile_scene(scale=0.06)
    rng = np.random.default_rng(71)
    n = 30000
    pp = rng.integers(0, sc.points.size, n).astype(np.uint32)
    ps = rng.integers(0, sc.sources.size, n).astype(np.uint32)
    square = gpu.upload(sc.scene, sc.sources)
    first = gpu.evaluate_pairs(square, sc.points, pp, ps, pair_bands=True)
    perm = rng.permutation(n)
    second = gpu.evaluate_pairs(square, sc.points, pp[perm], ps[perm], pair_bands=True)
    assert np.array_equal(second["failed"], first["failed"][perm])
    assert np.array_equal(second["sums"], first["sums"][perm])
    assert np.array_equal(second["pair_bands"], first["pair_bands"][perm])
    a, b = first["point_bands"], second["point_bands"]
    assert np.array_equal(a > 0, b > 0)
    assert np.allclose(a, b, rtol=1e-5, atol=0.0)
    # Candidate order (all of a source's pairs together, receivers in the caller's order) too.
    by_source = np.argsort(ps, kind="stable")
    third = gpu.evaluate_pairs(square, sc.points, pp[by_source], ps[by_source], point_bands=False)
    assert np.array_equal(third["sums"], first["sums"][by_source])
    pick = rng.choice(n, 2000, replace=False)
    want = oracle.evaluate_pairs(sc.scene, sc.points, sc.sources, pp[pick], ps[pick], point_bands=False)
    ok = want["failed"] == 0
    assert np.array_equal(first["failed"][pick] != 0, want["failed"] != 0)
    got_s, want_s = first["sums"][pick][ok][:, :3].astype(np.float64), want["sums"][ok][:, :3]
    assert np.array_equal(got_s > 0, want_s > 0)
    heard = want_s > 0
    err = np.abs(10.0 * np.log10(got_s[heard] / want_s[heard]))
    print(f"reordered pairs vs oracle: n={err.size} max {err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    square.close()
