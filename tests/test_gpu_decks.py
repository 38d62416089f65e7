"""Sources on bridge decks (physics/src/ray.rs RayEnds::source_ground_m, popup/src/evaluate.rs:
181-196, 243-252 deck_at, popup/src/candidates.rs:151-158 traced_ground_m) on the card against the
FP64 oracle, whose deck handling is pinned by ray_tests.rs a_source_on_a_deck_stands_on_it_above_the_
valley: point sources on decks above and below the terrain under them (with platforms, the terrain
cut to the deck within them) and line pieces on decks sloping between their ends, pair sums
within 0.05 dB and the same rays; decks move the pairs (the test is not vacuous); the surrogate
painter sends deck sources to the exact path."""
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


def db(x):
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.asarray(x, np.float64))


def on_decks(scene, sources, rng, share=0.5):
    """A share of the sources moved onto decks 8-30 m above (or 0-3 m below) the terrain under them."""
    out = sources.copy()
    pick = rng.uniform(size=out.size) < share
    ground = np.array([oracle.ground_at(scene.terrain, x, y)[1][0] for x, y in zip(out["x_m"], out["y_m"])])
    rise = np.where(rng.uniform(size=out.size) < 0.8, rng.uniform(8.0, 30.0, out.size), -rng.uniform(0.0, 3.0, out.size))
    out["on_deck"] = pick.astype(np.uint8)
    out["deck_m"] = np.where(pick, ground + rise, 0.0)
    out["platform_half_width_m"] = np.where(pick & (rng.uniform(size=out.size) < 0.5), rng.uniform(2.0, 12.0, out.size),
                                            out["platform_half_width_m"])
    return out, pick


def test_point_sources_on_decks_match_oracle(gpu):
    print("case: point sources on decks", flush=True)
    sc = scenes.tile_scene(scale=0.06)
    rng = np.random.default_rng(41)
    sources, deck = on_decks(sc.scene, sc.sources, rng)
    square = gpu.upload(sc.scene, sources)
    flat = gpu.upload(sc.scene, sc.sources)
    pp = rng.integers(0, sc.points.size, 4000).astype(np.uint32)
    ps = rng.integers(0, sources.size, 4000).astype(np.uint32)
    d = np.hypot(sources["x_m"][ps] - sc.points["x_m"][pp], sources["y_m"][ps] - sc.points["y_m"][pp])
    keep = d <= sc.cutoff_m
    pp, ps = pp[keep], ps[keep]
    got = gpu.evaluate_pairs(square, sc.points, pp, ps)
    want = oracle.evaluate_pairs(sc.scene, sc.points, sources, pp, ps)
    ok = (got["failed"] == 0) & (want["failed"] == 0)
    assert ok.mean() > 0.99
    err = np.abs(db(got["sums"][ok]) - db(want["sums"][ok]))
    assert not np.isinf(err).any(), "silent on one side only"
    err = err[~np.isnan(err)]
    print(f"deck pairs: n={int(ok.sum())} on decks {int(deck[ps[ok]].sum())} max={err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    # The decks move the pairs.
    base = gpu.evaluate_pairs(flat, sc.points, pp, ps)
    on = ok & deck[ps] & (base["failed"] == 0)
    moved = np.abs(db(got["sums"][on, 0]) - db(base["sums"][on, 0]))
    assert on.sum() > 100 and np.median(moved[np.isfinite(moved)]) > 0.5
    square.close()
    flat.close()


def test_line_pieces_on_decks_match_oracle(gpu):
    print("case: line pieces on decks", flush=True)
    sc = scenes.line_scene()
    rng = np.random.default_rng(42)
    pieces = sc.pieces.copy()
    pick = rng.uniform(size=pieces.size) < 0.6
    ga = np.array([oracle.ground_at(sc.scene.terrain, x, y)[1][0] for x, y in zip(pieces["ax_m"], pieces["ay_m"])])
    gb = np.array([oracle.ground_at(sc.scene.terrain, x, y)[1][0] for x, y in zip(pieces["bx_m"], pieces["by_m"])])
    pieces["on_deck"] = pick.astype(np.uint8)
    pieces["deck_a_m"] = np.where(pick, ga + rng.uniform(5.0, 25.0, pieces.size), 0.0)
    pieces["deck_b_m"] = np.where(pick, gb + rng.uniform(5.0, 25.0, pieces.size), 0.0)
    n = 2500
    pp = rng.integers(0, sc.points.size, n).astype(np.uint32)
    pl = rng.integers(0, pieces.size, n).astype(np.uint32)
    square = gpu.upload(sc.scene, sc.sources)
    got = gpu.evaluate_line_pairs(square, sc.points, pieces, pp, pl)
    want = oracle.evaluate_line_pairs(sc.scene, sc.points, pieces, pp, pl)
    ok = (got["failed"] == 0) & (want["failed"] == 0)
    assert np.array_equal(got["failed"] != 0, want["failed"] != 0)
    assert np.array_equal(got["rays"][ok], want["rays"][ok])
    err = np.abs(db(got["sums"][ok]) - db(want["sums"][ok]))
    assert not np.isinf(err).any(), "silent on one side only"
    err = err[~np.isnan(err)]
    print(f"deck line pairs: n={int(ok.sum())} on decks {int(pick[pl[ok]].sum())} max={err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    flat = gpu.evaluate_line_pairs(square, sc.points, sc.pieces, pp, pl)
    on = ok & pick[pl] & (flat["failed"] == 0)
    moved = np.abs(db(got["sums"][on, 0]) - db(flat["sums"][on, 0]))
    assert on.sum() > 100 and np.median(moved[np.isfinite(moved)]) > 0.2
    square.close()


def test_surrogate_sends_deck_sources_to_the_exact_path(gpu):
    from quietmap_amd import hip, surrogate
    from quietmap_amd.painter import RECEIVER_BYTES

    print("case: surrogate painter with deck sources", flush=True)
    sc = scenes.tile_scene(scale=0.03)
    nx, ny = sc.grid_shape
    n = nx * ny
    rng = np.random.default_rng(43)
    sources, deck = on_decks(sc.scene, sc.sources, rng, share=0.5)
    square = gpu.upload(sc.scene, sources)
    params = surrogate.with_output_bias(surrogate.init_params(9) * 0.5, np.linspace(3, 20, 8))
    net = surrogate.Mlp(gpu, params, max_batch=1 << 16)
    d_points = hip.DeviceBuffer.from_array(sc.points)
    d_rec = hip.DeviceBuffer(n * RECEIVER_BYTES)
    gpu.receivers_device(square, d_points.ptr, n, d_rec.ptr)
    d_e, d_p, d_f, d_s = (hip.DeviceBuffer(n * 24), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4),
                          hip.DeviceBuffer(n * 4))
    surrogate.paint_receivers_device(gpu, square, net, d_rec.ptr, nx, ny, sc.cutoff_m, d_e.ptr, d_p.ptr, d_f.ptr,
                                     d_s.ptr)
    hip.synchronize()
    paths = d_p.to_array(np.uint32, n).astype(np.int64)
    learned = d_s.to_array(np.uint32, n).astype(np.int64)
    # Every receiver hears all sources within the cutoff; the network answers only terrain sources.
    for r in rng.choice(n, 16, replace=False):
        d = np.hypot(sources["x_m"] - sc.points["x_m"][r], sources["y_m"] - sc.points["y_m"][r])
        near = d <= sc.cutoff_m
        assert paths[r] == near.sum()
        assert learned[r] <= (near & ~deck).sum()
    assert learned.sum() > 0
    square.close()
