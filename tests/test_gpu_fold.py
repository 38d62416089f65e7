"""The map's folding and packing of pixel sound on the card (paint/src/sone_fold.rs
fold_placed_for_map, popup/src/percentiles/weather.rs merge_close_states, paint/src/sone.rs
write_packed) against the FP64 oracle: the folded FP32 pixels bit for bit (NaN past each pixel's
end on both sides), and Nden of the card's folded pixels within 0.01 sone of the oracle's Nden of
the oracle's folded pixels."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import pixels

pytestmark = pytest.mark.gpu
CURVES = np.full((7, 3, 28), np.nan)


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def special_pixels():
    """The fold's edge cases: a collapse by f32-equal states, a weak event beside a floor, a Kurze
    line at its held threshold, a budget that folds half of twelve events, empty layers."""
    empty = (np.zeros(3), [])
    steady = lambda means, profile=None: pixels.pack_line_f64(
        (np.asarray(means) > 0).any(axis=1).astype(float), means, [np.inf] * 3, profile=profile)
    means = np.full((3, 10), 1_000_000.0)
    means[:, 5:] = 1_000_000.02
    guard = np.full((3, 10), 100.0)
    guard[:, 5:] = 10_000.0
    out = [pixels.pack_pixel_f64([(np.zeros(3), [steady(means), steady(guard, 0)])] + [empty] * 5, CURVES)]
    duty = pixels.pack_line_f64([50_400.0] * 3, np.full((3, 10), 50_400.0), [np.inf] * 3, duty=[0.5] * 3)
    out.append(pixels.pack_pixel_f64([empty, (np.full(3, 1_209_600.0), [duty])] + [empty] * 4, CURVES))
    lam, m = 0.5, 1000.0
    floor = m * oracle.relative_intensity(lam, 0.999) / 0.05 - m
    kurze = pixels.pack_line_f64([m] * 3, np.full((3, 10), m), [lam] * 3)
    for margin in (1.001, 0.999):
        out.append(pixels.pack_pixel_f64([(np.full(3, floor * margin), [kurze])] + [empty] * 5, CURVES))
    event = pixels.pack_line_f64([5000.0] * 3, np.full((3, 10), 5000.0), [np.inf] * 3, duty=[0.5] * 3)
    out.append(pixels.pack_pixel_f64([empty] * 5 + [(np.full(3, 100_000.0), [event] * 12)], CURVES))
    out.append(pixels.pack_pixel_f64([empty] * 6, CURVES))
    return pixels.concat_f64(out)


@pytest.mark.parametrize("case", ["synthetic", "many lines", "special"])
def test_fold_matches_oracle(gpu, case):
    print(f"case: fold {case}", flush=True)
    if case == "synthetic":
        values, offsets = pixels.synthetic_unfolded(seed=5, count=160)
    elif case == "many lines":
        values, offsets = pixels.synthetic_unfolded(seed=6, count=12, lines_max=60)
    else:
        values, offsets = special_pixels()
    want, used = oracle.fold_pixels(values, offsets)
    nd, got = gpu.pixel_loudness_unfolded(values, offsets, folded=True)
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    bad = np.flatnonzero(~same)
    print(f"folded values {got.size}: {bad.size} differ", flush=True)
    assert bad.size == 0, (bad[:8], got[bad[:8]], want[bad[:8]])
    ref = oracle.pixel_loudness(want, offsets)
    err = np.abs(nd - ref)
    print(f"Nden: max |err| {err.max():.2e} sone, max {ref.max():.2f} sone", flush=True)
    assert err.max() <= 0.01
    if case == "synthetic":
        assert (ref[:, 6] > 0).sum() > offsets.size // 2
        # The card's loudness of the card's folding agrees with its loudness of the oracle's folding.
        again = gpu.pixel_loudness(want, offsets)
        assert np.allclose(nd, again, rtol=1e-6, atol=1e-6)


def test_unfolded_malformed_fails(gpu):
    from quietmap_amd.painter import PainterError

    print("case: fold malformed", flush=True)
    good = pixels.pack_pixel_f64([(np.zeros(3), [])] * 6, CURVES)
    bad = good.copy()
    bad[3] = 2.5  # a line count that is not whole
    with pytest.raises(PainterError):
        gpu.pixel_loudness_unfolded(bad, np.zeros(1, np.uint64))
    assert gpu.pixel_loudness_unfolded(np.zeros(0), np.zeros(0, np.uint64)).shape == (0, 7)
