"""Pixel loudness on the card (level distributions per period, hour group and weather state,
mixed, then Nden on each curve) against the FP64 oracle on the same packed pixels: Nden within
0.01 sone per curve; every mixture's time and energy within 1e-6 of the oracle's (the
distributions lose none to the grid; FP32 bins summed from FP64 prefix sums); and bin by bin
within 2e-3 of the mixture's total (a sum lying within FP32 rounding of a bin edge may fall on the
other side of it)."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import pixels

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


@pytest.mark.parametrize("seed,count,kurze", [(11, 48, True), (12, 24, False)])
def test_pixel_loudness_matches_oracle(gpu, seed, count, kurze):
    print(f"case: pixel loudness seed {seed} pixels {count} kurze {kurze}", flush=True)
    px = pixels.synthetic_pixels(seed=seed, count=count, kurze=kurze)
    packed, offsets = pixels.concat(px)
    got, gmix = gpu.pixel_loudness(packed, offsets, mixtures=True)
    want, wmix = oracle.pixel_loudness(packed, offsets, mixtures=True)
    err = np.abs(got - want)
    print(f"Nden: max |err| {err.max():.2e} sone, max {want.max():.2f} sone", flush=True)
    assert err.max() <= 0.01
    for k in (0, 1):
        g = gmix[..., k].astype(np.float64).sum(axis=-1)
        w = wmix[..., k].sum(axis=-1)
        scale = np.maximum(w, 1e-30 if k else 1.0)
        rel = np.abs(g - w) / scale
        per_bin = np.abs(gmix[..., k].astype(np.float64) - wmix[..., k]).max(axis=-1) / scale
        print(f"mixtures {'energy' if k else 'time'}: max relative {rel.max():.2e}, bin by bin {per_bin.max():.2e}",
              flush=True)
        assert rel.max() < 1e-6
        assert per_bin.max() < 2e-3
    assert (want[:, 6] > 0).sum() > count // 2


def test_empty_and_silent_pixels(gpu):
    print("case: pixel loudness empty and silent", flush=True)
    curves = np.full((7, 3, 28), np.nan)
    silent = pixels.pack_pixel([(np.zeros(3), []) for _ in range(6)], curves)
    packed, offsets = pixels.concat([silent, silent])
    got = gpu.pixel_loudness(packed, offsets)
    assert np.all(got == 0.0)
    assert gpu.pixel_loudness(np.zeros(0, np.float32), np.zeros(0, np.uint64)).shape == (0, 7)


def test_malformed_counts_fail(gpu):
    from quietmap_amd.painter import PainterError

    print("case: pixel loudness malformed", flush=True)
    curves = np.full((7, 3, 28), np.nan)
    good = pixels.pack_pixel([(np.zeros(3), []) for _ in range(6)], curves)
    bad = good.copy()
    bad[3] = 1e6  # a line count the input does not hold
    packed, offsets = pixels.concat([bad])
    with pytest.raises(PainterError):
        gpu.pixel_loudness(packed, offsets)
