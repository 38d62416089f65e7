"""The HIP Zwicker loudness against the FP64 oracle (ISO 532-1) within 0.01 sone, and against
MoSQITo's values the reference pins (physics/src/loudness.rs:196-266); both input layouts."""
import numpy as np
import pytest

from oracle import oracle
from tests.test_oracle_loudness import RANDOM, octaves_to_thirds, quiet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def spectra(n, seed=4):
    rng = np.random.default_rng(seed)
    levels = rng.uniform(-10.0, 95.0, (n, 28))
    # Some quiet bands, some silent cells, some loud tonal cells.
    levels[rng.uniform(size=levels.shape) < 0.1] = -80.0
    levels[: n // 50] = -100.0
    return levels


def test_mosqito_values(gpu):
    print("case: loudness mosqito", flush=True)
    cases = []
    one = quiet()
    one[16] = 40.0
    cases.append((one.copy(), 0.927, 0.002))
    one[16] = 70.0
    cases.append((one.copy(), 6.944, 0.005))
    road = [98.35, 91.1, 89.4, 90.7, 95.6, 92.4, 84.6, 76.2]
    for offset, sone in ((-60.0, 2.48), (-40.0, 12.075), (-20.0, 43.34)):
        cases.append((octaves_to_thirds(road, offset), sone, 0.003 * sone))
    aircraft = [75.0, 78.0, 76.0, 72.0, 66.0, 58.0, 45.0, 30.0]
    cases.append((octaves_to_thirds(aircraft, -5.0), 21.39, 0.003 * 21.39))
    for levels, sone in RANDOM:
        cases.append((np.array(levels), sone, 0.003 * sone))
    got = gpu.loudness(np.stack([c[0] for c in cases]))
    for (levels, sone, tol), n in zip(cases, got):
        assert abs(n - sone) <= tol, (n, sone)


def test_random_spectra_within_0_01_sone(gpu):
    """The same FP32 levels on both sides: ISO 532-1's low-frequency table is discontinuous at
    its range limits (a band at exactly 80 dB switches rows), so rounding an FP64 level to FP32
    first may change the row."""
    print("case: loudness random", flush=True)
    levels = spectra(200_000).astype(np.float32)
    got = gpu.loudness(levels)
    want = oracle.zwicker_sone(levels.astype(np.float64))
    err = np.abs(got.astype(np.float64) - want)
    print(f"loudness: n={err.size} max={err.max():.6f} sone p99.9={np.quantile(err, 0.999):.6f}", flush=True)
    assert err.max() <= 0.01


def test_band_major_planes_match(gpu):
    print("case: loudness planes", flush=True)
    from quietmap_amd import hip

    levels = spectra(65_536, seed=9).astype(np.float32)
    planes = hip.DeviceBuffer.from_array(np.ascontiguousarray(levels.T))
    out = hip.DeviceBuffer(levels.shape[0] * 4)
    gpu.loudness_device(planes.ptr, levels.shape[0], 1, levels.shape[0], out.ptr)
    hip.synchronize()
    assert np.array_equal(out.to_array(np.float32, levels.shape[0]), gpu.loudness(levels))
