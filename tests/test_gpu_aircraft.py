"""The HIP Doc 29 aircraft grid against the FP64 oracle on the same seeded scene: Lden within
0.05 dB, N65/N70 event counts and the LAmax histogram bins bit-exact."""
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


@pytest.mark.parametrize("scale", [0.02, 0.05])
def test_aircraft_grid_matches_oracle(gpu, scale):
    print(f"case: aircraft scale {scale}", flush=True)
    sc = scenes.aircraft_scene(scale=scale)
    got = gpu.aircraft_grid(sc)
    rng = np.random.default_rng(1)
    n = sc.points_xy.shape[0]
    sample = rng.choice(n, min(400, n), replace=False)
    want = oracle.aircraft_points(sc, sc.points_xy[sample])
    assert np.array_equal(got["n65"][sample], want["n65"])
    assert np.array_equal(got["n70"][sample], want["n70"])
    assert np.array_equal(got["hist"][sample], want["hist"])
    from quietmap_amd.painter import lden_db

    hours = np.array([12.0, 4.0, 8.0]) * 3600.0
    g = lden_db(got["period_energy"][sample] / hours)
    w = lden_db(want["period_energy"] / hours)
    ok = np.isfinite(w)
    assert np.array_equal(np.isfinite(g), ok)
    err = np.abs(g[ok] - w[ok])
    print(f"aircraft Lden: n={ok.sum()} max={err.max():.5f} dB; events binned {int(want['hist'].sum())}, "
          f"N65 total {int(want['n65'].sum())}", flush=True)
    assert err.max() <= 0.05
    assert want["hist"].sum() > 100


def test_pass_classes_match_oracle(gpu):
    """The 48 pass classes per period: every heard segment's pass in the oracle's class (the
    passing shares, FP64 geometry, agree to 1e-9 and sit in the same classes), each class's energy
    (FP32 SEL energies summed in FP64) within the Lden bar of 0.05 dB."""
    print("case: aircraft pass classes scale 0.02", flush=True)
    sc = scenes.aircraft_scene(scale=0.02)
    got = gpu.aircraft_grid(sc, passes=True)
    rng = np.random.default_rng(2)
    n = sc.points_xy.shape[0]
    sample = rng.choice(n, min(200, n), replace=False)
    want = oracle.aircraft_points(sc, sc.points_xy[sample], passes=True)
    g, w = got["passes"][sample], want["passes"]
    assert np.array_equal(g[:, 1] > 0, w[:, 1] > 0)
    assert np.allclose(g[:, 1], w[:, 1], rtol=1e-9, atol=0)
    held = w[:, 0] > 0
    err = np.abs(10 * np.log10(g[:, 0][held] / w[:, 0][held]))
    print(f"pass classes: {int(held.sum())} classes held, energy max |err| {err.max():.5f} dB", flush=True)
    assert err.max() <= 0.05
    assert np.array_equal(got["hist"][sample], want["hist"])


def test_screened_grid_matches_oracle(gpu):
    """Horizon screening (terrain and building horizons, ground under the vertices): Lden within
    0.05 dB, N65/N70 and the histogram unchanged and bit-exact, pass classes in the oracle's
    classes; the screening lowers some receivers' energy."""
    print("case: aircraft screened scale 0.03", flush=True)
    sc = scenes.aircraft_scene(scale=0.03)
    scr = scenes.aircraft_horizons(sc)
    got = gpu.aircraft_grid(sc, passes=True, screening=scr)
    free = gpu.aircraft_grid(sc)
    rng = np.random.default_rng(4)
    n = sc.points_xy.shape[0]
    sample = rng.choice(n, min(300, n), replace=False)
    want = oracle.aircraft_points(sc, sc.points_xy[sample], passes=True,
                                  screening=dict(scr, horizon_of=scr["horizon_of"][sample]))
    for k in ("n65", "n70", "hist"):
        assert np.array_equal(got[k][sample], want[k])
        assert np.array_equal(got[k], free[k])
    from quietmap_amd.painter import lden_db

    hours = np.array([12.0, 4.0, 8.0]) * 3600.0
    g = lden_db(got["period_energy"][sample] / hours)
    w = lden_db(want["period_energy"] / hours)
    ok = np.isfinite(w)
    assert np.array_equal(np.isfinite(g), ok)
    err = np.abs(g[ok] - w[ok])
    drop = lden_db(free["period_energy"][sample] / hours)[ok] - w[ok]
    print(f"screened Lden: n={ok.sum()} max |err| {err.max():.5f} dB; screening up to {drop.max():.2f} dB", flush=True)
    assert err.max() <= 0.05
    assert drop.max() > 0.5
    gp, wp = got["passes"][sample], want["passes"]
    assert np.array_equal(gp[:, 1] > 0, wp[:, 1] > 0)
    assert np.allclose(gp[:, 1], wp[:, 1], rtol=1e-9, atol=0)
