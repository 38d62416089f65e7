"""The oracle's ISO 532-1 loudness against MoSQITo 1.2's values, as the reference pins its own
(physics/src/loudness.rs:196-266): 1 kHz at 40 / 70 dB, a road spectrum at three levels, an
aircraft spectrum and four random spectra; plus monotonicity and silence."""
import numpy as np

from oracle import oracle

SILENT = -60.0


def quiet():
    return np.full(28, SILENT)


def octaves_to_thirds(octaves, offset):
    """Each octave's energy shared evenly by its three thirds (loudness.rs:212-222)."""
    thirds = quiet()
    for octave, level in enumerate(octaves):
        for k in range(3):
            thirds[4 - 1 + 3 * octave + k] = max(level + offset - 10 * np.log10(3.0), SILENT)
    return thirds


RANDOM = [
    ([56.88, 77.29, 68.18, 26.89, 32.51, 75.52, 10.39, 71.59, 69.78, 45.10, 32.73, 30.88, 29.12, 43.38, 47.84, 51.51,
      84.66, 69.45, 56.66, 84.17, 26.15, 22.02, 55.94, 13.30, 12.68, 48.62, 44.97, 78.79], 53.64),
    ([57.19, 48.56, 47.27, 28.56, 10.88, 24.43, 61.90, 25.05, 37.72, 10.28, 72.25, 21.58, 30.07, 76.02, 48.23, 73.54,
      57.98, 65.63, 16.86, 50.59, 48.08, 75.35, 37.09, 54.86, 14.44, 39.07, 34.23, 21.26], 42.0),
    ([71.23, 38.46, 83.41, 54.25, 55.38, 57.85, 60.73, 21.31, 43.02, 27.97, 40.19, 17.25, 82.59, 26.13, 60.38, 32.53,
      75.56, 59.67, 19.87, 73.38, 80.87, 77.79, 52.73, 20.91, 24.43, 79.59, 51.42, 23.54], 67.0),
    ([76.30, 58.12, 52.73, 38.22, 40.82, 27.96, 12.85, 75.72, 45.08, 51.07, 34.16, 66.35, 11.89, 37.91, 12.28, 19.22,
      82.54, 59.33, 42.12, 49.28, 75.46, 35.82, 54.27, 61.28, 36.66, 48.93, 67.39, 78.19], 46.27),
]


def test_one_khz_third():
    one = quiet()
    one[16] = 40.0
    assert abs(oracle.zwicker_sone(one)[0] - 0.927) < 0.002
    one[16] = 70.0
    assert abs(oracle.zwicker_sone(one)[0] - 6.944) < 0.005


def test_road_and_aircraft_spectra():
    road = [98.35, 91.1, 89.4, 90.7, 95.6, 92.4, 84.6, 76.2]
    for offset, sone in ((-60.0, 2.48), (-40.0, 12.075), (-20.0, 43.34)):
        n = oracle.zwicker_sone(octaves_to_thirds(road, offset))[0]
        assert abs(n - sone) / sone < 0.003, (offset, n)
    aircraft = [75.0, 78.0, 76.0, 72.0, 66.0, 58.0, 45.0, 30.0]
    n = oracle.zwicker_sone(octaves_to_thirds(aircraft, -5.0))[0]
    assert abs(n - 21.39) / 21.39 < 0.003


def test_random_spectra():
    for levels, sone in RANDOM:
        n = oracle.zwicker_sone(levels)[0]
        assert abs(n - sone) / sone < 0.003, (n, sone)


def test_silence_and_monotonicity():
    assert oracle.zwicker_sone(quiet())[0] == 0.0
    one = quiet()
    values = []
    for level in range(0, 100, 5):
        one[16] = level
        values.append(oracle.zwicker_sone(one)[0])
    assert all(b >= a for a, b in zip(values, values[1:]))
