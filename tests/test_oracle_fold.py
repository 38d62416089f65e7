"""The oracle's map folding (paint/src/sone_fold.rs fold_placed_for_map with
popup/src/percentiles/weather.rs merge_close_states, then paint/src/sone.rs write_packed) pinned by
the reference's own tests restated on the same inputs, by closed forms of its three folding rules
and by the conservation of every period's energy."""
import numpy as np
import pytest

from oracle import oracle
from quietmap_amd import pixels

LINE = 43
CURVES = np.full((7, 3, 28), np.nan)


def steady_line(means, profile=None):
    """Line::steady (popup/src/percentiles/lines.rs:80-97) in the packed f64 layout."""
    means = np.broadcast_to(np.asarray(means, np.float64), (3, 10))
    energy = (means > 0).any(axis=1).astype(np.float64)
    return pixels.pack_line_f64(energy, means, [np.inf] * 3, profile=profile)


def unfolded_pixel(layers):
    """One unfolded pixel: per layer (steady per period, [lines]); six layers, the rest empty."""
    parts = []
    for layer in range(6):
        steady, lines = layers[layer] if layer < len(layers) else (np.zeros(3), [])
        parts.append(np.concatenate([np.asarray(steady, np.float64), [float(len(lines))]]))
        parts.extend(np.asarray(l, np.float64) for l in lines)
    parts.append(np.nan_to_num(CURVES, nan=np.nan).reshape(-1))
    return np.concatenate(parts)


def fold_one(layers, **kw):
    values = unfolded_pixel(layers)
    return oracle.fold_pixels(values, np.zeros(1, np.uint64), f64=True, **kw)


def read_layers(values):
    """(steady, lines) per layer of one packed pixel."""
    out, at = [], 0
    for _ in range(6):
        steady = values[at:at + 3]
        n = int(values[at + 3])
        at += 4
        out.append((steady, values[at:at + n * LINE].reshape(n, LINE)))
        at += n * LINE
    return out


def test_merge_close_states_reference_case():
    """popup/src/percentiles_tests.rs:470-505: at spread 1.2 the far line splits the states at 3,
    the near one's first three take their mean; lines apart by more than the spread keep theirs."""
    def line(means):
        means = np.array(means)
        return pixels.pack_line_f64([means.sum() / 10] * 3, np.tile(means, (3, 1)), [1e6] * 3)

    near = line([1.10, 1.05, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    far = line([4.0, 4.0, 4.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    merged_near, merged_far = oracle.merge_close_states([near[None], far[None]], 1.2)
    mn, mf = merged_near[0, 3:13], merged_far[0, 3:13]
    assert np.array_equal(mf, far[3:13])
    assert mn[0] == mn[2]
    assert abs(mn[0] - (1.10 + 1.05 + 1.0) / 3.0) < 1e-12
    assert np.array_equal(mn[3:], near[3 + 3:13])
    for merged in (merged_near, merged_far):
        assert abs(merged[0, 3:13].sum() / 10 - merged[0, 0]) < 1e-12
    apart = line([2.0, 1.5, 1.0, 0.7, 0.5, 0.35, 0.25, 0.18, 0.12, 0.09])
    (kept,) = oracle.merge_close_states([apart[None]], 1.2)
    assert np.array_equal(kept[0], apart, equal_nan=True)


def test_collapse_keeps_exact_mean_across_states():
    """paint/src/sone.rs:783-816: an unprofiled steady line whose states compare equal in f32 joins
    the steady energy with its exact f64 mean (1,000,000.01), not state 0's."""
    means = np.full((3, 10), 1_000_000.0)
    guard = np.full((3, 10), 100.0)
    means[:, 5:] = 1_000_000.02
    guard[:, 5:] = 10_000.0
    out, used, out64 = fold_one([(np.zeros(3), [steady_line(means), steady_line(guard, profile=0)])])
    (steady, lines), *_ = read_layers(out64)
    assert len(lines) == 1 and lines[0][36] == 0.0
    assert np.all(np.abs(steady - 1_000_000.01) < 1e-8), steady
    # Mutation: summing state 0 only would give 1,000,000.0.
    assert np.all(np.abs(steady - 1_000_000.0) > 5e-3)


def test_weak_duty_folds_like_steady():
    """paint/src/sone.rs:669-778 at the fold: a 4 % source on duty 0.5 beside a 96 % steady floor
    (on-peak 8 % <= 12 %, standard deviation 4 % <= 8 %) packs to the same bits as the steady one."""
    floor = np.full(3, 1_209_600.0)
    mean = np.full((3, 10), 50_400.0)
    duty = pixels.pack_line_f64([50_400.0] * 3, mean, [np.inf] * 3, duty=[0.5] * 3)
    steady = pixels.pack_line_f64([50_400.0] * 3, mean, [np.inf] * 3)
    a, ua = fold_one([(np.zeros(3), []), (floor, [duty])])[:2]
    b, ub = fold_one([(np.zeros(3), []), (floor, [steady])])[:2]
    assert ua[0] == ub[0]
    assert np.array_equal(a[:ua[0]].view(np.uint32), b[:ub[0]].view(np.uint32))
    (_, none), (s, lines), *_ = read_layers(a)
    assert len(lines) == 0 and np.all(s == np.float32(1_260_000.0))


@pytest.mark.parametrize("margin,kept", [(1.001, 0), (0.999, 1)])
def test_kurze_held_rule(margin, kept):
    """A Kurze line folds when its loud moments (the intensity exceeded 0.1 % of the time) stay
    below 5 % of the layer's held level (floor + the Kurze lines' means): at lambda 0.5 the floor
    at which m x I(0.999) = 0.05 (floor + m) parts folded from kept."""
    lam, m = 0.5, 1000.0
    exc = oracle.relative_intensity(lam, 0.999)
    floor = m * exc / 0.05 - m
    line = pixels.pack_line_f64([m] * 3, np.full((3, 10), m), [lam] * 3)
    out, used, out64 = fold_one([(np.full(3, floor * margin), [line])])
    (steady, lines), *_ = read_layers(out64)
    assert len(lines) == kept
    want = floor * margin + (m if kept == 0 else 0.0)
    assert np.allclose(steady, want, rtol=1e-14)


@pytest.mark.parametrize("d,count,folds", [(0.5, 12, 6), (0.35, 8, 2), (0.05, 10, 0)])
def test_duty_budget_closed_form(d, count, folds):
    """Identical weak events fold while their summed variance (m / total)^2 (1 - d) / d stays within
    0.08^2 of the layer: k = floor(0.0064 (total / m)^2 d / (1 - d)); events whose on-peak share
    m / d exceeds 12 % of the total never fold."""
    floor, m = 100_000.0, 5_000.0
    total = floor + count * m
    event = pixels.pack_line_f64([m] * 3, np.full((3, 10), m), [np.inf] * 3, duty=[d] * 3)
    out, used, out64 = fold_one([(np.full(3, floor), [event] * count)])
    (steady, lines), *_ = read_layers(out64)
    var = (m / total) ** 2 * (1 - d) / d
    k = 0 if m / d > 0.12 * total else min(count, int(np.floor(0.0064 / var)))
    assert k == folds
    # The folded events form one steady line of equal states, which joins the floor.
    assert len(lines) == count - k, (len(lines), count, k)
    assert np.allclose(steady, floor + k * m, rtol=1e-12)
    assert np.all(lines[:, 37:40] == d)


def layer_energy(values):
    """Per layer and period: 10 x steady + every line's positive means in the periods where it
    sounds (energy > 0, and a duty above 0 for an event)."""
    q = np.zeros((6, 3))
    for l, (steady, lines) in enumerate(read_layers(values)):
        q[l] = 10.0 * steady
        for line in lines:
            means = line[3:33].reshape(3, 10)
            sounds = line[0:3] > 0
            if not np.isnan(line[37]):
                sounds &= line[37:40] > 0
            q[l] += np.where(sounds, np.where(means > 0, means, 0).sum(axis=1), 0.0)
    return q


def test_fold_conserves_energy_and_never_grows():
    values, offsets = pixels.synthetic_unfolded(seed=3, count=96)
    out, used, out64 = oracle.fold_pixels(values, offsets, f64=True)
    sizes = np.diff(np.append(offsets, values.size)).astype(np.int64)
    used = used.astype(np.int64)
    assert np.all(used <= sizes)
    lines_in = lines_out = 0
    for i, (o, n) in enumerate(zip(offsets.astype(np.int64), sizes)):
        a, b = values[o:o + n], out64[o:o + used[i]]
        assert np.allclose(layer_energy(a), layer_energy(b), rtol=1e-12, atol=0)
        assert np.array_equal(a[n - 7 * 3 * 28:], b[used[i] - 7 * 3 * 28:], equal_nan=True)
        lines_in += sum(len(x[1]) for x in read_layers(a))
        lines_out += sum(len(x[1]) for x in read_layers(b))
        # The f32 pixel is the f64 one cast (duty and passes whole or NaN).
        assert np.array_equal(out[o:o + used[i]], b.astype(np.float32), equal_nan=True)
    assert lines_out < lines_in
