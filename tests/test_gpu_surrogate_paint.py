"""The surrogate's samples and painter on the GPU.

- Samples: the targets are the exact boundary terms, against the FP64 oracle's (0.05 dB); the
  inputs against a NumPy restatement of the feature definition over the oracle's ground and
  crossings.
- Painter: every receiver's energy against the same mixing recomputed on the host from the
  network's outputs for its pairs (and the exact pair energies outside the envelope).
- Gain: a network trained by the device kernels on one scene predicts the exact boundary terms of
  another scene far better than the best constant (the method without it), and its painted Lden
  lies far closer to the exact painter's.
"""
import math

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

LN10 = math.log(10.0)


@pytest.fixture(scope="module")
def gpu():
    from quietmap_amd.painter import Gpu

    g = Gpu(0)
    yield g
    g.close()


def features_reference(sc, point, altitude, source):
    """The 40 inputs and the envelope of one path, restated over the oracle's scene functions."""
    T, O = sc.scene.terrain, sc.scene.obstacles
    sx, sy, rx, ry = source["x_m"], source["y_m"], point["x_m"], point["y_m"]
    H = max(math.hypot(sx - rx, sy - ry), 1.0)
    failed, cr = oracle.crossings(O, (sx, sy), (rx, ry))
    keep = np.ones(cr.size, bool)
    if point["own_footprint"]:
        keep &= cr["footprint_id"] != point["own_footprint"]
    if source["footprint_id"]:
        marked = np.zeros(cr.size, bool)
        for i in range(cr.size):
            if not keep[i]:
                continue
            fid = cr["footprint_id"][i]
            count = np.sum(keep & (cr["bounds_roof"] != 0) & (cr["footprint_id"] == fid))
            marked[i] = bool(count % 2) or fid == source["footprint_id"]
        keep &= ~marked
    cr = cr[keep]
    _, s0 = oracle.ground_at(T, sx, sy)
    _, r1 = oracle.ground_at(T, rx, ry)
    source_alt = s0[0] + source["height_m"]
    slope = (altitude - source_alt) / H
    f = np.zeros(40)
    f[0] = H / 1000
    f[1] = source["height_m"] / 4
    f[2] = (altitude - r1[0]) / 4
    forest = s0[2] > 0 or r1[2] > 0
    g_sum = np.zeros(4)
    highest = -np.inf
    for j in range(48):
        t = (j + 0.5) / 48
        _, g = oracle.ground_at(T, sx + (rx - sx) * t, sy + (ry - sy) * t)
        above = (g[0] - (source_alt + slope * t * H)) / 10
        k = 4 + j // 4
        f[k] = max(f[k], above) if j % 4 else above
        if above > highest:
            highest, f[3] = above, t
        g_sum[j // 12] += g[1]
        forest = forest or g[2] > 0
    f[16:20] = g_sum / 12
    tops, roofed, open_fp = [], 0.0, None
    for c in cr:
        x = c["t"] * H
        _, g = oracle.ground_at(T, sx + (rx - sx) * c["t"], sy + (ry - sy) * c["t"])
        tops.append((x, (g[0] + c["height_m"] - (source_alt + slope * x)) / 10))
        if c["bounds_roof"]:
            if open_fp is not None and c["footprint_id"] == open_fp[0]:
                roofed += x - open_fp[1]
                open_fp = None
            elif open_fp is None:
                open_fp = (c["footprint_id"], x)
    kept = sorted(sorted(range(len(tops)), key=lambda i: -tops[i][1])[:8]) if len(tops) > 8 else range(len(tops))
    kept = list(kept)
    for k in range(8):
        f[20 + 2 * k] = tops[kept[k]][0] / H if k < len(kept) else 0.0
        f[21 + 2 * k] = tops[kept[k]][1] if k < len(kept) else -3.0
    f[36] = len(tops) / 8
    f[38] = source["ground_percent"] / 100 if source["ground_percent"] != 255 else s0[1]
    f[39] = min(roofed / H, 1.0)
    envelope = failed == 0 and H >= 500 and len(tops) <= 8 and not forest
    return f, envelope


def random_pairs(sc, n, seed, min_m=0.0):
    rng = np.random.default_rng(seed)
    pp = rng.integers(0, sc.points.size, 20 * n)
    ps = rng.integers(0, sc.sources.size, 20 * n)
    d = np.hypot(sc.sources["x_m"][ps] - sc.points["x_m"][pp], sc.sources["y_m"][ps] - sc.points["y_m"][pp])
    ok = (d <= sc.cutoff_m) & (d >= min_m)
    return pp[ok][:n].astype(np.uint32), ps[ok][:n].astype(np.uint32)


def test_samples_match_the_oracle(gpu):
    from quietmap_amd import hip, scenes, surrogate

    print("case: surrogate samples tile scale 0.05", flush=True)
    sc = scenes.tile_scene(scale=0.05)
    pp, ps = random_pairs(sc, 3000, 7)
    square = gpu.upload(sc.scene, sc.sources)
    d_points = hip.DeviceBuffer.from_array(sc.points)
    d_rec = hip.DeviceBuffer(sc.points.size * 256)
    gpu.receivers_device(square, d_points.ptr, sc.points.size, d_rec.ptr)
    n = pp.size
    d_x, d_t, d_f = hip.DeviceBuffer(n * 320), hip.DeviceBuffer(n * 64), hip.DeviceBuffer(n * 4)
    d_pp, d_ps = hip.DeviceBuffer.from_array(pp), hip.DeviceBuffer.from_array(ps)
    surrogate.samples_device(gpu, square, d_rec.ptr, d_pp.ptr, d_ps.ptr, n, d_x.ptr, d_t.ptr, d_f.ptr)
    hip.synchronize()
    x = d_x.to_array(np.float32, (n, 2, 40))
    t = d_t.to_array(np.float32, (n, 2, 8))
    env = d_f.to_array(np.uint32, n) != 0
    want, failed = oracle.pair_boundaries(sc.scene, sc.points, sc.sources, pp, ps)
    ok = failed == 0
    err = np.abs(t[ok] - want[ok])
    print(f"targets: n={ok.sum()} in envelope={env.sum()} max={err.max():.5f} dB p99.9={np.quantile(err, 0.999):.5f}",
          flush=True)
    assert err.max() <= 0.05
    assert np.all(x[:, 1, 37] == 1.0) and np.all(x[:, 0, 37] == 0.0)
    assert np.array_equal(x[:, 0, :37], x[:, 1, :37]) and np.array_equal(x[:, 0, 38:], x[:, 1, 38:])
    alt = oracle.receivers(sc.scene, sc.points)["altitude_m"]
    worst = 0.0
    for i in range(0, n, 15):
        f, e = features_reference(sc, sc.points[pp[i]], alt[pp[i]], sc.sources[ps[i]])
        worst = max(worst, float(np.max(np.abs(x[i, 0] - f))))
        assert e == bool(env[i]) or failed[i] != 0, i
    print(f"features: max |device - restatement| = {worst:.2e}", flush=True)
    assert worst < 2e-3
    square.close()


def host_energy(sc, rcv, pairs_src, att_rows, source_alt):
    """The exact path's mixing (no foliage) of per-band boundary terms, FP64, for one receiver."""
    total = np.zeros(3)
    for k, s in enumerate(pairs_src):
        src = sc.sources[s]
        H = max(math.hypot(rcv["x"] - src["x_m"], rcv["y"] - src["y_m"]), 1.0)
        dz = rcv["alt"] - source_alt[k]
        slant = max(math.hypot(H, dz), 1.0)
        az = math.atan2(rcv["y"] - src["y_m"], rcv["x"] - src["x_m"])
        bearing = (90.0 - math.degrees(az)) % 360.0
        pos = bearing * 16 / 360
        lo = int(math.floor(pos)) % 16
        fr = pos - math.floor(pos)
        p = [rcv["fav"][q][lo] + fr * (rcv["fav"][q][(lo + 1) % 16] - rcv["fav"][q][lo]) for q in range(3)]
        hd = max(math.hypot(rcv["x"] - src["x_m"], rcv["y"] - src["y_m"]), src["exclusion_radius_m"])
        d = max(math.hypot(hd, dz), 1.0)
        w = 1.0 / (d * d * 10 ** 1.1)
        for b in range(8):
            air = math.exp(-rcv["alpha"][b] * slant * 1e-3 * LN10 / 10)
            eh = air * math.exp(-att_rows[k][0][b] * LN10 / 10)
            ef = air * math.exp(-att_rows[k][1][b] * LN10 / 10)
            for q in range(3):
                total[q] += w * src["energy"][q][b] * (p[q] * ef + (1 - p[q]) * eh)
    return total


def test_surrogate_painter_mixes_as_the_exact_path(gpu):
    from quietmap_amd import hip, scenes, surrogate
    from quietmap_amd.painter import RECEIVER_BYTES

    print("case: surrogate painter plumbing tile scale 0.03", flush=True)
    sc = scenes.tile_scene(scale=0.03)
    nx, ny = sc.grid_shape
    n = nx * ny
    square = gpu.upload(sc.scene, sc.sources)
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
    energy = d_e.to_array(np.float64, (n, 3))
    paths = d_p.to_array(np.uint32, n)
    learned = d_s.to_array(np.uint32, n)
    print(f"paths {paths.sum()} network {learned.sum()}", flush=True)
    assert learned.sum() > 0.5 * paths.sum()
    rec = oracle.receivers(sc.scene, sc.points)
    rng = np.random.default_rng(2)
    worst = 0.0
    for r in rng.choice(n, 6, replace=False):
        d = np.hypot(sc.sources["x_m"] - sc.points["x_m"][r], sc.sources["y_m"] - sc.points["y_m"][r])
        near = np.flatnonzero(d <= sc.cutoff_m).astype(np.uint32)
        assert paths[r] == near.size
        m = near.size
        d_pp = hip.DeviceBuffer.from_array(np.full(m, r, np.uint32))
        d_ps = hip.DeviceBuffer.from_array(near)
        d_x, d_t, d_fl = hip.DeviceBuffer(m * 320), hip.DeviceBuffer(m * 64), hip.DeviceBuffer(m * 4)
        surrogate.samples_device(gpu, square, d_rec.ptr, d_pp.ptr, d_ps.ptr, m, d_x.ptr, d_t.ptr, d_fl.ptr)
        hip.synchronize()
        x = d_x.to_array(np.float32, (m * 2, 40))
        inside = d_fl.to_array(np.uint32, m) != 0
        att = net.infer(x).reshape(m, 2, 8)
        exact = gpu.evaluate_pairs(square, sc.points, np.full(m, r, np.uint32), near, point_bands=False)
        s_alt = np.array([oracle.ground_at(sc.scene.terrain, sc.sources["x_m"][s], sc.sources["y_m"][s])[1][0]
                          + sc.sources["height_m"][s] for s in near])
        rv = {"x": sc.points["x_m"][r], "y": sc.points["y_m"][r], "alt": rec["altitude_m"][r],
              "alpha": rec["alpha"][r], "fav": rec["favourable"][r]}
        want = host_energy(sc, rv, near[inside], att[inside], s_alt[inside])
        want += exact["sums"][~inside, :3].sum(axis=0)
        rel = np.max(np.abs(energy[r] - want) / want)
        worst = max(worst, rel)
    print(f"receiver energy vs host mixing: max relative {worst:.2e}", flush=True)
    assert worst < 2e-3
    net.close()
    square.close()


def test_trained_network_gains_over_the_constant(gpu):
    from quietmap_amd import hip, scenes, surrogate
    from quietmap_amd.painter import RECEIVER_BYTES, lden_db

    print("case: surrogate training gain", flush=True)
    train_scene = scenes.tile_scene(seed=11, scale=0.06)
    samples, _ = surrogate.generate_samples(gpu, train_scene, n_sources=400, seed=1, stride=2)
    means = surrogate.band_means(samples)
    params, history = surrogate.train(gpu, samples, surrogate.with_output_bias(surrogate.init_params(3), means),
                                      epochs=3, batch=4096, lr=1e-3, lr_end=5e-5)
    print(f"training rows {samples.rows}: loss {history[0][1]:.2f} -> {history[-1][1]:.2f} dB^2", flush=True)
    test_scene = scenes.tile_scene(seed=2, scale=0.05)
    held, _ = surrogate.generate_samples(gpu, test_scene, n_sources=200, seed=5, stride=2)
    rows = min(held.rows, 1 << 18)
    x = held.inputs.to_array(np.float32, (rows, 40))
    t = held.targets.to_array(np.float32, (rows, 8))
    net = surrogate.Mlp(gpu, params, max_batch=rows)
    y = net.infer(x)
    rmse_net = float(np.sqrt(np.mean((y - t) ** 2)))
    rmse_const = float(np.sqrt(np.mean((means[None, :] - t) ** 2)))
    print(f"held-out boundary terms: RMSE network {rmse_net:.3f} dB, constant {rmse_const:.3f} dB", flush=True)
    assert rmse_net < 0.5 * rmse_const
    # Painted Lden against the exact painter: the network and the constant in its place.
    nx, ny = test_scene.grid_shape
    n = nx * ny
    square = gpu.upload(test_scene.scene, test_scene.sources)
    d_points = hip.DeviceBuffer.from_array(test_scene.points)
    d_rec = hip.DeviceBuffer(n * RECEIVER_BYTES)
    gpu.receivers_device(square, d_points.ptr, n, d_rec.ptr)
    bufs = [hip.DeviceBuffer(n * 24), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4)]
    gpu.paint_receivers_device(square, d_rec.ptr, nx, ny, test_scene.cutoff_m, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr)
    hip.synchronize()
    exact = lden_db(bufs[0].to_array(np.float64, (n, 3)))
    const = np.zeros_like(params)
    const[-8:] = means
    errs = {}
    for name, p in (("network", params), ("constant", const)):
        m = surrogate.Mlp(gpu, p, max_batch=1)
        surrogate.paint_receivers_device(gpu, square, m, d_rec.ptr, nx, ny, test_scene.cutoff_m, bufs[0].ptr,
                                         bufs[1].ptr, bufs[2].ptr, bufs[3].ptr)
        hip.synchronize()
        got = lden_db(bufs[0].to_array(np.float64, (n, 3)))
        e = np.abs(got - exact)
        errs[name] = (float(np.mean(e)), float(np.quantile(e, 0.999)), float(e.max()))
        m.close()
    print(f"painted Lden vs exact (mean, p99.9, max dB): network {errs['network']}, constant {errs['constant']}",
          flush=True)
    assert errs["network"][0] < 0.5 * errs["constant"][0]
    net.close()
    square.close()
