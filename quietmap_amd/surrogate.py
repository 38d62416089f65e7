"""The learned excess-attenuation surrogate: its network on the device.

A network of 40 inputs, four hidden layers of 256 (tanh-form GELU) and 8 outputs stands in for
a path state's ground + diffraction attenuation per octave band (physics/src/cnossos/
diffraction.rs:64-196 computes the exact terms). Training and inference run in libqmpaint's BF16
MFMA kernels (include/qm_surrogate.h); this module holds the handle and the parameter
initialisation.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import hip
from .painter import PainterError, _check, library

INPUTS, HIDDEN, OUTPUTS = 40, 256, 8
SHAPES = [(HIDDEN, INPUTS), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (OUTPUTS, HIDDEN)]
PARAMS = sum(o * i + o for o, i in SHAPES)  # 209,928


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases, from a seeded generator; FP32, flat."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in SHAPES:
        bound = math.sqrt(6.0 / i)
        parts.append(rng.uniform(-bound, bound, o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)


class Mlp:
    """The network on a Gpu, with training buffers for `max_batch` rows a step (0: inference)."""

    def __init__(self, gpu, params, max_batch: int = 0):
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != PARAMS:
            raise PainterError(f"expected {PARAMS} parameters, got {params.size}")
        self.gpu, self.max_batch = gpu, int(max_batch)
        self._lib = library()
        err = C.create_string_buffer(512)
        handle = C.c_void_p()
        _check(self._lib.qm_mlp_new(gpu.handle, params.ctypes.data, self.max_batch, C.byref(handle), err, 512), err)
        self.handle = handle

    def close(self):
        if self.handle:
            self._lib.qm_mlp_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.empty(PARAMS, np.float32)
        err = C.create_string_buffer(512)
        _check(self._lib.qm_mlp_params(self.handle, out.ctypes.data, err, 512), err)
        return out

    def gradient_device(self, d_inputs, d_targets, rows, d_gradient, stream=None) -> float:
        loss = C.c_float()
        err = C.create_string_buffer(512)
        _check(self._lib.qm_mlp_gradient_device(self.handle, d_inputs, d_targets, rows, d_gradient, C.byref(loss),
                                                stream, err, 512), err)
        return float(loss.value)

    def train_step_device(self, d_inputs, d_targets, rows, lr, stream=None, want_loss=False):
        loss = C.c_float()
        err = C.create_string_buffer(512)
        _check(self._lib.qm_mlp_train_step_device(self.handle, d_inputs, d_targets, rows, C.c_float(lr),
                                                  C.byref(loss) if want_loss else None, stream, err, 512), err)
        return float(loss.value) if want_loss else None

    def infer_device(self, d_inputs, rows, d_outputs, stream=None):
        err = C.create_string_buffer(512)
        _check(self._lib.qm_mlp_infer_device(self.handle, d_inputs, rows, d_outputs, stream, err, 512), err)

    # Host-array conveniences (copies in and out).
    def gradient(self, inputs, targets):
        x = hip.DeviceBuffer.from_array(np.ascontiguousarray(inputs, np.float32))
        t = hip.DeviceBuffer.from_array(np.ascontiguousarray(targets, np.float32))
        g = hip.DeviceBuffer(PARAMS * 4)
        loss = self.gradient_device(x.ptr, t.ptr, len(inputs), g.ptr)
        return loss, g.to_array(np.float32, PARAMS)

    def infer(self, inputs):
        x = hip.DeviceBuffer.from_array(np.ascontiguousarray(inputs, np.float32))
        y = hip.DeviceBuffer(len(inputs) * OUTPUTS * 4)
        self.infer_device(x.ptr, len(inputs), y.ptr)
        hip.synchronize()
        return y.to_array(np.float32, (len(inputs), OUTPUTS))


# ---------------------------------------------------------------------------------------------
# Samples, training and the surrogate painter.

MIN_LENGTH_M = 500.0  # the envelope's shortest path (qm_features.hpp)


def samples_device(gpu, square, d_receivers, d_pair_receiver, d_pair_source, n_pairs, d_inputs, d_targets,
                   d_in_envelope, stream=None):
    """The network's input rows and the exact boundary terms of `n_pairs` device pairs."""
    err = C.create_string_buffer(512)
    _check(library().qm_surrogate_samples_device(gpu.handle, square.handle, d_receivers, d_pair_receiver,
                                                 d_pair_source, n_pairs, d_inputs, d_targets, d_in_envelope, stream,
                                                 err, 512), err)


def paint_receivers_device(gpu, square, mlp, d_receivers, nx, ny, cutoff_m, d_energy, d_paths, d_failures,
                           d_surrogate_paths=None, stream=None):
    """The receiver grid painted with the network inside its envelope (qm_paint_receivers_device's
    outputs, plus the paths the network answered per receiver)."""
    err = C.create_string_buffer(512)
    _check(library().qm_surrogate_paint_receivers_device(gpu.handle, square.handle, mlp.handle, d_receivers, nx, ny,
                                                         cutoff_m, d_energy, d_paths, d_failures, d_surrogate_paths,
                                                         stream, err, 512), err)


def source_major_pairs(scene, n_sources, seed, stride=8, min_m=MIN_LENGTH_M):
    """Pairs for training: `n_sources` random sources, each with the receivers of every `stride`-th
    grid row and column between `min_m` and the cutoff, source-major (a wave's lanes share the source
    and hold neighbouring receivers). Returns (receiver point indices, source indices)."""
    rng = np.random.default_rng(seed)
    gx, gy = scene.grid_shape
    sub = (np.arange(0, gy, stride)[:, None] * gx + np.arange(0, gx, stride)[None, :]).reshape(-1)
    px, py = scene.points["x_m"][sub], scene.points["y_m"][sub]
    srcs = rng.choice(scene.sources.size, size=min(n_sources, scene.sources.size), replace=False)
    pr, ps = [], []
    for s in np.sort(srcs):
        d = np.hypot(px - scene.sources["x_m"][s], py - scene.sources["y_m"][s])
        near = sub[(d >= min_m) & (d <= scene.cutoff_m)]
        pr.append(near.astype(np.uint32))
        ps.append(np.full(near.size, s, np.uint32))
    return np.concatenate(pr), np.concatenate(ps)


class Samples:
    """Training rows on the device: inputs (rows x 40), targets (rows x 8), in-envelope pairs only."""

    def __init__(self, inputs, targets, rows):
        self.inputs, self.targets, self.rows = inputs, targets, rows


def generate_samples(gpu, scene, n_sources, seed, stride=8):
    """Exact-kernel samples of a scene's in-envelope paths (two rows a path: calm, favourable)."""
    square = gpu.upload(scene.scene, scene.sources)
    d_points = hip.DeviceBuffer.from_array(scene.points)
    d_receivers = hip.DeviceBuffer(scene.points.size * 256)
    gpu.receivers_device(square, d_points.ptr, scene.points.size, d_receivers.ptr)
    pr, ps = source_major_pairs(scene, n_sources, seed, stride)
    # Pass 1: the envelope flags; pass 2: the kept pairs only, contiguous.
    keep = np.zeros(pr.size, bool)
    chunk = 1 << 21
    d_x = hip.DeviceBuffer(chunk * 2 * INPUTS * 4)
    d_t = hip.DeviceBuffer(chunk * 2 * OUTPUTS * 4)
    d_f = hip.DeviceBuffer(chunk * 4)
    for a in range(0, pr.size, chunk):
        b = min(a + chunk, pr.size)
        d_pr, d_ps = hip.DeviceBuffer.from_array(pr[a:b]), hip.DeviceBuffer.from_array(ps[a:b])
        samples_device(gpu, square, d_receivers.ptr, d_pr.ptr, d_ps.ptr, b - a, d_x.ptr, d_t.ptr, d_f.ptr)
        hip.synchronize()
        keep[a:b] = d_f.to_array(np.uint32, b - a) != 0
    pr, ps = pr[keep], ps[keep]
    n = pr.size
    inputs = hip.DeviceBuffer(max(n, 1) * 2 * INPUTS * 4)
    targets = hip.DeviceBuffer(max(n, 1) * 2 * OUTPUTS * 4)
    for a in range(0, n, chunk):
        b = min(a + chunk, n)
        d_pr, d_ps = hip.DeviceBuffer.from_array(pr[a:b]), hip.DeviceBuffer.from_array(ps[a:b])
        samples_device(gpu, square, d_receivers.ptr, d_pr.ptr, d_ps.ptr, b - a, inputs.ptr + a * 2 * INPUTS * 4,
                       targets.ptr + a * 2 * OUTPUTS * 4, d_f.ptr)
    hip.synchronize()
    square.close()
    return Samples(inputs, targets, 2 * n), (pr, ps)


def band_means(samples, limit=1 << 20):
    """Per-band mean of the targets (of the first `limit` rows)."""
    rows = min(samples.rows, limit)
    t = samples.targets.to_array(np.float32, (rows, OUTPUTS))
    return t.astype(np.float64).mean(axis=0)


def train(gpu, samples, params, epochs=2, batch=8192, lr=1e-3, lr_end=2e-5, seed=0, log=None):
    """Adam over the samples' rows in shuffled batches (a fixed permutation of batch starts per
    epoch), the learning rate decaying geometrically from `lr` to `lr_end`; returns the trained
    FP32 parameters and the loss at every 100th step."""
    net = Mlp(gpu, params, max_batch=batch)
    rng = np.random.default_rng(seed)
    starts = np.arange(0, samples.rows - batch + 1, batch)
    steps = epochs * starts.size
    history = []
    k = 0
    for _ in range(epochs):
        for s in rng.permutation(starts):
            rate = lr * (lr_end / lr) ** (k / max(steps - 1, 1))
            want = k % 100 == 0 or k == steps - 1
            loss = net.train_step_device(samples.inputs.ptr + int(s) * INPUTS * 4,
                                         samples.targets.ptr + int(s) * OUTPUTS * 4, batch, rate, want_loss=want)
            if want:
                history.append((k, loss))
                if log:
                    log(f"step {k}/{steps} lr {rate:.2e} loss {loss:.4f} dB^2")
            k += 1
    out = net.params()
    net.close()
    return out, history


def with_output_bias(params, bias):
    """`params` with the output layer's biases set to `bias` (8 values)."""
    p = np.array(params, np.float32)
    p[-OUTPUTS:] = np.asarray(bias, np.float32)
    return p
