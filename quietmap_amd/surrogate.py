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
