"""TEST INFRASTRUCTURE — FP64 restatement of the excess-attenuation surrogate network.

Only the tests and ``__graft_entry__.smoke()`` import this module, to check the HIP forward,
backward and Adam kernels against it; the product path (``quietmap_amd``) never does.

The network (no reference counterpart: the reference has no learned part) is
40 -> 256 -> 256 -> 256 -> 256 -> 8 with tanh-form GELU after each hidden layer. Parameters are one
flat FP32 vector, per layer the weight matrix [out][in] row-major followed by the bias [out].

The device multiplies BF16 operands with FP32 accumulation. With ``rounding=True`` this restatement
rounds to BF16 at exactly the points where the device stores a BF16 operand (the inputs, the
weights, every hidden activation, the output gradient and every pre-activation gradient) and
computes everything else in FP64, so the device must agree to FP32 accumulation-order error.
"""
from __future__ import annotations

import math

import numpy as np

SHAPES = [(256, 40), (256, 256), (256, 256), (256, 256), (8, 256)]
PARAMS = sum(o * i + o for o, i in SHAPES)  # 209,928
GELU_C = math.sqrt(2.0 / math.pi)
GELU_A = 0.044715


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    bits = f.view(np.uint32).astype(np.uint64)
    rounded = (bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    """The (W [out][in], b [out]) pairs of a flat parameter vector."""
    params = np.asarray(params, dtype=np.float64)
    out, k = [], 0
    for o, i in SHAPES:
        w = params[k:k + o * i].reshape(o, i)
        k += o * i
        b = params[k:k + o]
        k += o
        out.append((w, b))
    assert k == PARAMS
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in layers])


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def forward(params, x, rounding=True):
    """Outputs [rows][8] and the cache the backward pass uses."""
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params)
    a = r(x)
    cache = {"inputs": a, "z": [], "a": [a]}
    for li, (w, b) in enumerate(layers):
        z = a @ r(w).T + b
        if li + 1 < len(layers):
            cache["z"].append(z)
            a = r(gelu(z))
            cache["a"].append(a)
        else:
            return z, cache
    raise AssertionError


def loss_and_gradients(params, x, targets, rounding=True):
    """Mean squared error over rows and outputs and its gradient (flat, like `params`)."""
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params)
    y, cache = forward(params, x, rounding)
    n = y.size
    diff = y - np.asarray(targets, dtype=np.float64)
    loss = float(np.sum(diff * diff) / n)
    g = r(2.0 * diff / n)  # dL/dy as the device stores it
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        w, _ = layers[li]
        a_prev = cache["a"][li]
        grads[li] = (g.T @ a_prev, g.sum(axis=0))
        if li == 0:
            break
        da = g @ r(w)
        g = r(da * gelu_grad(cache["z"][li - 1]))
    return loss, pack(grads)


def adam(params, grads, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (bias-corrected) in FP64; returns new (params, m, v)."""
    m = beta1 * m + (1.0 - beta1) * grads
    v = beta2 * v + (1.0 - beta2) * grads * grads
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
