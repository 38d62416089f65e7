"""The FP64 restatement of the surrogate network against what the mathematics fixes: BF16
rounding at ties, GELU's derivative, the backward pass against central differences, and Adam's
first step (a step of the learning rate against the gradient's sign)."""
import numpy as np

from oracle import mlp
from quietmap_amd import surrogate


def test_bf16_rounds_to_nearest_even():
    one = 1.0
    assert mlp.bf16(one) == 1.0
    assert mlp.bf16(1.0 + 2.0 ** -8) == 1.0  # tie, 1.0 is even
    assert mlp.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie, rounds up to even
    assert mlp.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7  # above the tie
    assert mlp.bf16(-2.5) == -2.5


def test_gelu_gradient_matches_central_differences():
    z = np.linspace(-6, 6, 121)
    h = 1e-6
    numeric = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(numeric - mlp.gelu_grad(z))) < 1e-8


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(11)
    params = surrogate.init_params(5).astype(np.float64)
    params += rng.normal(0, 0.05, params.size) * (params == 0)  # non-zero biases
    x = rng.normal(0, 1, (6, mlp.SHAPES[0][1]))
    t = rng.normal(0, 1, (6, mlp.SHAPES[-1][0]))
    loss, grad = mlp.loss_and_gradients(params, x, t, rounding=False)
    # Parameters from every layer's weights and biases.
    offsets, k = [], 0
    for o, i in mlp.SHAPES:
        offsets.append((k, k + o * i, k + o * i + o))
        k += o * i + o
    picks = []
    for w0, b0, end in offsets:
        picks += list(rng.integers(w0, b0, 6)) + list(rng.integers(b0, end, 2))
    h = 1e-6
    for p in picks:
        up, down = params.copy(), params.copy()
        up[p] += h
        down[p] -= h
        lu, _ = mlp.loss_and_gradients(up, x, t, rounding=False)
        ld, _ = mlp.loss_and_gradients(down, x, t, rounding=False)
        numeric = (lu - ld) / (2 * h)
        assert abs(numeric - grad[p]) <= 1e-6 * max(1.0, abs(grad[p])), (p, numeric, grad[p])


def test_adam_first_step_moves_by_the_learning_rate():
    g = np.array([3.0, -0.5, 1e-3, -7.0])
    p, m, v = mlp.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), step=1, lr=0.01)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-5)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_parameter_count_and_initialisation():
    assert mlp.PARAMS == surrogate.PARAMS == 209_928
    p = surrogate.init_params(1)
    assert p.dtype == np.float32 and p.size == mlp.PARAMS
    w, b = mlp.unpack(p)[1]
    assert np.all(b == 0) and abs(w).max() <= np.sqrt(6 / 256)
