"""Pins the surrogate network's CPU oracle (oracle/mlp.py): back-propagation against central finite
differences of the loss in exact FP64, and the BF16 rounding against hand-checked bit patterns."""
import numpy as np

from oracle import mlp


def layers_for(n_in, width, depth, n_out):
    dims = [n_in] + [width] * depth + [n_out]
    out, off = [], 0
    for l in range(depth + 1):
        o, i = dims[l + 1], dims[l]
        out.append((off, off + o * i, o, i))
        off += o * i + o
    return out, off


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.0e38, 0.0], np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # exactly half an ulp (2^-7) above 1: ties to the even 1.0
    assert r[2] == 1.0 + 2 ** -6  # 1.5 ulp: ties to even rounds up
    assert r[3] == 1.0
    assert abs(r[4] / -3.0e38 - 1) < 2 ** -8
    assert r[5] == 0.0


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    layers, n_params = layers_for(5, 8, 2, 3)
    params = rng.normal(0, 0.5, n_params)
    x = rng.normal(size=(7, 5))
    t = rng.normal(size=(7, 3))
    grad, _ = mlp.backward_mse(params, layers, x, t, 3, rounding=False)
    fd = np.zeros(n_params)
    for i in range(n_params):
        e = np.zeros(n_params)
        e[i] = 1e-6
        fd[i] = (mlp.loss(params + e, layers, x, t, 3) - mlp.loss(params - e, layers, x, t, 3)) / 2e-6
    assert np.max(np.abs(grad - fd)) < 1e-7 * max(1.0, np.abs(fd).max())


def test_padding_outputs_carry_no_gradient():
    rng = np.random.default_rng(1)
    layers, n_params = layers_for(4, 8, 1, 6)
    params = rng.normal(0, 0.5, n_params)
    x = rng.normal(size=(5, 4))
    t = rng.normal(size=(5, 6))
    grad, _ = mlp.backward_mse(params, layers, x, t, 4)
    w, b, o, i = layers[-1]
    assert np.all(grad[w:w + o * i].reshape(o, i)[4:] == 0) and np.all(grad[b + 4:b + o] == 0)
