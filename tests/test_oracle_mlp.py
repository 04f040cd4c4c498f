"""Pins the FP64 MLP oracle (oracle/mlp.py) by mathematics: finite
differences for the gradient, weight patterns with closed-form outputs for
the forward pass, the bf16 rounding rule on exact cases and Adam's first
step."""
import math

import numpy as np

from oracle import mlp


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((B, 6)), 0.3 * rng.standard_normal(B)


def test_gradient_matches_central_differences():
    p = mlp.init_params(seed=3, scale=1.5)
    x, t = _batch(37, 4)
    L, g = mlp.loss_and_grad(p, x, t)
    assert abs(L - mlp.loss(p, x, t)) < 1e-15
    rng = np.random.default_rng(5)
    live = np.r_[0:6 * 64, mlp.OFF_B1:mlp.N_PARAMS]
    idx = np.r_[rng.choice(live, 120, replace=False), mlp.OFF_B3,
                mlp.OFF_W3, mlp.OFF_B2, mlp.OFF_B1]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mlp.loss(pp, x, t) - mlp.loss(pm, x, t)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i
    # the padding rows of W1 see zero inputs
    assert np.all(g[6 * 64:8 * 64] == 0)


def test_forward_single_path_closed_form():
    p = np.zeros(mlp.N_PARAMS)
    W1 = np.zeros((8, 64))
    W1[2, 5] = 0.7
    p[mlp.OFF_W1:mlp.OFF_B1] = W1.reshape(-1)
    p[mlp.OFF_B1 + 5] = -0.1
    W2 = np.zeros((64, 64))
    W2[5, 9] = 1.3          # input unit 5 -> output unit 9 (not 9 -> 5)
    p[mlp.OFF_W2:mlp.OFF_B2] = W2.reshape(-1)
    p[mlp.OFF_B2 + 9] = 0.2
    p[mlp.OFF_W3 + 9] = -0.8
    p[mlp.OFF_B3] = 0.25
    xs = np.array([[0.1, -2.0, 0.5, 3.0, 0.0, 1.0],
                   [0.0, 0.0, -1.5, 0.0, 7.0, 0.0]])
    y = mlp.forward(p, xs)
    for k in range(2):
        want = 0.25 - 0.8 * math.tanh(1.3 * math.tanh(0.7 * xs[k, 2] - 0.1)
                                      + 0.2)
        assert abs(y[k] - want) < 1e-15


def test_forward_odd_without_biases():
    p = mlp.init_params(seed=8)
    for off, n in ((mlp.OFF_B1, 64), (mlp.OFF_B2, 64), (mlp.OFF_B3, 1)):
        p[off:off + n] = 0
    x, _ = _batch(11, 9)
    assert np.allclose(mlp.forward(p, -x), -mlp.forward(p, x), atol=1e-15)


def test_bf16_rounding_exact_cases():
    one = 1.0
    ulp = 2.0 ** -7
    vals = np.array([one, one + ulp / 2, one + 1.5 * ulp, -3.0, 0.0,
                     one + 0.49 * ulp, 2.0 ** -130])
    want = np.array([one, one, one + 2 * ulp, -3.0, 0.0, one,
                     np.float32(2.0 ** -130)])
    got = mlp.to_bf16(vals)
    assert got[0] == want[0] and got[1] == want[1] and got[2] == want[2]
    assert got[3] == -3.0 and got[4] == 0.0 and got[5] == one
    # subnormal float32 input keeps its top bits
    assert 0 < got[6] <= want[6]


def test_bf16_forward_close_to_fp64():
    p = mlp.init_params(seed=1)
    x, _ = _batch(64, 2)
    d = np.abs(mlp.forward(p, x, bf16=True) - mlp.forward(p, x))
    assert d.max() < 3e-2 and d.max() > 0


def test_adam_first_step_is_signed_lr():
    p = mlp.init_params(seed=4)
    g = np.random.default_rng(6).standard_normal(mlp.N_PARAMS)
    m, v = np.zeros_like(p), np.zeros_like(p)
    q = mlp.adam(p.copy(), m, v, g, 1, lr=1e-3)
    live = np.ones(mlp.N_PARAMS, bool)
    live[6 * 64:8 * 64] = False
    step = (p - q)[live]
    gl = g[live]
    assert np.allclose(step, 1e-3 * gl / (np.abs(gl) + 1e-8), rtol=0,
                       atol=1e-13)
    assert np.all(q[~live] == p[~live])
