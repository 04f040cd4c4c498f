"""FP64 numpy restatement of the learned smoothing-length predictor.

TEST INFRASTRUCTURE ONLY: the checker of the HIP MLP kernels
(flecsph_amd/csrc/k_mlp.hip).  The predictor has no counterpart in the
reference (it replaces nothing there: the reference updates h in closed form,
include/physics/default_physics.h:826-848), so this module is pinned by
mathematics rather than by reference vectors: central finite differences of
`loss` pin `loss_and_grad`, and hand-built weight patterns with known outputs
pin `forward` (tests/test_oracle_mlp.py).

Network: y = w3 . tanh(W2^T tanh(W1^T x + b1) + b2) + b3 with x of 6
features; parameters are one flat vector laid out as
W1 [8][64] (rows 6 and 7 multiply zero padding), b1 [64], W2 [64][64]
(row = input unit, column = output unit), b2 [64], w3 [64], b3 [1].

`bf16=True` rounds exactly where the HIP kernels hold BF16 operands (x, W1,
H1 and W2 in the forward pass; the output-layer error dZ2 and H1 in the
backward products), so the kernels' FP32 results can be compared with a
tight tolerance; `bf16=False` is the exact FP64 model.
"""
import numpy as np

N_IN, N_INP, N_H = 6, 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_INP * N_H
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H
N_PARAMS = OFF_B3 + 1


def to_bf16(a):
    """Rounds float32 values to bfloat16 (round to nearest even), returned
    as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(N_INP, N_H)[:N_IN]
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_H, N_H)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def _layers(p, x, bf16):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, N_IN)
    if bf16:
        Z1 = to_bf16(x) @ to_bf16(W1) + b1
        H1 = to_bf16(np.tanh(Z1))
        Z2 = H1 @ to_bf16(W2) + b2
    else:
        Z1 = x @ W1 + b1
        H1 = np.tanh(Z1)
        Z2 = H1 @ W2 + b2
    H2 = np.tanh(Z2)
    y = H2 @ w3 + b3
    return x, H1, H2, y


def forward(p, x, bf16=False):
    return _layers(p, x, bf16)[3]


def loss(p, x, t, bf16=False):
    y = forward(p, x, bf16)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def loss_and_grad(p, x, t, bf16=False):
    """L = mean_b (y_b - t_b)^2 and dL/dp (flat, N_PARAMS; the padding rows
    of W1 get zero)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x, H1, H2, y = _layers(p, x, bf16)
    t = np.asarray(t, np.float64).reshape(-1)
    B = len(t)
    e = y - t
    L = float(np.mean(e * e))
    dy = 2.0 * e / B                          # dL/dy
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = H2.T @ dy
    g[OFF_B3] = dy.sum()
    dZ2 = dy[:, None] * w3[None, :] * (1.0 - H2 * H2)
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    dZ2m = to_bf16(dZ2) if bf16 else dZ2      # operand of the two products
    W2m = to_bf16(W2) if bf16 else W2
    g[OFF_W2:OFF_B2] = (H1.T @ dZ2m).reshape(-1)
    dH1 = dZ2m @ W2m.T
    dZ1 = dH1 * (1.0 - H1 * H1)
    gW1 = np.zeros((N_INP, N_H))
    gW1[:N_IN] = x.T @ dZ1
    g[OFF_W1:OFF_B1] = gW1.reshape(-1)
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    return L, g


def init_params(seed=0, scale=1.0):
    """Uniform(+-scale/sqrt(fan_in)) weights, small random biases; the
    padding rows of W1 are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS)
    W1 = np.zeros((N_INP, N_H))
    W1[:N_IN] = rng.uniform(-1, 1, (N_IN, N_H)) * scale / np.sqrt(N_IN)
    p[OFF_W1:OFF_B1] = W1.reshape(-1)
    p[OFF_B1:OFF_W2] = rng.uniform(-0.1, 0.1, N_H)
    p[OFF_W2:OFF_B2] = (rng.uniform(-1, 1, (N_H, N_H)) * scale /
                        np.sqrt(N_H)).reshape(-1)
    p[OFF_B2:OFF_W3] = rng.uniform(-0.1, 0.1, N_H)
    p[OFF_W3:OFF_B3] = rng.uniform(-1, 1, N_H) / np.sqrt(N_H)
    p[OFF_B3] = 0.05
    return p


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) in place; step counts
    from 1.  The padding rows of W1 are left unchanged."""
    keep = np.zeros(N_PARAMS, bool)
    keep[OFF_W1 + N_IN * N_H:OFF_B1] = True
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    upd = lr * mh / (np.sqrt(vh) + eps)
    upd[keep] = 0
    p -= upd
    return p
