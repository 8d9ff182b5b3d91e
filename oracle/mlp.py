"""CPU oracle of the dispatcher network (test infrastructure only).

The network has no counterpart in the reference crate; this is a plain float64
numpy restatement of what the HIP kernels compute, used to check their forward
outputs and gradients:

    h1 = relu(W1 x + b1)    W1: [64][n_in]
    h2 = relu(W2 h1 + b2)   W2: [64][64]
    y  = W3 h2 + b3         W3: [n_out][64]
    L  = sum_{s,o} w[s,o] (y[s,o] - t[s,o])^2 / sum w

Parameters travel as one flat vector in the order W1, b1, W2, b2, W3, b3
(row-major), the host layout of the C ABI.  Its gradient is pinned by central
finite differences (tests/test_oracle_mlp.py), not by retyping the formulas.
"""
import numpy as np

H = 64


def n_params(n_in, n_out):
    return H * n_in + H + H * H + H + n_out * H + n_out


def unpack(p, n_in, n_out):
    p = np.asarray(p, np.float64)
    shapes = [(H, n_in), (H,), (H, H), (H,), (n_out, H), (n_out,)]
    out, q = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(p[q:q + n].reshape(s))
        q += n
    assert q == p.size
    return out


def forward(p, x, n_out):
    x = np.asarray(x, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, x.shape[1], n_out)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3.T + b3
    return y, (x, z1, h1, z2, h2)


def loss(p, x, t, w):
    y, _ = forward(p, x, t.shape[1])
    w = np.asarray(w, np.float64)
    return float(np.sum(w * (y - t) ** 2) / np.sum(w))


def loss_grad(p, x, t, w):
    """(loss, flat gradient) by reverse-mode differentiation of the three layers."""
    t = np.asarray(t, np.float64)
    w = np.asarray(w, np.float64)
    y, (x, z1, h1, z2, h2) = forward(p, x, t.shape[1])
    _, _, W2, _, W3, _ = unpack(p, x.shape[1], t.shape[1])
    ws = np.sum(w)
    L = float(np.sum(w * (y - t) ** 2) / ws)
    dy = 2.0 * w * (y - t) / ws
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dz2 = (dy @ W3) * (z2 > 0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(0)
    g = np.concatenate([a.reshape(-1) for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return L, g
