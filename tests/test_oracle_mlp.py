"""Pins the dispatcher network's CPU oracle (oracle/mlp.py) by mathematics:
central finite differences of the loss, a closed-form one-hidden-path case,
and the ReLU gate.  These checks would fail for a dropped term, a transposed
weight or a wrong sign in the backward pass."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp as M  # noqa: E402


def _case(seed, n_in=5, n_out=3, B=7):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, M.n_params(n_in, n_out))
    x = rng.normal(0, 1, (B, n_in))
    t = rng.normal(0, 1, (B, n_out))
    w = (rng.random((B, n_out)) < 0.7).astype(np.float64) * rng.uniform(0.5, 2.0, (B, n_out))
    w[0, 0] = 1.0
    return p, x, t, w


def test_gradient_matches_central_differences():
    p, x, t, w = _case(1)
    L, g = M.loss_grad(p, x, t, w)
    assert abs(L - M.loss(p, x, t, w)) < 1e-12
    rng = np.random.default_rng(2)
    idx = rng.choice(p.size, 300, replace=False)
    eps = 1e-6
    for i in idx:
        d = np.zeros_like(p)
        d[i] = eps
        fd = (M.loss(p + d, x, t, w) - M.loss(p - d, x, t, w)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_every_parameter_block_gets_gradient():
    p, x, t, w = _case(3)
    _, g = M.loss_grad(p, x, t, w)
    blocks = M.unpack(g, 5, 3)
    for b in blocks:
        assert np.abs(b).sum() > 0


def test_closed_form_single_path():
    """One input, one active hidden path: y = w3 * relu(w2 * relu(w1 x + b1) + b2) + b3,
    with every other weight zero; the gradient follows by hand."""
    n_in, n_out = 1, 1
    p = np.zeros(M.n_params(n_in, n_out))
    W1, b1, W2, b2, W3, b3 = M.unpack(p, n_in, n_out)
    W1[0, 0], b1[0] = 2.0, 0.5
    W2[0, 0], b2[0] = -1.5, 4.0
    W3[0, 0], b3[0] = 3.0, -1.0
    x = np.array([[1.25]])
    t = np.array([[2.0]])
    w = np.array([[1.0]])
    h1 = 2.0 * 1.25 + 0.5          # 3.0
    h2 = -1.5 * h1 + 4.0           # -0.5 -> relu 0
    assert h2 < 0
    y = -1.0
    L, g = M.loss_grad(p, x, t, w)
    assert abs(L - (y - 2.0) ** 2) < 1e-12
    gW1, gb1, gW2, gb2, gW3, gb3 = M.unpack(g, n_in, n_out)
    assert abs(gb3[0] - 2 * (y - 2.0)) < 1e-12
    assert gW3[0, 0] == 0 and gW2[0, 0] == 0 and gW1[0, 0] == 0   # gated by relu(h2) = 0
    # open the gate: b2 = 10 -> h2 = 5.5, y = 15.5
    b2[0] = 10.0
    L, g = M.loss_grad(p, x, t, w)
    y = 3.0 * 5.5 - 1.0
    r = 2 * (y - 2.0)
    gW1, gb1, gW2, gb2, gW3, gb3 = M.unpack(g, n_in, n_out)
    assert abs(gW3[0, 0] - r * 5.5) < 1e-9
    assert abs(gb2[0] - r * 3.0) < 1e-9
    assert abs(gW2[0, 0] - r * 3.0 * h1) < 1e-9
    assert abs(gb1[0] - r * 3.0 * -1.5) < 1e-9
    assert abs(gW1[0, 0] - r * 3.0 * -1.5 * 1.25) < 1e-9


def test_zero_weight_targets_do_not_contribute():
    p, x, t, w = _case(4)
    w2 = w.copy()
    w2[:, 1] = 0
    t2 = t.copy()
    t2[:, 1] += 100.0
    L1, g1 = M.loss_grad(p, x, t2, w2)
    t3 = t.copy()
    L2, g2 = M.loss_grad(p, x, t3, w2)
    assert abs(L1 - L2) < 1e-12 and np.allclose(g1, g2)
