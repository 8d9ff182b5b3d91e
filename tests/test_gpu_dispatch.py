"""The learned dispatcher: its HIP network against the CPU oracle (forward
outputs and gradients, fp32 vs float64 within the stated tolerance), training
on the device, and the trained dispatcher judged by exhaustive timing (regret)
and against the default tiling (gain)."""
import os
import sys

import numpy as np
import pytest

from gpu_util import announce, rand_shards, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")
from rsgpu import dispatch as D  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp as M  # noqa: E402

# fp32 kernels vs the float64 oracle: sums of at most a few thousand fp32 products
RTOL = 2e-4


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _rand_problem(seed, n_in, n_out, B):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.25, M.n_params(n_in, n_out)).astype(np.float32)
    x = rng.normal(0, 1, (B, n_in)).astype(np.float32)
    t = rng.normal(0, 1, (B, n_out)).astype(np.float32)
    w = (rng.random((B, n_out)) < 0.8).astype(np.float32)
    w[0, 0] = 1.0  # at least one target
    return p, x, t, w


@pytest.mark.parametrize("n_in,n_out,B", [(14, 20, 700), (3, 1, 1), (16, 32, 257), (5, 7, 2600)])
def test_forward_and_gradient_match_oracle(n_in, n_out, B):
    announce(f"mlp grad check n_in={n_in} n_out={n_out} B={B}")
    p, x, t, w = _rand_problem(n_in * 100 + B, n_in, n_out, B)
    net = rsgpu.Mlp(n_in, n_out, seed=1)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    y = net.forward(_dev(x)).cpu().numpy()
    y_ref, _ = M.forward(p.astype(np.float64), x, n_out)
    assert np.allclose(y, y_ref, rtol=RTOL, atol=RTOL * np.abs(y_ref).max())
    loss, g = net.loss_grad(_dev(x), _dev(t), _dev(w))
    L_ref, g_ref = M.loss_grad(p.astype(np.float64), x, t, w)
    assert abs(loss - L_ref) <= RTOL * abs(L_ref)
    scale = np.abs(g_ref).max()
    assert np.allclose(g, g_ref, rtol=1e-3, atol=RTOL * scale), np.abs(g - g_ref).max() / scale
    # the step is deterministic: the same call gives the same bits
    loss2, g2 = net.loss_grad(_dev(x), _dev(t), _dev(w))
    assert loss2 == loss and np.array_equal(g, g2)


def test_adam_step_matches_oracle():
    """One device Adam step equals the textbook update applied to the oracle gradient."""
    announce("mlp adam step")
    n_in, n_out, B = 6, 4, 300
    p, x, t, w = _rand_problem(5, n_in, n_out, B)
    net = rsgpu.Mlp(n_in, n_out)
    net.set_params(p)
    lr = 1e-2
    net.train(_dev(x), _dev(t), _dev(w), 1, lr)
    _, g = M.loss_grad(p.astype(np.float64), x, t, w)
    m = 0.1 * g
    v = 0.001 * g * g
    expect = p - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    got = net.get_params()
    # The first Adam step moves each parameter by lr * g / (|g| + eps): its size
    # is fixed, its sign is the gradient's.  Where |g| is at fp32 noise level the
    # sign is not determined, so those entries are only checked for the step size.
    clear = np.abs(g) > 1e-3 * np.abs(g).max()
    assert clear.sum() > 0.5 * g.size
    assert np.allclose(got[clear], expect[clear], rtol=1e-4, atol=1e-5)
    assert np.all(np.abs(got - p) <= lr * (1 + 1e-4))


def test_training_on_device_fits_a_smooth_function():
    announce("mlp training")
    rng = np.random.default_rng(11)
    B, n_in, n_out = 2048, 4, 3
    x = rng.uniform(-1, 1, (B, n_in)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] ** 2, x[:, 2] * x[:, 3], np.abs(x[:, 0] - x[:, 3])],
                 1).astype(np.float32)
    w = np.ones_like(t)
    net = rsgpu.Mlp(n_in, n_out, seed=3)
    l0, _ = net.loss_grad(_dev(x), _dev(t), _dev(w))
    net.train(_dev(x), _dev(t), _dev(w), 1500, 3e-3)
    l1, _ = net.loss_grad(_dev(x), _dev(t), _dev(w))
    assert l1 < 0.05 * l0, (l0, l1)
    # the final loss agrees with the oracle at the trained parameters
    L_ref = M.loss(net.get_params().astype(np.float64), x, t, w)
    assert abs(l1 - L_ref) <= 1e-3 * L_ref + 1e-6


def test_dispatcher_trained_on_device_regret_and_gain():
    """Train on the committed timing set (80%) with the device kernels; on the
    held-out 20% the network's picks must be within 10% of the exhaustively
    measured best on average and keep most of the best's gain over the default
    tiling."""
    announce("dispatcher held-out regret")
    data = D.load(D.DEFAULT_TIMINGS)
    assert int(np.isfinite(data["T"]).sum()) > 40000
    tr, te = D.split(len(data["X"]))
    disp = D.Dispatcher(seed=0)
    loss = disp.fit(data["X"][tr], data["T"][tr], data["D"][tr], data["meta"][tr], steps=3000, lr=3e-3)
    assert loss < 0.05
    valid = np.isfinite(data["T"][te])
    picks = disp.pick(data["X"][te], valid)
    # only exact variants are picked, or the default tiling (-1)
    assert ((picks < 0) | valid[np.arange(len(picks)), np.maximum(picks, 0)]).all()
    r = D.regret(data["T"][te], data["D"][te], picks)
    print(f"held-out regret {r}", flush=True)
    assert r["mean"] <= 0.10, r
    assert r["total"] <= 0.05, r
    # over the default tiling, per call and in total time: at least 70% of the log
    # gain that the exhaustively measured best (default included) reaches on this set
    # (the default tiling is within about 11% of that best on average, so the
    # headroom is small; regret above is the primary measure)
    T, Dt = data["T"][te], data["D"][te]
    best = np.minimum(np.where(np.isfinite(T), T, np.inf).min(1), Dt)
    ceil_geo, ceil_tot = float(np.exp(np.log(Dt / best).mean())), float(Dt.sum() / best.sum())
    print(f"gain ceiling geomean {ceil_geo:.3f} total {ceil_tot:.3f}", flush=True)
    assert r["gain_vs_default_geomean"] >= ceil_geo ** 0.7, r
    assert r["gain_vs_default_total"] >= ceil_tot ** 0.7, r
    assert r["mean"] < r["default_mean"] / 4, r


def test_installed_tables_keep_outputs_exact():
    announce("dispatcher install exact")
    disp = D.Dispatcher.load()
    import oracle_lib as O
    for kind, k, m, n in [("gf8", 10, 4, 1 << 16), ("leopard_gf8", 12, 4, 8192)]:
        rs = D.make_codec(kind, k, m)
        disp.install(rs, kind)
        rng = np.random.default_rng(k + n)
        data = rand_shards(rng, k, n)
        if kind == "gf8":
            _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
            par = O.encode(O.GF8, mat, k, m, data)
        else:
            _, par = O.leo_encode(8, k, m, data)
        sh = [to_dev(d) for d in data] + [torch.zeros(n, dtype=torch.uint8, device="cuda")
                                         for _ in range(m)]
        rs.encode(sh)
        full = [to_host(x) for x in sh]
        for j in range(m):
            assert (full[k + j] == par[j]).all(), (kind, j)
        for erased in ([0, 3, k + 1, k + 2], [1]):
            work = [None if i in erased else to_dev(full[i]) for i in range(k + m)]
            rs.reconstruct(work)
            for i in erased:
                assert (to_host(work[i]) == full[i]).all(), (kind, erased, i)


def test_native_install_matches_python_table():
    """rsgpu_codec_install_dispatch (features, forward and argmin in the library)
    installs exactly the table the Python dispatcher computes."""
    announce("dispatcher native install")
    disp = D.Dispatcher.load()
    for kind, k, m in [("gf8", 10, 4), ("gf16", 6, 3), ("leopard_gf8", 32, 8), ("leopard_gf16", 64, 16),
                       ("gf8", 200, 30)]:
        rs = D.make_codec(kind, k, m)
        table = disp.table_for(rs, kind)
        disp.install_native(rs)
        n_ops, n_buckets, b0 = rsgpu.dispatch_dims()
        for op in range(n_ops):
            for e in range(m + 1):
                # an entry of -1 (or one not exact here) takes the library's default
                # choice: the matrix form (variant 12) for a Leopard reconstruct of
                # 1..4 erasures, the default tiling otherwise
                default = 12 if kind.startswith("leopard") and op == 2 and 1 <= e <= 4 else -1
                for b in range(n_buckets):
                    ln = int(round(2.0 ** ((b + b0) / 2.0)))
                    v = int(table[op, e, b])
                    want = v if v >= 0 and rs.variant_valid(op, e, v) else default
                    assert rs.variant_for(op, e, ln) == want, (kind, op, e, b)
