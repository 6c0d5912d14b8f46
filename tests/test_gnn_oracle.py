"""Pins the FP64 surrogate-network oracle (oracle/gnn_ref.py): its hand-written backward
pass against central finite differences, and its graph against brute force."""
import numpy as np
import pytest

from oracle import gnn_ref as G


def _problem(width=8, steps=2, n=14, n_recv=10, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1, (n, 2))
    send, recv, off = G.radius_graph(pos, n_recv, 0.45 ** 2)
    hist = rng.normal(0, 1, (5, n, 2))
    xn = G.node_features(hist, n_recv, 0.7)
    xe = G.edge_features(pos, send, recv, 1 / 0.45)
    params = G.init_params(width, steps, seed)
    for k in params:          # non-zero biases and residual branches for the check
        params[k] = params[k] + rng.normal(0, 0.05, params[k].shape)
    target = rng.normal(0, 1, (n, 2))
    mask = (np.arange(n) < n_recv).astype(float)
    return params, xn, xe, send, recv, off, target, mask, width, steps


def _loss(params, xn, xe, send, recv, off, target, mask, width, steps):
    out = G.forward(params, xn, xe, send, recv, off, width, steps)
    return G.masked_mse(out, target, mask)[0]


def test_graph_is_symmetric_among_receivers():
    params, xn, xe, send, recv, off, *_ = _problem()
    pairs = set(zip(recv.tolist(), send.tolist()))
    for i, j in pairs:
        if j < 10:
            assert (j, i) in pairs
    assert len(send) > 20


def test_backward_matches_finite_differences():
    params, xn, xe, send, recv, off, target, mask, width, steps = _problem()
    out, cache = G.forward(params, xn, xe, send, recv, off, width, steps, keep=True)
    loss, d_out = G.masked_mse(out, target, mask)
    grads = G.backward(params, cache, xn, xe, send, recv, d_out, width, steps)
    rng = np.random.default_rng(1)
    eps = 1e-6
    for name, w in params.items():
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in w.shape)
            old = w[idx]
            w[idx] = old + eps
            lp = _loss(params, xn, xe, send, recv, off, target, mask, width, steps)
            w[idx] = old - eps
            lm = _loss(params, xn, xe, send, recv, off, target, mask, width, steps)
            w[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert grads[name][idx] == pytest.approx(fd, rel=1e-5, abs=1e-8), name


def test_masked_loss_ignores_walls():
    params, xn, xe, send, recv, off, target, mask, width, steps = _problem()
    out = G.forward(params, xn, xe, send, recv, off, width, steps)
    t2 = target.copy()
    t2[mask == 0] += 100.0
    assert G.masked_mse(out, target, mask)[0] == G.masked_mse(out, t2, mask)[0]


def _sn_problem(width=8, steps=2, n=14, n_recv=10, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1, (n, 2))
    send, recv, off = G.radius_graph(pos, n_recv, 0.45 ** 2)
    hist = rng.normal(0, 1, (5, n, 2))
    xn = G.node_features(hist, n_recv, 0.7)
    re = G.sn_edge_features(pos, send, recv, 1 / 0.45)
    params = G.sn_init_params(width, steps, seed)
    for k in params:
        params[k] = params[k] + rng.normal(0, 0.05, params[k].shape)
    target = rng.normal(0, 1, (n, 2))
    mask = (np.arange(n) < n_recv).astype(float)
    return params, xn, re, send, recv, target, mask, width, steps


def test_rollout_net_backward_matches_finite_differences():
    """The rollout network's hand-written backward (sn_backward) against central finite
    differences of sn_forward, for every weight matrix."""
    params, xn, re, send, recv, target, mask, width, steps = _sn_problem()
    out, cache = G.sn_forward(params, xn, re, send, recv, width, steps, keep=True)
    loss, d_out = G.masked_mse(out, target, mask)
    grads = G.sn_backward(params, cache, xn, re, send, recv, d_out, width, steps)
    rng = np.random.default_rng(1)
    eps = 1e-6

    def L():
        return G.masked_mse(G.sn_forward(params, xn, re, send, recv, width, steps), target, mask)[0]
    for name, w in params.items():
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in w.shape)
            old = w[idx]
            w[idx] = old + eps
            lp = L()
            w[idx] = old - eps
            lm = L()
            w[idx] = old
            assert grads[name][idx] == pytest.approx((lp - lm) / (2 * eps), rel=1e-5, abs=1e-8), name


def test_rollout_net_message_sum_equals_summed_two_layer_messages():
    """Applying the edge MLP's second layer after the sum over a receiver's edges equals
    summing the two-layer messages [relu(z_e),1] D edge by edge (explicit loop)."""
    params, xn, re, send, recv, target, mask, width, steps = _sn_problem(seed=3)
    out, cache = G.sn_forward(params, xn, re, send, recv, width, steps, keep=True)
    n = xn.shape[0]
    h, z, sa, m, u2 = cache[0]
    D = params["p0_D"]
    expl = np.zeros((n, width))
    for e in range(len(send)):
        expl[recv[e]] += np.append(np.maximum(z[e], 0), 1.0) @ D
    assert np.allclose(m, expl, rtol=1e-12, atol=1e-12)
    # a receiver without edges gets the zero message
    lone = np.setdiff1d(np.arange(n), recv)
    assert len(lone) > 0 and np.all(m[lone] == 0)
