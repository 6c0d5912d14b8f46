"""The GPU surrogate network (radius graph, features, forward, loss, backward, Adam)
against the FP64 oracle statement oracle/gnn_ref.py, whose backward pass is itself
pinned by finite differences in test_gnn_oracle.py.

Tolerances: the GPU runs every dense layer with BF16 operands and FP32
accumulation, so outputs and gradients carry BF16 rounding of the activations
(relative 2^-9 per layer, compounding over 2 + 2 S + 2 layers).  The checks are
relative L2 errors of 3e-2 on outputs and 6e-2 on each weight gradient, which a
wrong index, transposed operand or dropped term exceeds by an order of magnitude."""
import numpy as np
import pytest

from oracle import gnn_ref
from uph_amd.devbuf import DevBuf, check, from_bf16, lib, to_bf16

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-gnn-net] {msg}", flush=True)


def _cloud(nf=400, nw=120, seed=0):
    """Jittered fluid lattice above a row of wall particles; spacing 1."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(nf)))
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), indexing="ij"), -1).reshape(-1, 2)
    fl = g[:nf].astype(np.float64) + rng.uniform(-0.2, 0.2, (nf, 2)) + [0.0, 1.5]
    wl = np.stack([np.linspace(-1, side, nw), np.zeros(nw)], 1)
    return np.concatenate([fl, wl]), nf


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_radius_graph_matches_brute_force():
    from uph_amd.gnn import Graph
    pos, nf = _cloud(seed=1)
    n = len(pos)
    dist = 2.6 ** 2
    dp = DevBuf.of(pos)
    g = Graph(n, nf, 64).build(dp, dist)          # also exercises the capacity retry
    off = g.off.download(n + 1, np.int32)
    send = g.send.download(g.E, np.int32)
    recv = g.recv.download(g.E, np.int32)
    soff = g.soff.download(n + 1, np.int32)
    sperm = g.sperm.download(g.E, np.int32)
    s_ref, r_ref, o_ref = gnn_ref.radius_graph(pos, nf, dist)
    _say(f"graph: {n} nodes, {g.E} edges")
    assert g.E == len(s_ref)
    assert np.array_equal(off, o_ref)
    assert np.array_equal(recv, r_ref)
    for i in range(nf):
        assert np.array_equal(np.sort(send[off[i]:off[i + 1]]), s_ref[o_ref[i]:o_ref[i + 1]])
    # sender view: the edges of node j as a sender, each exactly once
    assert soff[-1] == g.E and np.array_equal(np.sort(sperm), np.arange(g.E))
    for j in range(n):
        es = sperm[soff[j]:soff[j + 1]]
        assert np.all(send[es] == j)


def _features(pos, nf, g, seed):
    rng = np.random.default_rng(seed)
    n = len(pos)
    hist = rng.normal(0, 0.3, (5, n, 2))
    dh = DevBuf.of(hist)
    xn = DevBuf(n * 16 * 2)
    check(lib().uph_gns_node_feats(n, nf, dh.ptr, 1.0 / 0.5, xn.ptr, None))
    xe = DevBuf(g.E * 8 * 2)
    dp = DevBuf.of(pos)
    check(lib().uph_gns_edge_feats(g.E, g.send.ptr, g.recv.ptr, dp.ptr, 1.0 / 2.6, xe.ptr, None))
    check(lib().uph_dev_sync())
    xn_h = from_bf16(xn.download((n, 16), np.uint16)).astype(np.float64)
    xe_h = from_bf16(xe.download((g.E, 8), np.uint16)).astype(np.float64)
    return hist, xn, xe, xn_h, xe_h


def test_features_match_oracle():
    from uph_amd.gnn import Graph
    pos, nf = _cloud(seed=2)
    g = Graph(len(pos), nf, 1 << 16).build(DevBuf.of(pos), 2.6 ** 2)
    hist, xn, xe, xn_h, xe_h = _features(pos, nf, g, 5)
    send = g.send.download(g.E, np.int32)
    recv = g.recv.download(g.E, np.int32)
    rn = gnn_ref.node_features(hist, nf, 1.0 / 0.5)
    re = gnn_ref.edge_features(pos, send, recv, 1.0 / 2.6)
    assert np.all(np.abs(xn_h - rn) <= 2 ** -8 * np.abs(rn) + 1e-7)
    assert np.array_equal(xn_h[:, 10:], rn[:, 10:])
    assert np.abs(xe_h - re).max() < 2 ** -8 * max(1.0, np.abs(re).max())


@pytest.mark.parametrize("width,steps", [(32, 2), (128, 3)])
def test_forward_backward_vs_oracle(width, steps):
    from uph_amd.gnn import Graph, GraphNet
    pos, nf = _cloud(seed=3)
    n = len(pos)
    g = Graph(n, nf, 1 << 16).build(DevBuf.of(pos), 2.6 ** 2)
    hist, xn, xe, xn_h, xe_h = _features(pos, nf, g, 6)
    send = g.send.download(g.E, np.int32).astype(np.int64)
    recv = g.recv.download(g.E, np.int32).astype(np.int64)
    off = g.off.download(n + 1, np.int32)
    params = gnn_ref.init_params(width, steps, seed=11)
    rng = np.random.default_rng(4)
    # non-zero biases so the bias rows are exercised
    for k, w in params.items():
        if w.shape[0] == width + 1:
            w[-1] = rng.normal(0, 0.1, w.shape[1])
    # the oracle runs on the BF16-rounded weights the GPU multiplies with
    pb = {k: from_bf16(to_bf16(v.astype(np.float32))).astype(np.float64) for k, v in params.items()}
    net = GraphNet(width, steps, params=params)
    out_d = net.forward(g, xn, xe, train=True)
    out = net.output(n)
    ref, cache = gnn_ref.forward(pb, xn_h, xe_h, send, recv, off, width, steps, keep=True)
    e_out = _rel(out[:nf], ref[:nf])
    _say(f"W={width} S={steps}: forward rel L2 {e_out:.2e}")
    assert e_out < 3e-2
    # loss and backward
    target = rng.normal(0, 1.0, (n, 2)).astype(np.float32)
    mask = (np.arange(n) < nf).astype(np.int32)
    dt_, dm_ = DevBuf.of(target), DevBuf.of(mask)
    net.zero_grad()
    loss = net.loss(dt_, dm_, nf)
    l_ref, d_ref = gnn_ref.masked_mse(ref, target.astype(np.float64), mask)
    _say(f"loss {loss:.6f} oracle {l_ref:.6f}")
    assert abs(loss - l_ref) < 2e-2 * l_ref
    net.backward()
    check(lib().uph_dev_sync())
    gr = gnn_ref.backward(pb, cache, xn_h, xe_h, send, recv, d_ref, width, steps)
    got = net.grads()
    worst = 0.0
    for k in gr:
        e = _rel(got[k], gr[k])
        worst = max(worst, e)
        assert e < 6e-2, (k, e)
    _say(f"W={width} S={steps}: worst gradient rel L2 {worst:.2e}")
    # one Adam step moves every weight by at most lr (bias-corrected first step: exactly lr
    # where the gradient is non-zero)
    w0 = net.params()
    net.adam(lr=1e-3)
    w1 = net.params()
    for k in w0:
        d = np.abs(w1[k] - w0[k])
        assert d.max() <= 1e-3 * 1.001, k
        nz = np.abs(got[k]) > 1e-6
        assert np.allclose(d[nz], 1e-3, rtol=2e-2), k
    del out_d


def test_training_reduces_loss():
    """Sixty Adam steps on one graph fit a smooth target: the loss falls by 10x."""
    from uph_amd.gnn import Graph, GraphNet
    pos, nf = _cloud(seed=4)
    n = len(pos)
    g = Graph(n, nf, 1 << 16).build(DevBuf.of(pos), 2.6 ** 2)
    hist, xn, xe, _, _ = _features(pos, nf, g, 7)
    # target: a smooth function of the node's own newest velocity and its height
    target = np.stack([np.sin(hist[0, :, 0] / 0.5), 0.5 * (pos[:, 1] / pos[:, 1].max())], 1)
    target = target.astype(np.float32)
    mask = (np.arange(n) < nf).astype(np.int32)
    dt_, dm_ = DevBuf.of(target), DevBuf.of(mask)
    net = GraphNet(64, 2, seed=1)
    losses = []
    for it in range(60):
        net.forward(g, xn, xe, train=True)
        net.zero_grad()
        losses.append(net.loss(dt_, dm_, nf))
        net.backward()
        net.adam(lr=3e-3)
    _say(f"training loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.1 * losses[0]


def test_fused_edge_step_matches_unfused_path():
    """The fused edge kernel (both edge GEMMs, the gathers and the residual update in one
    launch) against the separate GEMM / gather / ReLU / add kernels: same network, same
    graph; they differ only in the FP32 summation order before each BF16 rounding."""
    from uph_amd.gnn import Graph, GraphNet
    pos, nf = _cloud(nf=900, nw=150, seed=8)
    n = len(pos)
    g = Graph(n, nf, 1 << 17).build(DevBuf.of(pos), 2.6 ** 2)
    hist, xn, xe, xn_h, xe_h = _features(pos, nf, g, 9)
    params = gnn_ref.init_params(128, 3, seed=5)
    rng = np.random.default_rng(6)
    for k, w in params.items():
        if w.shape[0] == 129:
            w[-1] = rng.normal(0, 0.1, w.shape[1])
    outs = {}
    for fused in (False, True):
        net = GraphNet(128, 3, params=params)
        net.forward(g, xn, xe, train=False, fused=fused)
        outs[fused] = net.output(n)[:nf]
    e = _rel(outs[True], outs[False])
    _say(f"fused vs unfused forward: rel L2 {e:.2e} (E = {g.E})")
    assert e < 1e-2
    # a partial last tile (192 edges per tile) is covered
    assert g.E % 192 != 0
