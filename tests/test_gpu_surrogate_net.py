"""The rollout network on the GPU (uph_amd/surrogate_net.py: BF16 MFMA node layers, the
edge sums of uph_sn_edge_agg / uph_sn_edge_bwd) against its FP64 statement
(oracle/gnn_ref.py sn_forward / sn_backward, pinned by finite differences in
test_gnn_oracle.py).

Tolerances: every node layer multiplies BF16 operands with FP32 accumulation, so
outputs and gradients carry the BF16 rounding of activations (2^-9 relative per layer,
compounding over the 2 + 4 S + 2 layers); the edge sums run in FP32.  The checks are
relative L2 errors of 3e-2 on outputs and 6e-2 on each weight gradient, which a wrong
index, a transposed operand or a dropped term exceeds by an order of magnitude."""
import numpy as np
import pytest

from oracle import gnn_ref
from uph_amd.devbuf import DevBuf, check, from_bf16, lib, to_bf16

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-sn] {msg}", flush=True)


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


def _check(pos, nf, R, width, steps, seed, hist_scale=0.3):
    from uph_amd.gnn import Graph
    from uph_amd.surrogate_net import SurrogateNet
    n = len(pos)
    dp = DevBuf.of(pos)
    g = Graph(n, nf, 1 << 16).build(dp, R * R)
    rng = np.random.default_rng(seed)
    hist = rng.normal(0, hist_scale, (5, n, 2))
    xn = DevBuf(n * 16 * 2)
    check(lib().uph_gns_node_feats(n, nf, DevBuf.of(hist).ptr, 1.0 / 0.5, xn.ptr, None))
    check(lib().uph_dev_sync())
    xn_h = from_bf16(xn.download((n, 16), np.uint16)).astype(np.float64)
    send = g.send.download(g.E, np.int32).astype(np.int64)
    recv = g.recv.download(g.E, np.int32).astype(np.int64)
    re = gnn_ref.sn_edge_features(pos, send, recv, 1.0 / R)
    params = gnn_ref.sn_init_params(width, steps, seed=11)
    for k, w in params.items():          # non-zero biases so every bias row is exercised
        if w.shape[0] == width + 1:
            w[-1] = rng.normal(0, 0.1, w.shape[1])
        if k.endswith("_A"):
            w[3] = rng.normal(0, 0.1, w.shape[1])
    # the oracle multiplies the BF16-rounded node weights the GPU uses; A stays FP32
    pb = {k: (v if k.endswith("_A") else from_bf16(to_bf16(v.astype(np.float32))).astype(np.float64))
          for k, v in params.items()}
    pb = {k: (v.astype(np.float32).astype(np.float64) if k.endswith("_A") else v) for k, v in pb.items()}
    net = SurrogateNet(width, steps, params=params)
    net.forward(g, dp, 1.0 / R, xn, train=True)
    out = net.output(n)
    ref, cache = gnn_ref.sn_forward(pb, xn_h, re, send, recv, width, steps, keep=True)
    e_out = _rel(out[:nf], ref[:nf])
    _say(f"W={width} S={steps} n={n} E={g.E}: forward rel L2 {e_out:.2e}")
    assert e_out < 3e-2
    target = rng.normal(0, 1.0, (n, 2)).astype(np.float32)
    mask = (np.arange(n) < nf).astype(np.int32)
    net.zero_grad()
    loss = net.loss(DevBuf.of(target), DevBuf.of(mask), nf)
    l_ref, d_ref = gnn_ref.masked_mse(ref, target.astype(np.float64), mask)
    _say(f"loss {loss:.6f} oracle {l_ref:.6f}")
    assert abs(loss - l_ref) < 2e-2 * l_ref
    net.backward()
    check(lib().uph_dev_sync())
    gr = gnn_ref.sn_backward(pb, cache, xn_h, re, send, recv, d_ref, width, steps)
    got = net.grads()
    worst, wk = 0.0, ""
    for k in gr:
        e = _rel(got[k], gr[k])
        if e > worst:
            worst, wk = e, k
        assert e < 6e-2, (k, e)
    _say(f"W={width} S={steps}: worst gradient rel L2 {worst:.2e} ({wk})")
    return net


@pytest.mark.parametrize("width,steps", [(32, 2), (128, 3)])
def test_forward_backward_vs_oracle(width, steps):
    pos, nf = _cloud(seed=3)
    net = _check(pos, nf, 2.6, width, steps, seed=6)
    # one Adam step moves each weight by at most lr
    w0 = net.params()
    net.adam(lr=1e-3)
    w1 = net.params()
    for k in w0:
        assert np.abs(w1[k] - w0[k]).max() <= 1e-3 * 1.001, k


def test_fused_inference_equals_the_layered_forward():
    """uph_sn_fused (the rollout's path: 1 + S launches at width 128) against the layered
    forward on the same graph and weights, and both against the FP64 oracle on the
    BF16-rounded weights, at W 128 and S 8.  The fused path also rounds the gathered zc
    to BF16 (the layered one keeps it in FP32), one more 2^-9 rounding per step, so both
    bounds are the layered test's 3e-2 (measured: 1.3e-2 between the two)."""
    from uph_amd.gnn import Graph
    from uph_amd.surrogate_net import SurrogateNet
    pos, nf = _cloud(nf=900, nw=200, seed=8)
    n, R = len(pos), 2.6
    dp = DevBuf.of(pos)
    g = Graph(n, nf, 1 << 16).build(dp, R * R, sender=False)
    rng = np.random.default_rng(2)
    xn = DevBuf(n * 16 * 2)
    check(lib().uph_gns_node_feats(n, nf, DevBuf.of(rng.normal(0, 0.3, (5, n, 2))).ptr, 2.0, xn.ptr, None))
    params = gnn_ref.sn_init_params(128, 8, seed=5)
    for k, w in params.items():
        if w.shape[0] == 129:
            w[-1] = rng.normal(0, 0.1, w.shape[1])
    for k, w in params.items():
        if k.endswith("_A"):
            w[3] = rng.normal(0, 0.1, w.shape[1])
    net = SurrogateNet(128, 8, params=params)
    net.forward(g, dp, 1.0 / R, xn, train=False)
    a = net.output(n)
    net.forward_fused(g, dp, 1.0 / R, xn)
    b = net.fused_output(n)
    xn_h = from_bf16(xn.download((n, 16), np.uint16)).astype(np.float64)
    send = g.send.download(g.E, np.int32).astype(np.int64)
    recv = g.recv.download(g.E, np.int32).astype(np.int64)
    pb = {k: (v.astype(np.float32).astype(np.float64) if k.endswith("_A")
              else from_bf16(to_bf16(v.astype(np.float32))).astype(np.float64)) for k, v in params.items()}
    ref = gnn_ref.sn_forward(pb, xn_h, gnn_ref.sn_edge_features(pos, send, recv, 1.0 / R), send, recv, 128, 8)
    e, ea, eb = _rel(b[:nf], a[:nf]), _rel(a[:nf], ref[:nf]), _rel(b[:nf], ref[:nf])
    _say(f"W128 S8, n={n} E={g.E}: fused vs layered rel L2 {e:.2e}; vs oracle: layered {ea:.2e}, fused {eb:.2e}")
    assert np.all(np.isfinite(b)) and e < 3e-2 and ea < 3e-2 and eb < 3e-2


def test_w128_s8_on_a_window_of_the_64k_dam_break():
    """The rollout workload's network size (W 128, S 8) on a subgraph of the 64,000-particle
    dam break after 100 plain steps: the fluid and wall particles of a 0.25 x 0.12 m window
    at the bottom-left corner, graph radius 2h."""
    from uph_amd import scenes
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(width=0.8, height=2.0, dx=0.005)
    g = GpuSolver(sc)
    g.step(100)
    x = g.state()["x"]
    nf = sc.n_fluid
    win = (x[:, 0] > -0.02) & (x[:, 0] < 0.23) & (x[:, 1] > -0.02) & (x[:, 1] < 0.10)
    fl = np.nonzero(win & (np.arange(sc.n) < nf))[0]
    wl = np.nonzero(win & (np.arange(sc.n) >= nf))[0]
    pos = np.concatenate([x[fl], x[wl]])
    _say(f"window: {len(fl)} fluid + {len(wl)} wall particles")
    assert len(fl) > 500 and len(wl) > 50
    _check(pos, len(fl), 2.0 * sc.h, 128, 8, seed=9, hist_scale=0.05)
