"""FP64 numpy statement of the surrogate graph network, forward and backward.

TEST INFRASTRUCTURE ONLY: the checker for the GPU network (uph_amd/gnn.py),
whose dense layers run as BF16 MFMA kernels.  Its own backward pass is pinned
by central finite differences in tests/test_gnn_oracle.py.

Architecture (encode-process-decode, Sanchez-Gonzalez et al. 2020):
  node encoder  h = L2(relu(L1(x_node)))            x_node: n x 16 (col 11 = 1)
  edge encoder  e = L2(relu(L1(x_edge)))            x_edge: E x 8  (col 3 = 1)
  step l = 1..S:
    z  = [e,1] A + ([h,1] B)[recv] + ([h,1] C)[send]
    m  = [relu(z),1] D                  edge message
    a  = segment sum of m over each receiver's edges
    y  = [h,1] Ew + a Fw
    dh = [relu(y),1] G
    e += m ;  h += dh
  decoder       out = [relu([h,1] P),1] Q       (n x 2, acceleration / ascale)
Every weight matrix holds its bias as the row that multiplies the constant-1
column; [v,1] is v with that column appended.
"""
from __future__ import annotations

import numpy as np

NODE_IN, EDGE_IN, OUT = 16, 8, 2


def layer_shapes(width: int, steps: int):
    """Name -> (rows, cols) of every weight matrix (rows include the bias row)."""
    W = width
    s = {"enc_n1": (NODE_IN, W), "enc_n2": (W + 1, W), "enc_e1": (EDGE_IN, W), "enc_e2": (W + 1, W),
         "dec1": (W + 1, W), "dec2": (W + 1, OUT)}
    for l in range(steps):
        s.update({f"p{l}_A": (W + 1, W), f"p{l}_B": (W + 1, W), f"p{l}_C": (W + 1, W),
                  f"p{l}_D": (W + 1, W), f"p{l}_E": (W + 1, W), f"p{l}_F": (W, W),
                  f"p{l}_G": (W + 1, W)})
    return s


def init_params(width: int, steps: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    out = {}
    for name, (r, c) in layer_shapes(width, steps).items():
        fan_in = r - 1 if r in (width + 1,) else r
        w = rng.normal(0.0, np.sqrt(2.0 / max(fan_in, 1)), (r, c))
        if r == width + 1:
            w[-1] = 0.0                      # bias row starts at zero
        if name in ("enc_n1",):
            w[11] = 0.0                      # the constant column of the node input
        if name in ("enc_e1",):
            w[3] = 0.0
        # residual branch outputs start small
        if name.endswith("_D") or name.endswith("_G"):
            w *= 0.1
        out[name] = w
    return out


def _one(x):
    return np.concatenate([x, np.ones((x.shape[0], 1))], axis=1)


def forward(params, xn, xe, send, recv, off, width, steps, keep=False):
    """Returns out (n x 2) and, with keep, the activations for backward."""
    n = xn.shape[0]
    c = {}
    z = xn @ params["enc_n1"]
    u = np.maximum(z, 0)
    h = _one(u) @ params["enc_n2"]
    c["enc_n"] = (u,)
    z = xe @ params["enc_e1"]
    u = np.maximum(z, 0)
    e = _one(u) @ params["enc_e2"]
    c["enc_e"] = (u,)
    for l in range(steps):
        P = lambda k: params[f"p{l}_{k}"]
        h1 = _one(h)
        zb = h1 @ P("B")
        zc = h1 @ P("C")
        z1 = _one(e) @ P("A") + zb[recv] + zc[send]
        u1 = np.maximum(z1, 0)
        m = _one(u1) @ P("D")
        a = np.zeros((n, width))
        np.add.at(a, recv, m)
        y = h1 @ P("E") + a @ P("F")
        u2 = np.maximum(y, 0)
        dh = _one(u2) @ P("G")
        c[l] = (h, e, u1, a, u2)
        e = e + m
        h = h + dh
    zp = _one(h) @ params["dec1"]
    up = np.maximum(zp, 0)
    out = _one(up) @ params["dec2"]
    c["dec"] = (h, up)
    return (out, c) if keep else out


def backward(params, cache, xn, xe, send, recv, d_out, width, steps):
    """Gradients of sum(d_out * out) with respect to every weight matrix."""
    n = xn.shape[0]
    g = {k: np.zeros_like(v) for k, v in params.items()}
    h, up = cache["dec"]
    g["dec2"] += _one(up).T @ d_out
    dup = (d_out @ params["dec2"].T)[:, :width] * (up > 0)
    g["dec1"] += _one(h).T @ dup
    dh = (dup @ params["dec1"].T)[:, :width]
    de = np.zeros((len(send), width))
    for l in reversed(range(steps)):
        P = lambda k: params[f"p{l}_{k}"]
        hp, ep, u1, a, u2 = cache[l]
        h1 = _one(hp)
        # node update: h_new = h + [relu(y),1] G,  y = [h,1] E + a F
        g[f"p{l}_G"] += _one(u2).T @ dh
        dy = (dh @ P("G").T)[:, :width] * (u2 > 0)
        g[f"p{l}_E"] += h1.T @ dy
        g[f"p{l}_F"] += a.T @ dy
        dhp = dh + (dy @ P("E").T)[:, :width]
        da = dy @ P("F").T
        # edge update: e_new = e + m,  m = [relu(z1),1] D
        dm = de + da[recv]
        g[f"p{l}_D"] += _one(u1).T @ dm
        dz1 = (dm @ P("D").T)[:, :width] * (u1 > 0)
        g[f"p{l}_A"] += _one(ep).T @ dz1
        dep = de + (dz1 @ P("A").T)[:, :width]
        dzb = np.zeros((n, width))
        np.add.at(dzb, recv, dz1)
        dzc = np.zeros((n, width))
        np.add.at(dzc, send, dz1)
        g[f"p{l}_B"] += h1.T @ dzb
        g[f"p{l}_C"] += h1.T @ dzc
        dhp = dhp + (dzb @ P("B").T)[:, :width] + (dzc @ P("C").T)[:, :width]
        dh, de = dhp, dep
    (u,) = cache["enc_e"]
    g["enc_e2"] += _one(u).T @ de
    du = (de @ params["enc_e2"].T)[:, :width] * (u > 0)
    g["enc_e1"] += xe.T @ du
    (u,) = cache["enc_n"]
    g["enc_n2"] += _one(u).T @ dh
    du = (dh @ params["enc_n2"].T)[:, :width] * (u > 0)
    g["enc_n1"] += xn.T @ du
    return g


def masked_mse(out, target, mask):
    """Mean over fluid rows and both components; returns (loss, d_out)."""
    cnt = int(mask.sum()) * out.shape[1]
    d = (out - target) * mask[:, None]
    return float((d * d).sum() / cnt), 2.0 * d / cnt


def node_features(hist, n_recv, inv_vscale):
    """hist: 5 x n x 2 velocities, newest first (same layout as the GPU kernel)."""
    n = hist.shape[1]
    x = np.zeros((n, NODE_IN))
    for k in range(5):
        x[:, 2 * k:2 * k + 2] = hist[k] * inv_vscale
    x[n_recv:, 10] = 1.0
    x[:, 11] = 1.0
    return x


def edge_features(pos, send, recv, invR):
    d = (pos[recv] - pos[send]) * invR
    x = np.zeros((len(send), EDGE_IN))
    x[:, 0:2] = d
    x[:, 2] = np.sqrt((d * d).sum(1))
    x[:, 3] = 1.0
    return x


def radius_graph(pos, n_recv, dist):
    """Edges (recv i <- send j) for every j != i within sqrt(dist), receivers i < n_recv,
    grouped by receiver, senders in ascending index (brute force, small inputs)."""
    send, recv = [], []
    off = [0]
    for i in range(pos.shape[0]):
        if i < n_recv:
            d = pos - pos[i]
            d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
            js = np.nonzero(d2 < dist)[0]
            js = js[js != i]
            send.extend(js.tolist())
            recv.extend([i] * len(js))
        off.append(len(send))
    return np.array(send, np.int64), np.array(recv, np.int64), np.array(off, np.int64)


# ---------------------------------------------------------------------------
# The rollout network (uph_amd/surrogate_net.py): the same encode-process-decode
# shape without an edge latent.  Each step's edge MLP reads the edge features and
# the two node latents, and its second layer is applied after the sum over a
# receiver's edges (it is linear, so sum_j ([relu(z_j),1] D) = [sum_j relu(z_j), deg] D):
#   node encoder  h = [relu(x_node L1),1] L2
#   step l:
#     z_e = r_e A + ([h,1] B)[recv] + ([h,1] C)[send]    r_e = ((x_i - x_j)/R, |r|/R, 1)
#     s   = segment sum of relu(z_e) over each receiver's edges; deg = edge count
#     m   = [s, deg] D
#     y   = [h,1] E + m F
#     h  += [relu(y),1] G
#   decoder       out = [relu([h,1] P),1] Q
# ---------------------------------------------------------------------------

SN_EDGE_IN = 4


def sn_layer_shapes(width: int, steps: int):
    W = width
    s = {"enc1": (NODE_IN, W), "enc2": (W + 1, W), "dec1": (W + 1, W), "dec2": (W + 1, OUT)}
    for l in range(steps):
        s.update({f"p{l}_A": (SN_EDGE_IN, W), f"p{l}_B": (W + 1, W), f"p{l}_C": (W + 1, W),
                  f"p{l}_D": (W + 1, W), f"p{l}_E": (W + 1, W), f"p{l}_F": (W, W),
                  f"p{l}_G": (W + 1, W)})
    return s


def sn_init_params(width: int, steps: int, seed: int = 0):
    """He-normal weights; bias rows (the last row of a (W+1)-row matrix, the fourth row
    of A, the constant column 11 of the node input) start at zero; the residual
    branch's output layer G starts small."""
    rng = np.random.default_rng(seed)
    out = {}
    for name, (r, c) in sn_layer_shapes(width, steps).items():
        fan_in = r - 1 if r == width + 1 else (3 if r == SN_EDGE_IN else r)
        w = rng.normal(0.0, np.sqrt(2.0 / max(fan_in, 1)), (r, c))
        if r == width + 1:
            w[-1] = 0.0
        if name == "enc1":
            w[11] = 0.0
        if name.endswith("_A"):
            w[3] = 0.0
        if name.endswith("_D"):
            w *= 1.0 / np.sqrt(20.0)      # about 20 edges per receiver
        if name.endswith("_G"):
            w *= 0.1
        out[name] = w
    return out


def sn_edge_features(pos, send, recv, invR):
    d = (pos[recv] - pos[send]) * invR
    x = np.ones((len(send), SN_EDGE_IN))
    x[:, 0:2] = d
    x[:, 2] = np.sqrt((d * d).sum(1))
    return x


def sn_forward(params, xn, re, send, recv, width, steps, keep=False):
    n = xn.shape[0]
    c = {}
    u = np.maximum(xn @ params["enc1"], 0)
    h = _one(u) @ params["enc2"]
    c["enc"] = (u,)
    deg = np.bincount(recv, minlength=n).astype(np.float64)
    for l in range(steps):
        P = lambda k: params[f"p{l}_{k}"]
        h1 = _one(h)
        zb = h1 @ P("B")
        zc = h1 @ P("C")
        z = re @ P("A") + zb[recv] + zc[send]
        s = np.zeros((n, width))
        np.add.at(s, recv, np.maximum(z, 0))
        sa = np.concatenate([s, deg[:, None]], axis=1)
        m = sa @ P("D")
        y = h1 @ P("E") + m @ P("F")
        u2 = np.maximum(y, 0)
        c[l] = (h, z, sa, m, u2)
        h = h + _one(u2) @ P("G")
    up = np.maximum(_one(h) @ params["dec1"], 0)
    out = _one(up) @ params["dec2"]
    c["dec"] = (h, up)
    return (out, c) if keep else out


def sn_backward(params, cache, xn, re, send, recv, d_out, width, steps):
    n = xn.shape[0]
    g = {k: np.zeros_like(v) for k, v in params.items()}
    h, up = cache["dec"]
    g["dec2"] += _one(up).T @ d_out
    dup = (d_out @ params["dec2"].T)[:, :width] * (up > 0)
    g["dec1"] += _one(h).T @ dup
    dh = (dup @ params["dec1"].T)[:, :width]
    for l in reversed(range(steps)):
        P = lambda k: params[f"p{l}_{k}"]
        hp, z, sa, m, u2 = cache[l]
        h1 = _one(hp)
        g[f"p{l}_G"] += _one(u2).T @ dh
        dy = (dh @ P("G").T)[:, :width] * (u2 > 0)
        g[f"p{l}_E"] += h1.T @ dy
        g[f"p{l}_F"] += m.T @ dy
        dhp = dh + (dy @ P("E").T)[:, :width]
        dm = dy @ P("F").T
        g[f"p{l}_D"] += sa.T @ dm
        ds = (dm @ P("D").T)[:, :width]
        ge = ds[recv] * (z > 0)
        g[f"p{l}_A"] += re.T @ ge
        dzb = np.zeros((n, width))
        np.add.at(dzb, recv, ge)
        dzc = np.zeros((n, width))
        np.add.at(dzc, send, ge)
        g[f"p{l}_B"] += h1.T @ dzb
        g[f"p{l}_C"] += h1.T @ dzc
        dh = dhp + (dzb @ P("B").T)[:, :width] + (dzc @ P("C").T)[:, :width]
    (u,) = cache["enc"]
    g["enc2"] += _one(u).T @ dh
    du = (dh @ params["enc2"].T)[:, :width] * (u > 0)
    g["enc1"] += xn.T @ du
    return g
