"""The rollout network of the surrogate: an encode-process-decode graph network with
no per-edge latent, on the GPU (forward, backward, Adam).

Equations (FP64 statement: oracle/gnn_ref.py, sn_forward / sn_backward):
  node encoder  h = [relu(x_node L1),1] L2                     x_node: n x 16
  step l:
    z_e = r_e A + ([h,1] B)[recv] + ([h,1] C)[send]          r_e = ((x_i - x_j)/R, |r|/R, 1)
    s   = sum of relu(z_e) over each receiver's edges;  deg = its edge count
    m   = [s, deg] D        (the edge MLP's second layer, applied after the sum)
    y   = [h,1] E + m F
    h  += [relu(y),1] G
  decoder       out = [relu([h,1] P),1] Q                       (acceleration / ascale)

The second layer of the edge MLP is linear, so summing the two-layer messages equals
applying it once to the summed hidden layer with the edge count multiplying its
bias: no edge-sized tensor is ever stored.  The node layers run as the library's
BF16 MFMA GEMMs (FP32 accumulation, biases as the weight row that multiplies a
constant-1 column); the edge sum and its backward are uph_sn_edge_agg /
uph_sn_edge_bwd, with the edge features formed from the FP64 positions.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from .devbuf import DevBuf, check, lib
from .gnn import OUT, Graph, GraphNet, Lin

NODE_IN, EDGE_IN = 16, 4


def _check_sn(rc):
    if rc != 0:
        raise RuntimeError(lib().uph_sn_last_error().decode() or lib().uph_gnn_last_error().decode())


def layer_shapes(width: int, steps: int):
    W = width
    s = {"enc1": (NODE_IN, W), "enc2": (W + 1, W), "dec1": (W + 1, W), "dec2": (W + 1, OUT)}
    for l in range(steps):
        s.update({f"p{l}_A": (EDGE_IN, W), f"p{l}_B": (W + 1, W), f"p{l}_C": (W + 1, W),
                  f"p{l}_D": (W + 1, W), f"p{l}_E": (W + 1, W), f"p{l}_F": (W, W),
                  f"p{l}_G": (W + 1, W)})
    return s


def init_params(width: int, steps: int, seed: int = 0):
    """He-normal weights; bias rows start at zero; D scaled for about 20 edges per
    receiver and the residual branch's G by 0.1 (the oracle's sn_init_params)."""
    rng = np.random.default_rng(seed)
    out = {}
    for name, (r, c) in layer_shapes(width, steps).items():
        fan_in = r - 1 if r == width + 1 else (3 if r == EDGE_IN else r)
        w = rng.normal(0.0, np.sqrt(2.0 / max(fan_in, 1)), (r, c))
        if r == width + 1:
            w[-1] = 0.0
        if name == "enc1":
            w[11] = 0.0
        if name.endswith("_A"):
            w[3] = 0.0
        if name.endswith("_D"):
            w *= 1.0 / np.sqrt(20.0)
        if name.endswith("_G"):
            w *= 0.1
        out[name] = w
    return out


class SurrogateNet(GraphNet):
    """The rollout network with device buffers sized on demand (GraphNet supplies the
    parameter handling, the GEMM helpers, the loss and Adam)."""

    def __init__(self, width: int = 128, steps: int = 8, params: Optional[Dict] = None, seed: int = 0):
        if width % 8 or width > 128:
            raise ValueError("width must be a multiple of 8, at most 128")
        self.W, self.S = width, steps
        params = params if params is not None else init_params(width, steps, seed)
        shapes = layer_shapes(width, steps)
        self.lin = {k: Lin(*shapes[k]) for k in shapes}
        for k, L in self.lin.items():
            L.set(np.asarray(params[k], np.float32))
        self.adam_t = 0
        self._cap = (0, 0, False)
        self.work = None

    @classmethod
    def from_state(cls, d) -> "SurrogateNet":
        W, S = int(d["_width"]), int(d["_steps"])
        return cls(W, S, params={k: d[k] for k in d if not k.startswith("_")})

    def state_dict(self) -> Dict[str, np.ndarray]:
        d = super().state_dict()
        d["_arch"] = np.array("rollout")
        return d

    # ------------------------------------------------------------ buffers

    def _ensure(self, n: int, E: int, train: bool):
        n = max(n, 1)
        cn, _, ct = self._cap
        if n <= cn and (ct or not train):
            return
        n = max(n if n <= cn else int(n * 1.1) + 64, cn)
        train = bool(train or ct)
        W, Wl = self.W, self.W + 8
        nb = lambda rows, cols, es: DevBuf(rows * cols * es)
        ns = self.S if train else 1
        b = self.b = {}
        b["h"] = nb(n, W, 4)                                   # FP32 latent
        b["zn"] = nb(n, Wl, 4)                                 # FP32 GEMM output
        b["out"] = nb(n, 8, 4)
        b["zb"] = [nb(n, W, 4) for _ in range(ns)]
        b["zc"] = [nb(n, W, 4) for _ in range(ns)]
        b["s"] = nb(n, Wl, 4)
        b["m"] = nb(n, W, 4)
        b["dh"] = nb(n, W, 4)
        # BF16 activations; hb, ub, un, up carry the constant 1 at column W
        for key, k, cols in (("hb", self.S + 1 if train else 1, Wl), ("ub", ns, Wl), ("sb", ns, Wl),
                             ("mb", ns, W), ("un", 1, Wl), ("up", 1, Wl)):
            b[key] = [nb(n, cols, 2) for _ in range(k)]
            for buf in b[key]:
                buf.zero()
                if key in ("hb", "ub", "un", "up"):
                    check(lib().uph_fill_col_bf16(n, Wl, W, 1.0, buf.ptr, None))
        if train:
            for key, cols, es in (("gh", Wl, 4), ("t1", Wl, 4), ("ds", Wl, 4), ("dzb", W, 4),
                                  ("dzc", W, 4), ("g16a", Wl, 2), ("g16b", Wl, 2), ("g16c", W, 2),
                                  ("dout", 8, 4), ("doutb", 8, 2)):
                b[key] = nb(n, cols, es)
            for key in ("g16a", "g16b", "g16c"):
                b[key].zero()
            b["lpart"] = nb((n * 8 + 255) // 256 + 1, 1, 4)
            b["part"] = nb(1024 * W, 1, 4)
        self.work = DevBuf(4 * max(n * Wl, 64 * Wl * Wl))
        self._cap = (n, 0, train)

    # ------------------------------------------------------------ forward

    def forward(self, g: Graph, pos: DevBuf, invR: float, xn: DevBuf, train: bool = False) -> DevBuf:
        """The network on graph g (receivers: rows < g.n_recv) with node features xn
        (BF16 n x 16) and positions pos (FP64 n x 2, device)."""
        n, W, Wl = g.n, self.W, self.W + 8
        self._ensure(n, g.E, train)
        b, L, Lb = self.b, self.lin, lib()
        self._lin_fwd(L["enc1"], xn.ptr, NODE_IN, n, b["zn"].ptr, Wl)
        check(Lb.uph_relu_to_bf16(n, W, b["zn"].ptr, Wl, b["un"][0].ptr, Wl, None))
        self._lin_fwd(L["enc2"], b["un"][0].ptr, Wl, n, b["h"].ptr, W)
        check(Lb.uph_rows_to_bf16(n, W, b["h"].ptr, W, b["hb"][0].ptr, Wl, None))
        for l in range(self.S):
            s = l if train else 0
            P = lambda k: L[f"p{l}_{k}"]
            hb = b["hb"][s].ptr
            zb, zc = b["zb"][s].ptr, b["zc"][s].ptr
            self._lin_fwd(P("B"), hb, Wl, n, zb, W)
            self._lin_fwd(P("C"), hb, Wl, n, zc, W)
            A = P("A")
            _check_sn(Lb.uph_sn_edge_agg(n, g.n_recv, W, g.off.ptr, g.send.ptr, pos.ptr, invR, A.w32.ptr, A.Np,
                                         zb, W, zc, W, b["s"].ptr, Wl, None))
            sb, mb, ub = b["sb"][s].ptr, b["mb"][s].ptr, b["ub"][s].ptr
            check(Lb.uph_rows_to_bf16(n, Wl, b["s"].ptr, Wl, sb, Wl, None))
            self._lin_fwd(P("D"), sb, Wl, n, b["m"].ptr, W)
            check(Lb.uph_rows_to_bf16(n, W, b["m"].ptr, W, mb, W, None))
            self._lin_fwd(P("E"), hb, Wl, n, b["zn"].ptr, Wl)
            self._gemm(n, W, W, mb, W, P("F").wt.ptr, P("F").Kp, b["zn"].ptr, Wl, trans_b=1, accumulate=1)
            check(Lb.uph_relu_to_bf16(n, W, b["zn"].ptr, Wl, ub, Wl, None))
            self._lin_fwd(P("G"), ub, Wl, n, b["dh"].ptr, W)
            check(Lb.uph_add_f32(n * W, b["dh"].ptr, b["h"].ptr, None))
            s1 = l + 1 if train else 0
            check(Lb.uph_rows_to_bf16(n, W, b["h"].ptr, W, b["hb"][s1].ptr, Wl, None))
        hbf = b["hb"][self.S if train else 0].ptr
        self._lin_fwd(L["dec1"], hbf, Wl, n, b["zn"].ptr, Wl)
        check(Lb.uph_relu_to_bf16(n, W, b["zn"].ptr, Wl, b["up"][0].ptr, Wl, None))
        self._lin_fwd(L["dec2"], b["up"][0].ptr, Wl, n, b["out"].ptr, 8)
        self._g, self._xn, self._pos, self._invR = g, xn, pos, float(invR)
        return b["out"]

    # ------------------------------------------------------------ backward

    def backward(self, d_out: Optional[DevBuf] = None):
        """Accumulates the weight gradients of sum(d_out * out) (default: the loss's d_out)."""
        g, b, L, Lb = self._g, self.b, self.lin, lib()
        n, W, Wl = g.n, self.W, self.W + 8
        dout = d_out if d_out is not None else b["dout"]
        check(Lb.uph_rows_to_bf16(n, 8, dout.ptr, 8, b["doutb"].ptr, 8, None))
        self._lin_dw(L["dec2"], b["up"][0].ptr, Wl, n, b["doutb"].ptr, 8)
        self._lin_dx(L["dec2"], b["doutb"].ptr, 8, n, b["t1"].ptr, Wl)
        check(Lb.uph_relu_bwd_rows(n, W, b["t1"].ptr, Wl, b["up"][0].ptr, Wl, b["g16a"].ptr, Wl, None))
        self._lin_dw(L["dec1"], b["hb"][self.S].ptr, Wl, n, b["g16a"].ptr, Wl)
        self._lin_dx(L["dec1"], b["g16a"].ptr, Wl, n, b["gh"].ptr, Wl)      # dh in columns < W
        for l in reversed(range(self.S)):
            P = lambda k: L[f"p{l}_{k}"]
            hb, ub, sb, mb = b["hb"][l].ptr, b["ub"][l].ptr, b["sb"][l].ptr, b["mb"][l].ptr
            zb, zc = b["zb"][l].ptr, b["zc"][l].ptr
            # h_new = h + [u,1] G,  u = relu(y),  y = [h,1] E + m F
            check(Lb.uph_rows_to_bf16(n, W, b["gh"].ptr, Wl, b["g16a"].ptr, Wl, None))
            self._lin_dw(P("G"), ub, Wl, n, b["g16a"].ptr, Wl)
            self._lin_dx(P("G"), b["g16a"].ptr, Wl, n, b["t1"].ptr, Wl)
            check(Lb.uph_relu_bwd_rows(n, W, b["t1"].ptr, Wl, ub, Wl, b["g16b"].ptr, Wl, None))
            dy = b["g16b"].ptr
            self._lin_dw(P("E"), hb, Wl, n, dy, Wl)
            self._lin_dw(P("F"), mb, W, n, dy, Wl)
            self._lin_dx(P("E"), dy, Wl, n, b["t1"].ptr, Wl)
            check(Lb.uph_add_rows_f32(n, W, b["t1"].ptr, Wl, b["gh"].ptr, Wl, None))
            # dm = dy F^T;  m = [s, deg] D
            self._gemm(n, W, W, dy, Wl, P("F").wb.ptr, P("F").Np, b["m"].ptr, W, trans_b=1)
            check(Lb.uph_rows_to_bf16(n, W, b["m"].ptr, W, b["g16c"].ptr, W, None))
            self._lin_dw(P("D"), sb, Wl, n, b["g16c"].ptr, W)
            self._lin_dx(P("D"), b["g16c"].ptr, W, n, b["ds"].ptr, Wl)
            # the edge sum: dzb, dzc and dA
            A = P("A")
            _check_sn(Lb.uph_sn_edge_bwd(n, g.n_recv, W, g.off.ptr, g.send.ptr, g.soff.ptr, g.sperm.ptr,
                                         g.recv.ptr, self._pos.ptr, self._invR, A.w32.ptr, A.Np, zb, W, zc, W,
                                         b["ds"].ptr, Wl, b["dzb"].ptr, b["dzc"].ptr, W, A.g.ptr,
                                         b["part"].ptr, None))
            # zb = [h,1] B, zc = [h,1] C
            check(Lb.uph_rows_to_bf16(n, W, b["dzb"].ptr, W, b["g16a"].ptr, Wl, None))
            check(Lb.uph_rows_to_bf16(n, W, b["dzc"].ptr, W, b["g16b"].ptr, Wl, None))
            self._lin_dw(P("B"), hb, Wl, n, b["g16a"].ptr, Wl)
            self._lin_dw(P("C"), hb, Wl, n, b["g16b"].ptr, Wl)
            self._lin_dx(P("B"), b["g16a"].ptr, Wl, n, b["t1"].ptr, Wl)
            check(Lb.uph_add_rows_f32(n, W, b["t1"].ptr, Wl, b["gh"].ptr, Wl, None))
            self._lin_dx(P("C"), b["g16b"].ptr, Wl, n, b["t1"].ptr, Wl)
            check(Lb.uph_add_rows_f32(n, W, b["t1"].ptr, Wl, b["gh"].ptr, Wl, None))
        # encoder
        check(Lb.uph_rows_to_bf16(n, W, b["gh"].ptr, Wl, b["g16a"].ptr, Wl, None))
        self._lin_dw(L["enc2"], b["un"][0].ptr, Wl, n, b["g16a"].ptr, Wl)
        self._lin_dx(L["enc2"], b["g16a"].ptr, Wl, n, b["t1"].ptr, Wl)
        check(Lb.uph_relu_bwd_rows(n, W, b["t1"].ptr, Wl, b["un"][0].ptr, Wl, b["g16b"].ptr, Wl, None))
        self._lin_dw(L["enc1"], self._xn.ptr, NODE_IN, n, b["g16b"].ptr, Wl)

    # ------------------------------------------------------------ fused inference

    def forward_fused(self, g: Graph, pos: DevBuf, invR: float, xn: DevBuf) -> DevBuf:
        """The same network as forward(), inference only, in 1 + S launches of
        uph_sn_fused (width 128): the encoder, then one launch per message-passing step
        (edge sums, node MLP, residual, next projections; the last one the decoder).
        Returns the output buffer (n x 8 FP32, columns 0..1)."""
        if self.W != 128:
            raise ValueError("the fused path is built for width 128")
        n = g.n
        fb = getattr(self, "fb", None)
        if fb is None or fb["n"] < n:
            m = int(n * 1.1) + 64
            fb = self.fb = {"n": m, "h": DevBuf(4 * 128 * m), "out": DevBuf(4 * 8 * m),
                            "zb": [DevBuf(4 * 128 * m) for _ in range(2)],
                            "zc": [DevBuf(2 * 128 * m) for _ in range(2)]}
        L, Lb = self.lin, lib()
        wt = lambda k: L[k].wt.ptr
        _check_sn(Lb.uph_sn_fused(0, n, g.n_recv, None, None, None, 0.0, None, None, None, None, None, None,
                                  None, fb["h"].ptr, wt("p0_B"), wt("p0_C"), fb["zb"][0].ptr, fb["zc"][0].ptr,
                                  None, None, None, xn.ptr, wt("enc1"), wt("enc2"), None))
        for l in range(self.S):
            cur, nxt = l % 2, 1 - l % 2
            last = l == self.S - 1
            P = lambda k: wt(f"p{l}_{k}")
            _check_sn(Lb.uph_sn_fused(
                2 if last else 1, n, g.n_recv, g.off.ptr, g.send.ptr, pos.ptr, float(invR), L[f"p{l}_A"].w32.ptr,
                fb["zb"][cur].ptr, fb["zc"][cur].ptr, P("D"), P("E"), P("F"), P("G"), fb["h"].ptr,
                None if last else wt(f"p{l + 1}_B"), None if last else wt(f"p{l + 1}_C"),
                None if last else fb["zb"][nxt].ptr, None if last else fb["zc"][nxt].ptr,
                wt("dec1") if last else None, wt("dec2") if last else None, fb["out"].ptr if last else None,
                None, None, None, None))
        return fb["out"]

    def fused_output(self, n: int) -> np.ndarray:
        check(lib().uph_dev_sync())
        return self.fb["out"].download((n, 8), np.float32)[:, :OUT]
