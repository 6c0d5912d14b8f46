"""The surrogate graph network on the GPU: forward, backward and Adam.

Same architecture as the oracle statement (encode-process-decode with L
message-passing steps, width W; see oracle/gnn_ref.py for the equations).
Every dense layer is one call of the library's BF16 MFMA GEMM (FP32
accumulation); biases are the weight row that multiplies a constant-1 column
kept at index W of every BF16 activation buffer (row stride W + 8).  Latent
streams (residuals) stay in FP32; aggregations are segment sums over the
receiver-ordered CSR, and their transposes use the sender-ordered
permutation, so a forward/backward pass is deterministic.
"""
from __future__ import annotations

import math
from typing import Dict, Optional

import numpy as np

from .devbuf import DevBuf, check, from_bf16, lib, to_bf16

NODE_IN, EDGE_IN, OUT = 16, 8, 2
FUSED_WIDTH = 128          # width the fused edge kernel (uph_gns_edge_step) is built for


def _pad8(x: int) -> int:
    return (x + 7) // 8 * 8


class Graph:
    """Radius graph on device: CSR by receiver plus the sender permutation."""

    def __init__(self, n: int, n_recv: int, capacity: int):
        self.n, self.n_recv = n, n_recv
        self.cap = capacity
        self.off = DevBuf(4 * (n + 1))
        self.soff = DevBuf(4 * (n + 1))
        self.send = DevBuf(4 * capacity)
        self.recv = DevBuf(4 * capacity)
        self.sperm = DevBuf(4 * capacity)
        self.E = 0

    def build(self, pos: DevBuf, dist: float, sender: bool = True):
        """sender=False skips the sender view (soff, sperm), which only backward passes read."""
        import ctypes as C
        ne = C.c_long()
        so = lambda: (self.soff.ptr, self.sperm.ptr) if sender else (None, None)
        rc = lib().uph_graph_build(pos.ptr, self.n, self.n_recv, float(dist), self.off.ptr,
                                   self.send.ptr, self.recv.ptr, *so(), self.cap, C.byref(ne))
        if rc == -2:
            self.cap = int(ne.value * 1.25) + 64
            self.send, self.recv, self.sperm = (DevBuf(4 * self.cap), DevBuf(4 * self.cap),
                                                DevBuf(4 * self.cap))
            rc = lib().uph_graph_build(pos.ptr, self.n, self.n_recv, float(dist), self.off.ptr,
                                       self.send.ptr, self.recv.ptr, *so(), self.cap, C.byref(ne))
        if rc != 0:
            raise RuntimeError("uph_graph_build: " + lib().uph_graph_last_error().decode())
        self.E = int(ne.value)
        return self


class Lin:
    """Weight matrix padded to (Kp x Np); FP32 master, BF16 copy, BF16 transpose, Adam state."""

    def __init__(self, K: int, N: int):
        self.K, self.N = K, N
        self.Kp, self.Np = _pad8(K), _pad8(N)
        sz = self.Kp * self.Np
        self.w32 = DevBuf(4 * sz)
        self.wb = DevBuf(2 * sz)
        self.wt = DevBuf(2 * sz)
        self.g = DevBuf(4 * sz)
        self.m = DevBuf(4 * sz)
        self.v = DevBuf(4 * sz)
        for b in (self.g, self.m, self.v):
            b.zero()

    def set(self, w: np.ndarray):
        full = np.zeros((self.Kp, self.Np), np.float32)
        full[:w.shape[0], :w.shape[1]] = w
        self.w32.upload(full)
        self.sync_copies()

    def get(self) -> np.ndarray:
        return self.w32.download((self.Kp, self.Np), np.float32)[:self.K, :self.N]

    def grad(self) -> np.ndarray:
        return self.g.download((self.Kp, self.Np), np.float32)[:self.K, :self.N]

    def sync_copies(self):
        n = self.Kp * self.Np
        check(lib().uph_cast_f32_bf16(n, self.w32.ptr, self.wb.ptr, None))
        check(lib().uph_transpose_bf16(self.Kp, self.Np, self.w32.ptr, self.wt.ptr, None))


def _init_params(width: int, steps: int, seed: int):
    """He-normal weights with zero bias rows; residual-branch outputs scaled by 0.1."""
    rng = np.random.default_rng(seed)
    W = width
    shapes = {"enc_n1": (NODE_IN, W), "enc_n2": (W + 1, W), "enc_e1": (EDGE_IN, W),
              "enc_e2": (W + 1, W), "dec1": (W + 1, W), "dec2": (W + 1, OUT)}
    for l in range(steps):
        shapes.update({f"p{l}_A": (W + 1, W), f"p{l}_B": (W + 1, W), f"p{l}_C": (W + 1, W),
                       f"p{l}_D": (W + 1, W), f"p{l}_E": (W + 1, W), f"p{l}_F": (W, W),
                       f"p{l}_G": (W + 1, W)})
    out = {}
    for name, (r, c) in shapes.items():
        fan_in = r - 1 if r == W + 1 else r
        w = rng.normal(0.0, math.sqrt(2.0 / max(fan_in, 1)), (r, c))
        if r == W + 1:
            w[-1] = 0.0
        if name == "enc_n1":
            w[11] = 0.0
        if name == "enc_e1":
            w[3] = 0.0
        if name.endswith("_D") or name.endswith("_G"):
            w *= 0.1
        out[name] = w
    return out, shapes


class GraphNet:
    """Encode-process-decode graph network with device buffers sized on demand."""

    def __init__(self, width: int = 128, steps: int = 8, params: Optional[Dict] = None,
                 seed: int = 0):
        self.W, self.S = width, steps
        init, shapes = _init_params(width, steps, seed)
        params = params if params is not None else init
        self.lin = {k: Lin(*shapes[k]) for k in shapes}
        for k, L in self.lin.items():
            L.set(np.asarray(params[k], np.float32))
        self.adam_t = 0
        self._cap = (0, 0, False)
        self.work = None

    # ------------------------------------------------------------ parameters

    def params(self) -> Dict[str, np.ndarray]:
        return {k: L.get() for k, L in self.lin.items()}

    def grads(self) -> Dict[str, np.ndarray]:
        return {k: L.grad() for k, L in self.lin.items()}

    def zero_grad(self):
        for L in self.lin.values():
            L.g.zero()

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.adam_t += 1
        for L in self.lin.values():
            n = L.Kp * L.Np
            check(lib().uph_adam(n, L.w32.ptr, L.g.ptr, L.m.ptr, L.v.ptr, L.wb.ptr, lr, b1, b2,
                                 eps, self.adam_t, None))
            check(lib().uph_transpose_bf16(L.Kp, L.Np, L.w32.ptr, L.wt.ptr, None))

    def state_dict(self) -> Dict[str, np.ndarray]:
        d = self.params()
        d["_width"] = np.array(self.W)
        d["_steps"] = np.array(self.S)
        return d

    @classmethod
    def from_state(cls, d) -> "GraphNet":
        W, S = int(d["_width"]), int(d["_steps"])
        return cls(W, S, params={k: d[k] for k in d if not k.startswith("_")})

    # ------------------------------------------------------------ buffers

    def _ensure(self, n: int, E: int, train: bool):
        n, E = max(n, 1), max(E, 1)
        cn, cE, ct = self._cap
        if n <= cn and E <= cE and (ct or not train):
            return
        # grow with headroom: rollouts see the edge count fluctuate step to step
        n = max(n if n <= cn else int(n * 1.1) + 64, cn)
        E = max(E if E <= cE else int(E * 1.25) + 1024, cE)
        train = bool(train or ct)
        W, Wl = self.W, self.W + 8
        nb = lambda rows, cols, es: DevBuf(rows * cols * es)
        self.b = b = {}
        b["h"], b["e"] = nb(n, W, 4), nb(E, W, 4)                # FP32 latent streams
        b["zn"], b["ze"] = nb(n, Wl, 4), nb(E, Wl, 4)            # FP32 GEMM outputs
        b["zb"], b["zc"] = nb(n, W, 4), nb(n, W, 4)
        b["m"], b["a"] = nb(E, W, 4), nb(n, W, 4)
        b["out"] = nb(n, 8, 4)
        nsave = self.S + 1 if train else 1
        # BF16 activations carrying the constant-1 column at index W; training keeps
        # one set per message-passing step for the backward pass
        for key, rows in (("hb", n), ("eb", E), ("u1", E), ("u2", n), ("un", n), ("ue", E),
                          ("up", n)):
            k = nsave if key in ("hb", "eb", "u1", "u2") else 1
            b[key] = [nb(rows, Wl, 2) for _ in range(k)]
            for buf in b[key]:
                buf.zero()
                check(lib().uph_fill_col_bf16(rows, Wl, W, 1.0, buf.ptr, None))
        b["abs"] = [nb(n, W, 2) for _ in range(nsave)]
        if train:
            for key, rows, cols, es in (("dh", n, Wl, 4), ("de", E, W, 4), ("t1", n, Wl, 4),
                                        ("t2", E, Wl, 4), ("dm", E, W, 4), ("da", n, W, 4),
                                        ("dzb", n, W, 4), ("dzc", n, W, 4), ("gb16n", n, Wl, 2),
                                        ("gb16e", E, Wl, 2), ("gb16n2", n, Wl, 2),
                                        ("dout", n, 8, 4), ("doutb", n, 8, 2),
                                        ("lpart", (n * 8 + 255) // 256 + 1, 1, 4)):
                b[key] = nb(rows, cols, es)
            for key in ("gb16n", "gb16e", "gb16n2"):
                b[key].zero()
        self.work = DevBuf(4 * max(n * Wl, E * W, 64 * Wl * Wl))
        self._cap = (n, E, train)

    # ------------------------------------------------------------ GEMM helpers

    def _gemm(self, M, N, K, A, lda, B, ldb, C, ldc, trans_a=0, trans_b=0, out_bf16=0, act=0,
              accumulate=0):
        """trans_b: B is given as its N x K transpose (the weight copies kept for that)."""
        splits = 1
        if trans_a:
            splits = max(1, min(64, K // 4096))
        work = self.work.ptr if (splits > 1 or accumulate) else None
        check(lib().uph_gemm_bf16(M, N, K, A, lda, int(trans_a) | (int(trans_b) << 1), B, ldb, None,
                                  C, ldc, out_bf16, act, splits, work, accumulate, None))

    def _lin_fwd(self, L: Lin, X, ldx, rows, out, ldo):
        """out (FP32) = X (BF16, rows x L.Kp) . W   (W read through its transpose)"""
        self._gemm(rows, L.Np, L.Kp, X, ldx, L.wt.ptr, L.Kp, out, ldo, trans_b=1)

    def _lin_dw(self, L: Lin, X, ldx, rows, dZb, ldz):
        """grad += X^T dZ  (X BF16 rows x Kp, dZ BF16 rows x Np)"""
        self._gemm(L.Kp, L.Np, rows, X, ldx, dZb, ldz, L.g.ptr, L.Np, trans_a=1, accumulate=1)

    def _lin_dx(self, L: Lin, dZb, ldz, rows, out, ldo):
        """out (FP32, rows x Kp) = dZ W^T   (W^T given through W, its N x K form)"""
        self._gemm(rows, L.Kp, L.Np, dZb, ldz, L.wb.ptr, L.Np, out, ldo, trans_b=1)

    # ------------------------------------------------------------ forward

    def forward(self, g: Graph, xn: DevBuf, xe: DevBuf, train: bool = False,
                fused: Optional[bool] = None) -> DevBuf:
        """fused (default: width 128): the edge MLP of every step as one kernel."""
        n, E, W, Wl = g.n, g.E, self.W, self.W + 8
        fused = (W == FUSED_WIDTH) if fused is None else bool(fused)
        if fused and W != FUSED_WIDTH:
            raise ValueError(f"the fused edge kernel is built for width {FUSED_WIDTH}")
        self._ensure(n, E, train)
        b, L = self.b, self.lin
        Lb = lib()
        # encoders
        self._lin_fwd(L["enc_n1"], xn.ptr, NODE_IN, n, b["zn"].ptr, Wl)
        check(Lb.uph_relu_to_bf16(n, W, b["zn"].ptr, Wl, b["un"][0].ptr, Wl, None))
        self._lin_fwd(L["enc_n2"], b["un"][0].ptr, Wl, n, b["h"].ptr, W)
        self._lin_fwd(L["enc_e1"], xe.ptr, EDGE_IN, E, b["ze"].ptr, Wl)
        check(Lb.uph_relu_to_bf16(E, W, b["ze"].ptr, Wl, b["ue"][0].ptr, Wl, None))
        self._lin_fwd(L["enc_e2"], b["ue"][0].ptr, Wl, E, b["e"].ptr, W)
        check(Lb.uph_rows_to_bf16(n, W, b["h"].ptr, W, b["hb"][0].ptr, Wl, None))
        check(Lb.uph_rows_to_bf16(E, W, b["e"].ptr, W, b["eb"][0].ptr, Wl, None))
        for l in range(self.S):
            s = l if train else 0
            s1 = l + 1 if train else 0
            P = lambda k: L[f"p{l}_{k}"]
            hb, eb = b["hb"][s].ptr, b["eb"][s].ptr
            self._lin_fwd(P("B"), hb, Wl, n, b["zb"].ptr, W)
            self._lin_fwd(P("C"), hb, Wl, n, b["zc"].ptr, W)
            u1 = b["u1"][s].ptr
            if fused:
                # both edge GEMMs, the gathers and the residual update in one kernel
                check(Lb.uph_gns_edge_step(E, eb, b["eb"][s1].ptr, b["e"].ptr, b["zb"].ptr,
                                           b["zc"].ptr, g.recv.ptr, g.send.ptr, P("A").wt.ptr,
                                           P("D").wt.ptr, b["m"].ptr, u1 if train else None,
                                           None))
            else:
                self._lin_fwd(P("A"), eb, Wl, E, b["ze"].ptr, Wl)
                check(Lb.uph_gather_add2(E, W, g.recv.ptr, b["zb"].ptr, g.send.ptr, b["zc"].ptr,
                                         W, b["ze"].ptr, Wl, None))
                check(Lb.uph_relu_to_bf16(E, W, b["ze"].ptr, Wl, u1, Wl, None))
                self._lin_fwd(P("D"), u1, Wl, E, b["m"].ptr, W)
            check(Lb.uph_segsum(n, W, g.off.ptr, None, b["m"].ptr, W, 0, b["a"].ptr, W, 0, None))
            ab = b["abs"][s].ptr
            check(Lb.uph_rows_to_bf16(n, W, b["a"].ptr, W, ab, W, None))
            self._lin_fwd(P("E"), hb, Wl, n, b["zn"].ptr, Wl)
            self._gemm(n, W, W, ab, W, P("F").wt.ptr, P("F").Kp, b["zn"].ptr, Wl, trans_b=1,
                       accumulate=1)
            u2 = b["u2"][s].ptr
            check(Lb.uph_relu_to_bf16(n, W, b["zn"].ptr, Wl, u2, Wl, None))
            self._lin_fwd(P("G"), u2, Wl, n, b["zb"].ptr, W)
            if not fused:
                check(Lb.uph_add_f32(E * W, b["m"].ptr, b["e"].ptr, None))
                check(Lb.uph_rows_to_bf16(E, W, b["e"].ptr, W, b["eb"][s1].ptr, Wl, None))
            check(Lb.uph_add_f32(n * W, b["zb"].ptr, b["h"].ptr, None))
            check(Lb.uph_rows_to_bf16(n, W, b["h"].ptr, W, b["hb"][s1].ptr, Wl, None))
        hbf = b["hb"][self.S if train else 0].ptr
        self._lin_fwd(L["dec1"], hbf, Wl, n, b["zn"].ptr, Wl)
        check(Lb.uph_relu_to_bf16(n, W, b["zn"].ptr, Wl, b["up"][0].ptr, Wl, None))
        self._lin_fwd(L["dec2"], b["up"][0].ptr, Wl, n, b["out"].ptr, 8)
        self._g, self._xn, self._xe = g, xn, xe
        return b["out"]

    def output(self, n: int) -> np.ndarray:
        check(lib().uph_dev_sync())
        return self.b["out"].download((n, 8), np.float32)[:, :OUT]

    # ------------------------------------------------------------ loss / backward

    def loss(self, target: DevBuf, mask: DevBuf, count: int) -> float:
        """Masked MSE of the last forward output; leaves d_out in the buffers."""
        g, b = self._g, self.b
        n = g.n
        b["dout"].zero()
        nbk = (n * OUT + 255) // 256
        check(lib().uph_mse(n, OUT, b["out"].ptr, 8, target.ptr, OUT, mask.ptr, count * OUT,
                            b["dout"].ptr, b["lpart"].ptr, None))
        check(lib().uph_dev_sync())
        return float(b["lpart"].download(nbk, np.float32).sum())

    def backward(self, d_out: Optional[DevBuf] = None):
        """Accumulate weight gradients of sum(d_out * out) (default: the loss's d_out)."""
        g, b, L = self._g, self.b, self.lin
        Lb = lib()
        n, E, W, Wl = g.n, g.E, self.W, self.W + 8
        dout = d_out if d_out is not None else b["dout"]
        check(Lb.uph_rows_to_bf16(n, 8, dout.ptr, 8, b["doutb"].ptr, 8, None))
        # decoder
        self._lin_dw(L["dec2"], b["up"][0].ptr, Wl, n, b["doutb"].ptr, 8)
        self._lin_dx(L["dec2"], b["doutb"].ptr, 8, n, b["t1"].ptr, Wl)
        check(Lb.uph_relu_bwd_rows(n, W, b["t1"].ptr, Wl, b["up"][0].ptr, Wl, b["gb16n"].ptr, Wl,
                                   None))
        self._lin_dw(L["dec1"], b["hb"][self.S].ptr, Wl, n, b["gb16n"].ptr, Wl)
        self._lin_dx(L["dec1"], b["gb16n"].ptr, Wl, n, b["dh"].ptr, Wl)
        b["de"].zero()
        for l in reversed(range(self.S)):
            P = lambda k: L[f"p{l}_{k}"]
            hb, eb, u1, u2, ab = (b["hb"][l].ptr, b["eb"][l].ptr, b["u1"][l].ptr, b["u2"][l].ptr,
                                  b["abs"][l].ptr)
            # node update
            check(Lb.uph_rows_to_bf16(n, W, b["dh"].ptr, Wl, b["gb16n"].ptr, Wl, None))
            self._lin_dw(P("G"), u2, Wl, n, b["gb16n"].ptr, Wl)
            self._lin_dx(P("G"), b["gb16n"].ptr, Wl, n, b["t1"].ptr, Wl)
            check(Lb.uph_relu_bwd_rows(n, W, b["t1"].ptr, Wl, u2, Wl, b["gb16n2"].ptr, Wl, None))
            dy = b["gb16n2"].ptr
            self._lin_dw(P("E"), hb, Wl, n, dy, Wl)
            self._lin_dw(P("F"), ab, W, n, dy, Wl)
            self._lin_dx(P("E"), dy, Wl, n, b["t1"].ptr, Wl)
            check(Lb.uph_add_rows_f32(n, W, b["t1"].ptr, Wl, b["dh"].ptr, Wl, None))
            self._gemm(n, W, W, dy, Wl, P("F").wb.ptr, P("F").Np, b["da"].ptr, W, trans_b=1)
            # edge update: dm = de + da[recv]
            check(Lb.uph_gather_rows(E, W, g.recv.ptr, b["da"].ptr, W, b["dm"].ptr, W, 4, None))
            check(Lb.uph_add_f32(E * W, b["de"].ptr, b["dm"].ptr, None))
            check(Lb.uph_rows_to_bf16(E, W, b["dm"].ptr, W, b["gb16e"].ptr, Wl, None))
            self._lin_dw(P("D"), u1, Wl, E, b["gb16e"].ptr, Wl)
            self._lin_dx(P("D"), b["gb16e"].ptr, Wl, E, b["t2"].ptr, Wl)
            check(Lb.uph_relu_bwd_rows(E, W, b["t2"].ptr, Wl, u1, Wl, b["gb16e"].ptr, Wl, None))
            dz1 = b["gb16e"].ptr
            self._lin_dw(P("A"), eb, Wl, E, dz1, Wl)
            self._lin_dx(P("A"), dz1, Wl, E, b["t2"].ptr, Wl)
            check(Lb.uph_add_rows_f32(E, W, b["t2"].ptr, Wl, b["de"].ptr, W, None))
            check(Lb.uph_segsum(n, W, g.off.ptr, None, dz1, Wl, 1, b["dzb"].ptr, W, 0, None))
            check(Lb.uph_segsum(n, W, g.soff.ptr, g.sperm.ptr, dz1, Wl, 1, b["dzc"].ptr, W, 0,
                                None))
            check(Lb.uph_rows_to_bf16(n, W, b["dzb"].ptr, W, b["gb16n"].ptr, Wl, None))
            check(Lb.uph_rows_to_bf16(n, W, b["dzc"].ptr, W, b["gb16n2"].ptr, Wl, None))
            self._lin_dw(P("B"), hb, Wl, n, b["gb16n"].ptr, Wl)
            self._lin_dw(P("C"), hb, Wl, n, b["gb16n2"].ptr, Wl)
            self._lin_dx(P("B"), b["gb16n"].ptr, Wl, n, b["t1"].ptr, Wl)
            check(Lb.uph_add_rows_f32(n, W, b["t1"].ptr, Wl, b["dh"].ptr, Wl, None))
            self._lin_dx(P("C"), b["gb16n2"].ptr, Wl, n, b["t1"].ptr, Wl)
            check(Lb.uph_add_rows_f32(n, W, b["t1"].ptr, Wl, b["dh"].ptr, Wl, None))
        # encoders
        check(Lb.uph_rows_to_bf16(E, W, b["de"].ptr, W, b["gb16e"].ptr, Wl, None))
        self._lin_dw(L["enc_e2"], b["ue"][0].ptr, Wl, E, b["gb16e"].ptr, Wl)
        self._lin_dx(L["enc_e2"], b["gb16e"].ptr, Wl, E, b["t2"].ptr, Wl)
        check(Lb.uph_relu_bwd_rows(E, W, b["t2"].ptr, Wl, b["ue"][0].ptr, Wl, b["gb16e"].ptr, Wl,
                                   None))
        self._lin_dw(L["enc_e1"], self._xe.ptr, EDGE_IN, E, b["gb16e"].ptr, Wl)
        check(Lb.uph_rows_to_bf16(n, W, b["dh"].ptr, Wl, b["gb16n"].ptr, Wl, None))
        self._lin_dw(L["enc_n2"], b["un"][0].ptr, Wl, n, b["gb16n"].ptr, Wl)
        self._lin_dx(L["enc_n2"], b["gb16n"].ptr, Wl, n, b["t1"].ptr, Wl)
        check(Lb.uph_relu_bwd_rows(n, W, b["t1"].ptr, Wl, b["un"][0].ptr, Wl, b["gb16n2"].ptr,
                                   Wl, None))
        self._lin_dw(L["enc_n1"], self._xn.ptr, NODE_IN, n, b["gb16n2"].ptr, Wl)
