"""Learned acceleration surrogate around the plain solver.

A graph network (uph_amd.gnn) predicts each fluid particle's acceleration over a
macro step of `substeps` plain time steps from its last five macro-step
velocities and the positions of its neighbours within the SPH support (2h).
The state advances by semi-implicit Euler on the prediction.  The rollout
controller keeps the plain solver authoritative:

* every `every` network macro steps, the state goes back to the plain solver
  (positions, macro velocities, and densities carried across the block as
  rho * rho_sum(end) / rho_sum(start), where rho_sum is the SPH summation
  density over the graph), which runs one macro step of plain steps: the
  correction;
* if the mean summation density of the fluid drifts by more than `drift_tol`
  (1 %) over a block, or anything is non-finite, the block is discarded and the
  plain solver, which still holds the state of the block's start, runs it
  instead.  So is a block that leaves a fluid particle outside the scene's box
  (a state the plain solver cannot continue).  The mean alone lets a block
  through where a few particles are flung apart, and the next blocks build on
  it, so by default a second test (`spike_tol`, None to switch it off) also
  rejects a block where one particle's summation density changes by more than
  spike_tol * rho0 (counting particles with at least 0.9 rho0 at either end, so
  that free-surface particles, whose support is cut, do not count).

Training data are plain-solver trajectories sampled every macro step; the
target is the second difference of positions, (x[t+1] - 2 x[t] + x[t-1]) / Dt^2,
with Gaussian noise on the input velocities (the target is corrected so that
the noisy state still lands on the true next position).

Scope: scenes with static walls (the network advances fluid particles only,
and the velocity history of a wall would have to come from its kinematics).
No counterpart in the reference; the plain path it wraps follows
sph.cc:90-168.
"""
from __future__ import annotations

import dataclasses
import json
import math
import time
from typing import List, Optional, Sequence

import numpy as np

from . import scenes as scn
from .devbuf import DevBuf, check, lib
from .gnn import Graph, GraphNet
from .solver import GpuSolver
from .surrogate_net import SurrogateNet

HIST = 5


@dataclasses.dataclass
class SurrogateConfig:
    substeps: int = 20          # plain steps per macro step
    width: int = 128
    mp_steps: int = 8           # message-passing steps
    radius_h: float = 2.0       # graph radius in units of h
    noise: float = 0.1          # input-velocity noise, in units of ascale * macro dt
    every: int = 10             # network macro steps between plain corrections
    drift_tol: float = 0.01     # fallback threshold on the mean summation density
    spike_tol: Optional[float] = 0.05   # ... and on any one particle's (see Rollout.block; None: off)
    arch: str = "rollout"       # "rollout": SurrogateNet (no edge latent); "gns": GraphNet
    lr: float = 1e-3
    lr_final: float = 1e-5
    iters: int = 2000
    seed: int = 0


def record(scene: scn.Scene, frames: int, substeps: int) -> np.ndarray:
    """Positions (frames + 1, n, 2) of a plain run, every `substeps` steps."""
    if scene.motion != scn.MOTION_NONE:
        raise ValueError("the surrogate covers scenes with static walls")
    g = GpuSolver(scene)
    X = np.empty((frames + 1, scene.n, 2))
    X[0] = g.state()["x"]
    for f in range(frames):
        g.step(substeps)
        X[f + 1] = g.state()["x"]
    g.close()
    return X


class Dataset:
    """Trajectories with their scenes; normalisation statistics over the fluid."""

    def __init__(self, scenes: Sequence[scn.Scene], trajs: Sequence[np.ndarray], dt_macro: float):
        self.scenes, self.trajs, self.dt = list(scenes), list(trajs), float(dt_macro)
        vs, acc = [], []
        for sc, X in zip(self.scenes, self.trajs):
            V = np.diff(X[:, :sc.n_fluid], axis=0) / self.dt
            vs.append(V.reshape(-1))
            acc.append((np.diff(V, axis=0) / self.dt).reshape(-1))
        self.vscale = float(np.sqrt(np.mean(np.concatenate(vs) ** 2)))
        self.ascale = float(np.sqrt(np.mean(np.concatenate(acc) ** 2)))

    def sample(self, s: int, t: int, rng: np.random.Generator, noise: float):
        """(positions, history 5 x n x 2, normalised target n x 2) at frame t of trajectory s."""
        sc, X = self.scenes[s], self.trajs[s]
        nf = sc.n_fluid
        hist = np.stack([(X[t - k] - X[t - k - 1]) / self.dt for k in range(HIST)])
        vnext = (X[t + 1] - X[t]) / self.dt
        if noise > 0:
            hist[:, :nf] += rng.normal(0.0, noise * self.ascale * self.dt, (HIST, nf, 2))
        tgt = np.zeros((sc.n, 2), np.float32)
        tgt[:nf] = ((vnext[:nf] - hist[0, :nf]) / self.dt / self.ascale).astype(np.float32)
        return X[t], np.ascontiguousarray(hist), tgt


class _Feats:
    """Per-scene device buffers for graph, features and inputs (edge features only for
    the GraphNet form; the rollout network forms them from the positions)."""

    def __init__(self, sc: scn.Scene, cap: int = 0, edge_feats: bool = False):
        n = sc.n
        self.n, self.nf = n, sc.n_fluid
        self.graph = Graph(n, sc.n_fluid, max(cap, 32 * n))
        self.pos = DevBuf(16 * n)
        self.hist = DevBuf(16 * HIST * n)
        self.xn = DevBuf(2 * 16 * n)
        self.edge_feats = edge_feats
        self.xe = DevBuf(2 * 8 * self.graph.cap) if edge_feats else None

    def build(self, dist: float, invR: float, inv_vscale: float, sender: bool = True):
        g = self.graph.build(self.pos, dist, sender=sender)
        check(lib().uph_gns_node_feats(self.n, self.nf, self.hist.ptr, inv_vscale, self.xn.ptr, None))
        if self.edge_feats:
            if self.xe.nbytes < 16 * g.E:
                self.xe = DevBuf(16 * g.cap)
            check(lib().uph_gns_edge_feats(g.E, g.send.ptr, g.recv.ptr, self.pos.ptr, invR, self.xe.ptr,
                                           None))
        return g


def make_net(cfg: SurrogateConfig, seed: int = 0):
    return (SurrogateNet(cfg.width, cfg.mp_steps, seed=seed) if cfg.arch == "rollout"
            else GraphNet(cfg.width, cfg.mp_steps, seed=seed))


def net_forward(net, f: "_Feats", g: Graph, invR: float, train: bool):
    if isinstance(net, SurrogateNet):
        if not train and net.W == 128:
            return net.forward_fused(g, f.pos, invR, f.xn)
        return net.forward(g, f.pos, invR, f.xn, train=train)
    return net.forward(g, f.xn, f.xe, train=train)


def train(net: GraphNet, data: Dataset, cfg: SurrogateConfig, log=None) -> List[float]:
    rng = np.random.default_rng(cfg.seed)
    feats = [_Feats(sc, edge_feats=not isinstance(net, SurrogateNet)) for sc in data.scenes]
    tg = [DevBuf(8 * sc.n) for sc in data.scenes]
    masks = [DevBuf.of((np.arange(sc.n) < sc.n_fluid).astype(np.int32)) for sc in data.scenes]
    losses = []
    decay = (cfg.lr_final / cfg.lr) ** (1.0 / max(cfg.iters - 1, 1))
    for it in range(cfg.iters):
        s = int(rng.integers(len(data.scenes)))
        sc = data.scenes[s]
        t = int(rng.integers(HIST, len(data.trajs[s]) - 1))
        x, hist, tgt = data.sample(s, t, rng, cfg.noise)
        f = feats[s]
        f.pos.upload(x)
        f.hist.upload(hist)
        tg[s].upload(tgt)
        R = cfg.radius_h * sc.h
        g = f.build(R * R, 1.0 / R, 1.0 / data.vscale)
        net_forward(net, f, g, 1.0 / R, train=True)
        net.zero_grad()
        losses.append(net.loss(tg[s], masks[s], sc.n_fluid))
        net.backward()
        net.adam(lr=cfg.lr * decay ** it)
        if log and (it % 200 == 0 or it == cfg.iters - 1):
            log(f"iter {it}: loss {np.mean(losses[-50:]):.4f}")
    return losses


class Rollout:
    """Network macro steps with plain corrections and density-drift fallback."""

    def __init__(self, scene: scn.Scene, net: Optional[GraphNet], vscale: float, ascale: float,
                 cfg: SurrogateConfig):
        if scene.motion != scn.MOTION_NONE:
            raise ValueError("the surrogate covers scenes with static walls")
        self.sc, self.net, self.cfg = scene, net, cfg
        self.vscale, self.ascale = vscale, ascale
        self.dt = cfg.substeps * scene.dt
        self.P = GpuSolver(scene)
        n, nf = scene.n, scene.n_fluid
        self.f = _Feats(scene, edge_feats=net is not None and not isinstance(net, SurrogateNet))
        self.vel = DevBuf(16 * n)
        self.rho = DevBuf(8 * n)
        self.rho2 = DevBuf(8 * n)
        self.prev = DevBuf(16 * n)
        self.hist2 = DevBuf(16 * HIST * n)
        self.mass = DevBuf.of(np.ascontiguousarray(scene.mass, np.float64))
        self.rs0, self.rs1 = DevBuf(8 * nf), DevBuf(8 * nf)
        self.blocks = self.fallbacks = self.macro = self.escapes = 0
        self.kept_drifts: List[float] = []
        self.drifts: List[float] = []
        self.spikes: List[float] = []
        self.t_net = self.t_plain = 0.0
        R = cfg.radius_h * scene.h
        self.dist, self.invR = R * R, 1.0 / R
        self._warm = False

    # ------------------------------------------------------------ helpers

    def _push_velocity(self, v_ptr: int):
        """hist <- [v, hist[0:4]]"""
        n = self.sc.n
        row = 16 * n
        check(lib().uph_dev_copy(self.hist2.ptr + row, self.f.hist.ptr, (HIST - 1) * row))
        check(lib().uph_dev_copy(self.hist2.ptr, v_ptr, row))
        self.f.hist, self.hist2 = self.hist2, self.f.hist

    def _plain_macro(self):
        """One macro step of the plain solver; pushes its macro velocity."""
        t0 = time.perf_counter()
        self.P.step(self.cfg.substeps)
        self.P.sync()
        self.t_plain += time.perf_counter() - t0
        self.P.state_to_device(self.f.pos.ptr, self.vel.ptr, self.rho.ptr)
        n2 = 2 * self.sc.n
        check(lib().uph_lincomb_f64(n2, 1.0 / self.dt, self.f.pos.ptr, -1.0 / self.dt,
                                    self.prev.ptr, self.vel.ptr, None))
        check(lib().uph_dev_copy(self.prev.ptr, self.f.pos.ptr, 8 * n2))
        self._push_velocity(self.vel.ptr)
        self.macro += 1

    def _rs(self, buf: DevBuf) -> np.ndarray:
        check(lib().uph_dev_sync())
        return buf.download(self.sc.n_fluid, np.float64)

    def warmup(self):
        """HIST plain macro steps fill the velocity history."""
        self.P.state_to_device(self.prev.ptr, self.vel.ptr, self.rho.ptr)
        self.f.hist.zero()
        for _ in range(HIST):
            self._plain_macro()
        self._warm = True

    # ------------------------------------------------------------ one block

    def block(self) -> bool:
        """`every` network macro steps and one plain correction, or the plain fallback.
        Returns True when the network's block was kept."""
        if not self._warm:
            self.warmup()
        cfg, sc, n, nf = self.cfg, self.sc, self.sc.n, self.sc.n_fluid
        L = lib()
        self.blocks += 1
        # the plain solver's state at the block start (it stays there until committed)
        self.P.state_to_device(self.f.pos.ptr, self.vel.ptr, self.rho.ptr)
        check(L.uph_dev_copy(self.vel.ptr, self.f.hist.ptr, 16 * n))     # macro velocity
        hist_start = DevBuf(16 * HIST * n)
        check(L.uph_dev_copy(hist_start.ptr, self.f.hist.ptr, 16 * HIST * n))
        prev_start = DevBuf(16 * n)
        check(L.uph_dev_copy(prev_start.ptr, self.prev.ptr, 16 * n))
        ok = True
        t0 = time.perf_counter()
        for k in range(cfg.every):
            g = self.f.build(self.dist, self.invR, 1.0 / self.vscale, sender=False)
            if k == 0:
                check(L.uph_sph_density(nf, g.off.ptr, g.send.ptr, self.f.pos.ptr, self.mass.ptr,
                                        sc.h, self.rs0.ptr, None))
            out = net_forward(self.net, self.f, g, self.invR, train=False)
            check(L.uph_gns_integrate(nf, self.f.pos.ptr, self.vel.ptr, out.ptr, 8, self.dt,
                                      self.ascale, None))
            self._push_velocity(self.vel.ptr)
        check(L.uph_dev_sync())
        self.t_net += time.perf_counter() - t0
        try:
            g = self.f.graph.build(self.f.pos, self.dist, sender=False)
        except RuntimeError:                  # non-finite positions
            ok = False
        if ok:
            # a state the plain solver can continue: every fluid position finite and inside
            # the scene's box (a particle flung out of the tank keeps no neighbours, and a
            # mean over the fluid hardly sees a few of them)
            x = self.f.pos.download((n, 2), np.float64)[:nf]
            x0, y0, x1, y1 = sc.box
            ok = bool(np.all(np.isfinite(x)) and np.all((x[:, 0] > x0) & (x[:, 0] < x1) &
                                                         (x[:, 1] > y0) & (x[:, 1] < y1)))
            if not ok:
                self.escapes += 1
        if ok:
            check(L.uph_sph_density(nf, g.off.ptr, g.send.ptr, self.f.pos.ptr, self.mass.ptr, sc.h,
                                    self.rs1.ptr, None))
            r0, r1 = self._rs(self.rs0), self._rs(self.rs1)
            drift = abs(r1.mean() - r0.mean()) / r0.mean()
            full = np.maximum(r0, r1) >= 0.9 * sc.rho0
            spike = float(np.abs(r1 - r0)[full].max() / sc.rho0) if full.any() else 0.0
            self.drifts.append(drift)
            self.spikes.append(spike)
            ok = (math.isfinite(drift) and math.isfinite(spike) and drift <= cfg.drift_tol
                  and (cfg.spike_tol is None or spike <= cfg.spike_tol))
            if ok:
                self.kept_drifts.append(drift)
        if ok:
            # densities carried across the block; the plain solver takes the state and
            # runs the correcting macro step
            check(L.uph_scale_ratio_f64(nf, self.rho.ptr, self.rs1.ptr, self.rs0.ptr, self.rho2.ptr,
                                        None))
            check(L.uph_dev_copy(self.prev.ptr, self.f.pos.ptr, 16 * n))
            self.P.state_from_device(self.f.pos.ptr, self.vel.ptr, self.rho2.ptr, restart=True)
            self.macro += cfg.every
            self._plain_macro()
            return True
        # fallback: the plain solver runs the block from its start
        self.fallbacks += 1
        check(L.uph_dev_copy(self.f.hist.ptr, hist_start.ptr, 16 * HIST * n))
        check(L.uph_dev_copy(self.prev.ptr, prev_start.ptr, 16 * n))
        for _ in range(cfg.every + 1):
            self._plain_macro()
        return False

    def run(self, macro_steps: int):
        """Advances by at least `macro_steps` macro steps (whole blocks after the warm-up)."""
        if not self._warm:
            self.warmup()
        while self.macro < macro_steps:
            self.block()
        check(lib().uph_dev_sync())

    def state(self):
        return self.P.state()


def save(path: str, net: GraphNet, data: Dataset, cfg: SurrogateConfig):
    d = net.state_dict()
    d["_vscale"], d["_ascale"] = np.array(data.vscale), np.array(data.ascale)
    d["_cfg"] = np.array(json.dumps(dataclasses.asdict(cfg)))
    np.savez(path, **d)


def load(path: str):
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    cfg = SurrogateConfig(**json.loads(str(d.pop("_cfg"))))
    vs, as_ = float(d.pop("_vscale")), float(d.pop("_ascale"))
    arch = str(d.pop("_arch", "gns"))
    net = SurrogateNet.from_state(d) if arch == "rollout" else GraphNet.from_state(d)
    return net, vs, as_, cfg


def training_scenes(n_scenes: int, dx: float, eos_height: float, seed: int,
                    width=(0.15, 0.4), height=(0.8, 2.0)) -> List[scn.Scene]:
    """Narrow dam-break columns at one resolution and sound speed (seeded sizes)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        w = float(rng.uniform(*width))
        hgt = float(rng.uniform(*height))
        out.append(scn.dam_break(width=w, height=hgt, dx=dx, eos_height=eos_height,
                                 name=f"train_{k}"))
    return out


def errors(a: dict, b: dict, sc: scn.Scene) -> dict:
    nf = sc.n_fluid
    dxy = np.linalg.norm(a["x"][:nf] - b["x"][:nf], axis=1) / sc.dx
    drho = np.abs(a["rho"][:nf] - b["rho"][:nf]) / b["rho"][:nf]
    return dict(pos_err_mean_dx=float(dxy.mean()), pos_err_max_dx=float(dxy.max()),
                rho_err_max=float(drho.max()))


def time_macro_step(dx: float = 0.005, width: float = 0.8, height: float = 2.0, reps: int = 20) -> dict:
    """The rollout network's macro step on a dam break (default: the 64,000-particle
    one) against the plain solver, with initial weights (the cost does not depend on
    them): ms for the radius-graph build, the node features and the fused W128 / S8
    forward, and for 20 plain steps; HIP work timed by device synchronisation."""
    sc = scn.dam_break(width=width, height=height, dx=dx)
    g = GpuSolver(sc)
    g.step(50)
    st = g.state()
    n, nf = sc.n, sc.n_fluid
    pos = DevBuf.of(np.ascontiguousarray(st["x"]))
    R = 2.0 * sc.h
    graph = Graph(n, nf, 32 * n)
    hist = DevBuf.of(np.random.default_rng(0).normal(0.0, 0.1, (HIST, n, 2)))
    xn = DevBuf(2 * 16 * n)
    net = SurrogateNet(128, 8)

    def t(f, k=reps):
        f()
        check(lib().uph_dev_sync())
        t0 = time.perf_counter()
        for _ in range(k):
            f()
        check(lib().uph_dev_sync())
        return 1e3 * (time.perf_counter() - t0) / k
    out = {"fluid": nf, "particles": n, "edges": graph.build(pos, R * R, sender=False).E}
    out["graph_ms"] = t(lambda: graph.build(pos, R * R, sender=False))
    out["node_feats_ms"] = t(lambda: check(lib().uph_gns_node_feats(n, nf, hist.ptr, 1.0, xn.ptr, None)))
    out["fused_forward_ms"] = t(lambda: net.forward_fused(graph, pos, 1.0 / R, xn))
    out["plain_20_steps_ms"] = t(lambda: (g.step(20), g.sync()), max(reps // 4, 1))
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Train the surrogate on dam breaks and measure a "
                                             "corrected rollout against the plain solver.")
    ap.add_argument("--dx", type=float, default=0.005)
    ap.add_argument("--width", type=float, default=0.8, help="rollout column width (m)")
    ap.add_argument("--height", type=float, default=2.0, help="rollout column height (m)")
    ap.add_argument("--train-scenes", type=int, default=6)
    ap.add_argument("--frames", type=int, default=400)
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--net-width", type=int, default=128)
    ap.add_argument("--mp-steps", type=int, default=8)
    ap.add_argument("--substeps", type=int, default=20)
    ap.add_argument("--macro-steps", type=int, default=400)
    ap.add_argument("--model", default=None, help="save the trained model (.npz)")
    ap.add_argument("--arch", choices=["rollout", "gns"], default="rollout",
                    help="rollout: the network without edge latents (default); gns: edge latents")
    ap.add_argument("--spike-tol", type=float, default=0.05,
                    help="also fall back when one particle's summation density moves by more than "
                         "this fraction of rho0 over a block (0: off)")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--time-only", action="store_true",
                    help="only time the network's macro step against 20 plain steps (time_macro_step)")
    a = ap.parse_args(argv)
    if a.time_only:
        res = time_macro_step(dx=a.dx, width=a.width, height=a.height)
        print(json.dumps(res), flush=True)
        return res
    say = lambda m: print(f"[surrogate] {m}", flush=True)
    cfg = SurrogateConfig(substeps=a.substeps, width=a.net_width, mp_steps=a.mp_steps,
                          iters=a.iters, arch=a.arch, spike_tol=a.spike_tol or None, lr=a.lr)
    tr = training_scenes(a.train_scenes, a.dx, a.height, seed=20261203)
    t0 = time.perf_counter()
    trajs = [record(sc, a.frames, cfg.substeps) for sc in tr]
    data = Dataset(tr, trajs, cfg.substeps * tr[0].dt)
    t_rec = time.perf_counter() - t0
    say(f"recorded {len(tr)} trajectories ({sum(s.n_fluid for s in tr)} fluid) x {a.frames} "
        f"frames in {t_rec:.1f} s; vscale {data.vscale:.3f} m/s, ascale {data.ascale:.2f} m/s^2")
    net = make_net(cfg, seed=cfg.seed)
    t0 = time.perf_counter()
    losses = train(net, data, cfg, log=say)
    t_train = time.perf_counter() - t0
    if a.model:
        save(a.model, net, data, cfg)
    sc = scn.dam_break(width=a.width, height=a.height, dx=a.dx)
    assert abs(sc.dt - tr[0].dt) < 1e-15 * sc.dt
    # plain reference over the same span
    ref = GpuSolver(sc)
    ro = Rollout(sc, net, data.vscale, data.ascale, cfg)
    ro.warmup()
    steps_total = None
    check(lib().uph_dev_sync())
    t0 = time.perf_counter()
    ro.run(a.macro_steps)
    t_roll = time.perf_counter() - t0
    say(f"rollout: {ro.blocks} blocks, {ro.fallbacks} fallbacks; drift per block "
        f"{np.round(ro.drifts, 4).tolist()}; spike per block {np.round(ro.spikes, 4).tolist()}")
    steps_total = ro.macro * cfg.substeps
    t0 = time.perf_counter()
    ref.step(steps_total)
    ref.sync()
    t_plain = time.perf_counter() - t0
    err = errors(ro.state(), ref.state(), sc)
    # ablation: the same controller with the network's output forced to zero
    zero = type(net)(cfg.width, cfg.mp_steps, params={k: 0 * v for k, v in net.params().items()})
    rz = Rollout(sc, zero, data.vscale, data.ascale, cfg)
    rz.run(a.macro_steps)
    err0 = errors(rz.state(), ref.state(), sc)
    nf = sc.n_fluid
    rollout_span = (ro.macro - HIST) * cfg.substeps
    res = {
        "workload": f"dam break {a.width}x{a.height} m, dx={a.dx}, {nf} fluid",
        "substeps": cfg.substeps, "every": cfg.every, "net": f"W{cfg.width} S{cfg.mp_steps}",
        "train": {"scenes": len(tr), "frames": a.frames, "iters": cfg.iters,
                  "loss_first50": float(np.mean(losses[:50])),
                  "loss_last50": float(np.mean(losses[-50:])), "seconds": t_train},
        "rollout": {"macro_steps": ro.macro, "blocks": ro.blocks, "fallbacks": ro.fallbacks,
                    "max_drift": float(max(ro.drifts) if ro.drifts else 0.0),
                    "max_spike": float(max(ro.spikes) if ro.spikes else 0.0),
                    "seconds": t_roll, "seconds_network": ro.t_net,
                    "seconds_plain": ro.t_plain,
                    "ms_per_network_macro_step": 1e3 * ro.t_net / max(
                        (ro.blocks) * cfg.every, 1),
                    "plain_dt_equiv_updates_per_s": nf * rollout_span / t_roll, **err},
        "zero_network": {"fallbacks": rz.fallbacks, **err0},
        "plain": {"steps": steps_total, "seconds": t_plain,
                  "updates_per_s": nf * steps_total / t_plain},
    }
    print(json.dumps(res), flush=True)
    return res


if __name__ == "__main__":
    main()
