"""The surrogate's rollout controller on the GPU: state exchange with the plain
solver, the density-drift fallback, and the ablation against a network whose
output is forced to zero (constant-velocity macro steps with the same
corrections)."""
import numpy as np
import pytest

from uph_amd import scenes
from uph_amd.devbuf import DevBuf

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-surrogate] {msg}", flush=True)


def test_state_exchange_roundtrip_and_restart():
    """Export then import without restart changes nothing (bit-identical continuation);
    import with restart equals a fresh solver built from the exported state."""
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(jitter=0.1, seed=2)
    n = sc.n
    a, b = GpuSolver(sc), GpuSolver(sc)
    a.step(30)
    b.step(30)
    pos, vel, rho = DevBuf(16 * n), DevBuf(16 * n), DevBuf(8 * n)
    a.state_to_device(pos.ptr, vel.ptr, rho.ptr)
    st = a.state()
    assert np.array_equal(pos.download((n, 2), np.float64), st["x"])
    assert np.array_equal(vel.download((n, 2), np.float64), st["v"])
    assert np.array_equal(rho.download(n, np.float64), st["rho"])
    a.state_from_device(pos.ptr, vel.ptr, rho.ptr, restart=False)
    a.step(20)
    b.step(20)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    # restart: the next step takes the first-step form with localVelocity = velocity
    a.state_to_device(pos.ptr, vel.ptr, rho.ptr)
    a.state_from_device(pos.ptr, vel.ptr, rho.ptr, restart=True)
    assert a.itime == 0
    s2 = sc.copy()
    cur = a.state()
    nf = sc.n_fluid
    s2.pos[:nf], s2.vel[:nf], s2.rho[:nf] = cur["x"][:nf], cur["v"][:nf], cur["rho"][:nf]
    s2.lvel[:nf] = cur["v"][:nf]
    c = GpuSolver(s2)
    a.step(25)
    c.step(25)
    sa, sc_ = a.state(), c.state()
    for k in sa:
        assert np.array_equal(sa[k], sc_[k]), k


def _small_training(iters):
    from uph_amd import surrogate as S
    from uph_amd.gnn import GraphNet
    cfg = S.SurrogateConfig(substeps=20, width=64, mp_steps=4, iters=iters)
    tr = S.training_scenes(4, 0.02, 2.0, seed=5, width=(0.2, 0.5), height=(1.0, 2.0))
    trajs = [S.record(sc, 120, cfg.substeps) for sc in tr]
    data = S.Dataset(tr, trajs, cfg.substeps * tr[0].dt)
    net = GraphNet(cfg.width, cfg.mp_steps, seed=0)
    losses = S.train(net, data, cfg)
    return S, cfg, data, net, losses


def test_fallback_on_density_drift_is_the_plain_solver():
    """A network with its decoder scaled by 1e3 scrambles the particles: every block
    exceeds the 1 % drift threshold or flings particles out of the tank, falls back,
    and the result is bit-identical to an uninterrupted plain run."""
    from uph_amd import surrogate as S
    from uph_amd.gnn import GraphNet
    from uph_amd.solver import GpuSolver
    cfg = S.SurrogateConfig(substeps=20, width=32, mp_steps=2)
    sc = scenes.dam_break()
    bad = GraphNet(cfg.width, cfg.mp_steps, seed=3)
    p = bad.params()
    p["dec2"] = p["dec2"] * 1e3
    bad = GraphNet(cfg.width, cfg.mp_steps, params=p)
    ro = S.Rollout(sc, bad, vscale=1.0, ascale=10.0, cfg=cfg)
    ro.run(40)
    _say(f"sabotaged: {ro.blocks} blocks, {ro.fallbacks} fallbacks, drifts {np.round(ro.drifts, 3)}")
    assert ro.blocks >= 3 and ro.fallbacks == ro.blocks
    assert ro.escapes + sum(d > cfg.drift_tol for d in ro.drifts) == ro.blocks
    ref = GpuSolver(sc)
    ref.step(ro.macro * cfg.substeps)
    a, b = ro.state(), ref.state()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_trained_network_beats_zero_network():
    """Ablation on the 5000-particle dam break: the trained network's rollout stays
    closer to the plain solver than constant-velocity macro steps do."""
    from uph_amd.gnn import GraphNet
    from uph_amd.solver import GpuSolver
    S, cfg, data, net, losses = _small_training(2000)
    l0, l1 = np.mean(losses[:50]), np.mean(losses[-100:])
    _say(f"training loss {l0:.3f} -> {l1:.3f}")
    assert l1 < 0.25 * l0
    sc = scenes.dam_break()
    res = {}
    for name, nn in (("trained", net),
                     ("zero", GraphNet(cfg.width, cfg.mp_steps,
                                       params={k: 0 * v for k, v in net.params().items()}))):
        ro = S.Rollout(sc, nn, data.vscale, data.ascale, cfg)
        ro.run(60)
        ref = GpuSolver(sc)
        ref.step(ro.macro * cfg.substeps)
        res[name] = S.errors(ro.state(), ref.state(), sc)
        _say(f"{name}: {ro.blocks} blocks, {ro.fallbacks} fallbacks, {res[name]}")
        st = ro.state()
        assert np.all(np.isfinite(st["x"])) and np.all(np.isfinite(st["rho"]))
    assert res["trained"]["pos_err_mean_dx"] < 0.7 * res["zero"]["pos_err_mean_dx"]


def test_restart_does_not_reuse_k_th_keys_of_earlier_steps():
    """After a restart the step counter starts again at 1.  The K-th keys that capped
    queries recorded before it must not pass for the new steps' keys: with K = 15 the
    dam-break lattice caps every interior particle (20 in range); the restart state
    spreads the right half of the column (spacing 1.35 dx, about 12 in range), so its
    particles no longer cap while the left half still does.  The restarted solver
    equals a fresh one built from the restart state."""
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break()
    sc.K = 15
    n, nf = sc.n, sc.n_fluid
    a = GpuSolver(sc)
    a.step(2)
    assert a.sync() > 0, "the cap should bind before the restart"
    cur = a.state()
    x = cur["x"].copy()
    right = np.arange(n) < nf
    right &= x[:, 0] > 0.5
    x[right] = np.array([0.5, 0.0]) + (x[right] - np.array([0.5, 0.0])) * 1.35
    pos, vel, rho = DevBuf.of(x), DevBuf.of(cur["v"]), DevBuf.of(cur["rho"])
    a.state_from_device(pos.ptr, vel.ptr, rho.ptr, restart=True)
    s2 = sc.copy()
    s2.pos[:nf], s2.vel[:nf], s2.rho[:nf] = x[:nf], cur["v"][:nf], cur["rho"][:nf]
    s2.lvel[:nf] = cur["v"][:nf]
    b = GpuSolver(s2)
    a.step(2)
    b.step(2)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
