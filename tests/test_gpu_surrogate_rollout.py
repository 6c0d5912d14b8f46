"""The learned surrogate at its workload: the W128 / S8 rollout network (uph_amd/
surrogate_net.py, the fused inference path at width 128) trained on the GPU by the
project's own forward and backward kernels on plain-solver dam-break trajectories at the
rollout's resolution, then rolled out on a 64,000-particle dam break (0.8 x 2.0 m at
dx = 5 mm) with a plain-solver correction every 10 network macro steps (20 plain steps
each) and the fallback to the plain solver.

Reported (and asserted only where the controller guarantees it): plain-dt-equivalent
throughput against the plain GPU path on the same span, density and position errors
against an uninterrupted plain run, blocks kept and fallen back, under the stated 1 %
mean-drift rule alone and with the per-particle test added.  The ablation runs with the
fallback disabled, so it measures the network rather than the plain solver: the trained
network against the same controller with the network's output forced to zero."""
import dataclasses
import time

import numpy as np
import pytest

from uph_amd import scenes

pytestmark = pytest.mark.gpu

TRAIN_SCENES, FRAMES, ITERS = 6, 200, 2000


def _say(msg):
    print(f"[gpu-rollout] {msg}", flush=True)


@pytest.fixture(scope="module")
def trained():
    from uph_amd import surrogate as S
    cfg = S.SurrogateConfig(substeps=20, width=128, mp_steps=8, iters=ITERS)
    tr = S.training_scenes(TRAIN_SCENES, 0.005, 2.0, seed=20261203)
    t0 = time.perf_counter()
    trajs = [S.record(sc, FRAMES, cfg.substeps) for sc in tr]
    data = S.Dataset(tr, trajs, cfg.substeps * tr[0].dt)
    net = S.make_net(cfg, seed=0)
    losses = S.train(net, data, cfg)
    t = time.perf_counter() - t0
    _say(f"trained W128/S8 on {len(tr)} columns ({sum(s.n_fluid for s in tr)} fluid) x {FRAMES} frames, "
         f"{ITERS} iterations in {t:.1f} s: loss {np.mean(losses[:50]):.3g} -> {np.mean(losses[-200:]):.3g}")
    return S, cfg, data, net, losses


def _scene():
    sc = scenes.dam_break(width=0.8, height=2.0, dx=0.005)
    assert sc.n_fluid == 64000
    return sc


def test_training_converges(trained):
    S, cfg, data, net, losses = trained
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-200:]) < 1e-2 * np.mean(losses[:50])


@pytest.mark.parametrize("rule", ["mean_drift", "mean_drift_and_particle"])
def test_rollout_64k_with_corrections_and_fallback(trained, rule):
    from uph_amd.solver import GpuSolver
    S, cfg, data, net, _ = trained
    c = dataclasses.replace(cfg, spike_tol=None if rule == "mean_drift" else 0.05)
    sc = _scene()
    ro = S.Rollout(sc, net, data.vscale, data.ascale, c)
    ro.warmup()
    t0 = time.perf_counter()
    ro.run(120)
    t_roll = time.perf_counter() - t0
    steps = ro.macro * c.substeps
    ref = GpuSolver(sc)
    ref.step(steps)
    ref.sync()
    t1 = time.perf_counter()
    ref2 = GpuSolver(sc)
    ref2.step((ro.macro - S.HIST) * c.substeps)
    ref2.sync()
    t_plain = time.perf_counter() - t1
    err = S.errors(ro.state(), ref.state(), sc)
    nf = sc.n_fluid
    span = (ro.macro - S.HIST) * c.substeps
    thr, thr_plain = nf * span / t_roll, nf * span / t_plain
    net_ms = 1e3 * ro.t_net / max(ro.blocks * c.every, 1)
    _say(f"{rule}: {ro.blocks} blocks, {ro.fallbacks} fallen back ({ro.escapes} left the tank); kept-block "
         f"drifts max {max(ro.kept_drifts, default=0):.4f}; network macro step {net_ms:.2f} ms (20 plain "
         f"steps {1e3 * t_plain / span * 20:.2f} ms); plain-dt-equivalent {thr:.3g} updates/s = "
         f"{thr / thr_plain:.3f} x plain; errors vs plain {err}")
    st = ro.state()
    assert np.all(np.isfinite(st["x"])) and np.all(np.isfinite(st["rho"]))
    assert all(d <= c.drift_tol for d in ro.kept_drifts)
    assert ro.blocks >= 10


def test_ablation_without_fallback(trained):
    """Fallback disabled (every block kept): the trained network against the zero-output
    network under the same corrections, three blocks of 10 macro steps."""
    from uph_amd.solver import GpuSolver
    S, cfg, data, net, _ = trained
    c = dataclasses.replace(cfg, drift_tol=float("inf"), spike_tol=None)
    sc = _scene()
    res = {}
    for name, nn in (("trained", net),
                     ("zero", type(net)(cfg.width, cfg.mp_steps,
                                        params={k: 0 * v for k, v in net.params().items()}))):
        ro = S.Rollout(sc, nn, data.vscale, data.ascale, c)
        ro.warmup()
        for _ in range(3):
            ro.block()
        ref = GpuSolver(sc)
        ref.step(ro.macro * c.substeps)
        res[name] = S.errors(ro.state(), ref.state(), sc)
        _say(f"no fallback, {name}: {ro.blocks} blocks ({ro.fallbacks} fallbacks, {ro.escapes} left the tank): "
             f"{res[name]}")
    assert res["trained"]["pos_err_mean_dx"] < res["zero"]["pos_err_mean_dx"]
