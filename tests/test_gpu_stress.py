"""stressType 2 of the reference's run-phase scheme on the GPU: ClassicShearStress
(derivatives.c:1686-1721) by k_stress before the pair pass, and ViscousForcesTwoStep
(derivatives.c:1724-1768) in dvdt, against the FP64 oracle.

Rates are compared to 1e-5 of their largest magnitude, and the viscous part alone
(the acceleration at viscosity mu minus the one at a vanishing viscosity, on both
sides) to 1e-5 of its own; trajectories to 1e-6 relative density after the stated
steps.  The compacted and the direct pair kernel stay bit-identical; the gather form
conserves momentum.
"""
import os
import subprocess

import numpy as np
import pytest

from uph_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    print(f"[gpu-stress] {msg}", flush=True)


def _two(sc, mu=0.5, kernel=scenes.KERNEL_CUBIC):
    sc.mode, sc.stress_type, sc.kernel = scenes.MODE_REF, 2, kernel
    sc.kappa, sc.mu, sc.eps, sc.alpha_av = 0.015, mu, 1.0, 0.0
    return sc


def _moving(sc, seed=1, amp=0.3):
    rng = np.random.default_rng(seed)
    nf = sc.n_fluid
    sc.vel[:nf] = rng.normal(0, amp, (nf, 2))
    sc.lvel[:nf] = sc.vel[:nf] + rng.normal(0, 0.1 * amp, (nf, 2))
    sc.pressure[:nf] = np.abs(rng.normal(500.0, 200.0, nf))
    return sc


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


def _rates(sc, steps):
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    g = GpuSolver(sc, keep_rates=True)
    assert g.resorts() == -1          # a stressType 2 handle sorts every step
    g.step(steps)
    assert g.sync() == 0
    rg = g.rates()
    rg.update(g.ref_rates())
    o = Sim(sc, nthreads=0)
    o.step(steps)
    ro = o.rates()
    ro.update(o.ref())
    return rg, ro


@pytest.mark.parametrize("kernel", [scenes.KERNEL_CUBIC, scenes.KERNEL_POLY7])
def test_stress_two_rates_and_viscous_part(kernel):
    """Steps 1 and 2 of a jittered dam break with random velocities and pressures; the
    viscous part at step 1, where both runs start from the same state."""
    nf = None
    for steps in (1, 2):
        sc = _moving(_two(scenes.dam_break(jitter=0.1, seed=4), kernel=kernel))
        nf = sc.n_fluid
        rg, ro = _rates(sc, steps)
        err = {k: _rel(rg[k], ro[k][:nf]) for k in ("R", "Rt", "a", "at", "RP", "RPt")}
        _say(f"kernel {kernel} step {steps}: rel-to-max errors {err}")
        for k, e in err.items():
            assert e < 1e-5, (k, e)
        if steps == 2:
            continue   # after a step the two runs' states differ, and so do their other terms
        # a vanishing viscosity, not zero: the reference's effective viscosity of the
        # predicting form is a harmonic mean, 0/0 at viscosity0 = 0
        sc0 = _moving(_two(scenes.dam_break(jitter=0.1, seed=4), mu=1e-12, kernel=kernel))
        rg0, ro0 = _rates(sc0, steps)
        vg = rg["a"] - rg0["a"]
        vo = ro["a"][:nf] - ro0["a"][:nf]
        ev = _rel(vg, vo)
        _say(f"kernel {kernel} step {steps}: viscous part max {np.abs(vo).max():.3e}, rel-to-max error {ev:.2e}")
        assert np.abs(vo).max() > 1e-3 * np.abs(ro["a"][:nf]).max()
        assert ev < 1e-5


def test_stress_two_dam_break_300_steps():
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    sc = _two(scenes.dam_break(), mu=0.05)
    g = GpuSolver(sc)
    g.step(300)
    assert g.sync() == 0
    o = Sim(sc, nthreads=0)
    o.step(300)
    sg, so = g.state(), o.state()
    nf = sc.n_fluid
    drho = float(np.max(np.abs(sg["rho"][:nf] - so["rho"][:nf]) / so["rho"][:nf]))
    dP = _rel(g.pressure()[:nf], o.ref()["P"][:nf])
    dx = float(np.abs(sg["x"][:nf] - so["x"][:nf]).max()) / sc.dx
    _say(f"300 steps: rho {drho:.2e}, P rel-to-max {dP:.2e}, |dx|/dx {dx:.2e}")
    assert drho < 1e-6 and dP < 1e-5 and dx < 1e-4


def test_stress_two_compact_equals_direct_and_batch_equals_single():
    from uph_amd.solver import GpuSolver
    sc = _moving(_two(scenes.dam_break(width=0.5, height=0.8, jitter=0.1, seed=12)))
    a, b = GpuSolver(sc, pair_variant=0), GpuSolver(sc, pair_variant=1)
    a.step(20)
    b.step(20)
    for key in ("x", "v", "rho"):
        assert np.array_equal(a.state()[key], b.state()[key]), key
    assert np.array_equal(a.pressure(), b.pressure())
    other = _two(scenes.dam_break(width=0.3, height=0.4, dx=0.015))
    gb = GpuSolver([sc, other])
    gb.step(20)
    for key in ("x", "v", "rho"):
        assert np.array_equal(gb.state(0)[key], a.state()[key]), key


def test_stress_two_conserves_momentum():
    """Fluid only, random v and P: the pair terms of the gather form cancel pairwise."""
    from uph_amd.solver import GpuSolver
    n = 30
    dx = 0.01
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    rng = np.random.default_rng(2)
    pos = np.stack([(i.ravel() + 0.5) * dx, (j.ravel() + 0.5) * dx], 1) + rng.uniform(-2e-3, 2e-3, (n * n, 2))
    N = len(pos)
    c0, B = scenes.tait_constants(1000.0, 7.0, 9.81, 1.0)
    sc = scenes.Scene(pos=pos, vel=rng.normal(0, 0.5, (N, 2)), rho=np.full(N, 1000.0),
                      pressure=rng.uniform(0, 300, N), mass=np.full(N, 0.1), itype=np.ones(N, np.int32),
                      lvel=np.zeros((N, 2)), n_fluid=N, n_fixed=0, n_moving=0, dx=dx, h=1.3 * dx, c0=c0,
                      B=B, gravity=0, dt=0.25 * 1.3 * dx / c0, box=(-0.05, -0.05, 0.35, 0.35))
    sc.lvel[:] = sc.vel
    _two(sc)
    g = GpuSolver(sc, keep_rates=True)
    g.step(1)
    r = g.rates()
    ma = 0.1 * r["a"]
    _say(f"momentum: |sum m a| {np.abs(ma.sum(0)).max():.2e} of sum |m a| {np.abs(ma).sum():.2e}")
    assert np.abs(ma.sum(0)).max() < 1e-6 * np.abs(ma).sum()


def test_stress_two_limits():
    from uph_amd.solver import GpuSolver
    capped = _two(scenes.dam_break(width=0.2, height=0.2, h_factor=2.691))
    with pytest.raises(RuntimeError, match="stressType 2"):
        GpuSolver(capped)
    a = _two(scenes.dam_break(width=0.2, height=0.2))
    b = _two(scenes.dam_break(width=0.2, height=0.3))
    b.stress_type = 1
    with pytest.raises(RuntimeError, match="stressType"):
        GpuSolver([a, b])
    g = GpuSolver(a)
    with pytest.raises(RuntimeError, match="stressType 2"):
        g.step_persistent(2)
    # a patch compressed until more than K lie in range: the stresses would take every
    # neighbour in range, so the step reports it
    sc = _two(scenes.dam_break(width=0.5, height=0.6, dx=0.02))
    sc.K = 29
    nf = sc.n_fluid
    x = sc.pos[:nf]
    patch = (x[:, 0] < 0.16) & (x[:, 1] < 0.16)
    sc.pos[:nf][patch] = x[patch] * 0.7
    g = GpuSolver(sc)
    g.step(2)
    with pytest.raises(RuntimeError, match="stressType 2"):
        g.sync()


def test_c_host_driver_stress_two(tmp_path):
    """A deck with stressType 2 in its parameters.dat: uph_run's device loop and its
    --compat loop over the derivatives() seam agree, and match the oracle."""
    from oracle.oracle import Sim
    from uph_amd.scenes import load_deck_dir, write_deck_dir
    drv = os.path.join(ROOT, "uph_amd", "uph_run")
    assert os.path.exists(drv), "uph_run not built"
    sc = _two(scenes.dam_break(width=0.3, height=0.3, dx=0.01, tank=(0.6, 0.5)), mu=0.05)
    write_deck_dir(str(tmp_path), sc)
    steps = 40
    args = [drv, str(steps), "--scheme", "ref", "--no-output"]
    for extra, name in (([], "device.dat"), (["--compat"], "compat.dat")):
        out = subprocess.run(args + extra + ["--out", name], cwd=tmp_path, capture_output=True, text=True,
                             timeout=300)
        assert out.returncode == 0, out.stderr
    a = np.loadtxt(tmp_path / "device.dat", skiprows=1)
    b = np.loadtxt(tmp_path / "compat.dat", skiprows=1)
    dk = load_deck_dir(str(tmp_path), mode=scenes.MODE_REF)
    assert dk.stress_type == 2
    drho = float(np.max(np.abs(a[:, 4] - b[:, 4]) / b[:, 4]))
    _say(f"uph_run stressType 2: device vs compat rho {drho:.2e}")
    assert drho < 1e-8
    o = Sim(dk, nthreads=0)
    o.step(steps)
    so = o.state()
    n = dk.n_fluid
    drho_o = float(np.max(np.abs(a[:, 4] - so["rho"][:n]) / so["rho"][:n]))
    _say(f"uph_run stressType 2 vs oracle: rho {drho_o:.2e}")
    assert drho_o < 1e-6
