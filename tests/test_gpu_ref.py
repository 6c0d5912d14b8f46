"""The reference's run-phase scheme (UPH_MODE_REF) on the GPU against the FP64 oracle.

Pair sums run in FP32 with compensated accumulation on the GPU, FP64 in the oracle;
density, DynamicPressure, velocity and position are integrated in FP64 on both.
Tolerances: rates to 1e-5 of their largest magnitude (fluctuations, whose sums cancel
to a few per mille of their terms, to 1e-4), states to 1e-6 relative after the stated
number of steps.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden_scene
from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-ref] {msg}", flush=True)


def _ref(sc, mu=1e-3):
    sc.mode = scenes.MODE_REF
    sc.kappa, sc.mu, sc.eps = 0.015, mu, 1.0
    sc.alpha_av = 0.0
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


def _step_rates(sc, name, steps=1):
    """Rates of the steps-th derivatives() call on the GPU and in the oracle."""
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    _say(f"rates {name}: n={sc.n}, step {steps}")
    g = GpuSolver(sc, keep_rates=True)
    g.step(steps)
    g.sync()
    rg = g.rates()
    rg.update(g.ref_rates())
    o = Sim(sc, nthreads=0)
    o.step(steps)
    ro = o.rates()
    ro.update(o.ref())
    nf = sc.n_fluid
    keys = ("R", "Rt", "a", "at", "RP", "RPt", "drho", "dP", "dv")
    err = {k: _rel(rg[k], ro[k][:nf]) for k in keys if np.abs(ro[k][:nf]).max() > 0}
    _say(f"rates {name}: rel-to-max errors {err}")
    return err, g, rg, ro


def _check(err):
    for k, e in err.items():
        assert e < (1e-4 if k in ("drho", "dP", "dv") else 1e-5), (k, e)


def _check_steps_1_and_2(sc, name):
    """Step 1 starts from localDensity = rho0 everywhere (input.c:109): with equal
    masses the velocity fluctuation's coefficient m_p/rholoc_q - m_q/rholoc_p is 0 in
    the oracle and at FP32 rounding on the GPU.  Step 2's local fields carry the
    first step's fluctuations, so all nine fields are compared there."""
    err, g, rg, ro = _step_rates(sc, name, 1)
    _check(err)
    if "dv" not in err:
        assert np.abs(rg["dv"]).max() < 1e-6
    err, g, rg, ro = _step_rates(sc, name, 2)
    assert "dv" in err
    _check(err)
    return g


def test_ref_rates_jittered_moving_fluid():
    sc = _moving(_ref(scenes.dam_break(jitter=0.1, seed=4)))
    _check_steps_1_and_2(sc, "ref mode, dam break 5K jittered, random v, P")


def test_ref_rates_poly7_reference_dambreak_deck():
    """The reference's own scheme on its own deck: PolyOrderSeven at h = 2.691 dx,
    K = 40 binding for every particle, its kappa, viscosity, epsilon and springs."""
    sc = load_golden_scene("dambreak.npz", mode=scenes.MODE_REF, kernel=scenes.KERNEL_POLY7, alpha_av=0.0)
    sc = _moving(sc, amp=0.1)
    g = _check_steps_1_and_2(sc, "ref mode, PolyOrderSeven, tests/dambreak")
    assert g.sync() == 2


def _run_both(sc, steps, name, **gopts):
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    _say(f"{name}: n={sc.n} steps={steps}")
    g = GpuSolver(sc, **gopts)
    g.step(steps)
    g.sync()
    sg, Pg = g.state(), g.pressure()
    o = Sim(sc, nthreads=0)
    o.step(steps)
    so, Po = o.state(), o.ref()["P"]
    nf = sc.n_fluid
    drho = float(np.max(np.abs(sg["rho"][:nf] - so["rho"][:nf]) / so["rho"][:nf]))
    dP = _rel(Pg[:nf], Po[:nf])
    dx = float(np.abs(sg["x"][:nf] - so["x"][:nf]).max()) / sc.dx
    _say(f"{name}: max rel rho {drho:.3e}, P rel-to-max {dP:.3e}, max |dx|/dx {dx:.3e}")
    return drho, dP, dx, g


def test_ref_dam_break_5k_300_steps():
    sc = _ref(scenes.dam_break())
    drho, dP, dx, _ = _run_both(sc, 300, "ref mode dam break 5K 300 steps")
    assert drho < 1e-6 and dP < 1e-5 and dx < 1e-4


@pytest.mark.parametrize("fixture", ["hydrostatic_init.npz", "waves.npz"])
def test_ref_reference_decks_with_their_kernel(fixture):
    """The unmodified reference scheme as configured by its own decks: PolyOrderSeven,
    hsml = eta dx, DynamicPressure from the deck, spring walls, the waves deck's
    piston; 100 steps."""
    sc = load_golden_scene(fixture, mode=scenes.MODE_REF, kernel=scenes.KERNEL_POLY7, alpha_av=0.0)
    drho, dP, dx, _ = _run_both(sc, 100, f"ref mode PolyOrderSeven {fixture} 100 steps")
    assert drho < 1e-6 and dP < 1e-5 and dx < 1e-4


def test_ref_compact_variant_equals_direct():
    from uph_amd.solver import GpuSolver
    sc = _moving(_ref(scenes.dam_break(width=0.5, height=0.8, jitter=0.1, seed=12)))
    _say("ref mode compact vs direct")
    a = GpuSolver(sc, pair_variant=0)
    b = GpuSolver(sc, pair_variant=1)
    a.step(20)
    b.step(20)
    sa, sb = a.state(), b.state()
    for key in ("x", "v", "rho"):
        assert np.array_equal(sa[key], sb[key]), key
    assert np.array_equal(a.pressure(), b.pressure())


def test_mode_must_be_uniform_across_a_batch():
    from uph_amd.solver import GpuSolver
    a = scenes.dam_break(width=0.2, height=0.2)
    b = _ref(scenes.dam_break(width=0.2, height=0.3))
    with pytest.raises(RuntimeError, match="scheme"):
        GpuSolver([a, b])
    with pytest.raises(RuntimeError):
        GpuSolver(a, keep_rates=True).ref_rates()


def test_ref_hydrostatic_deck_against_the_reference_recorded_run():
    """The reference's own recorded run of tests/hydrostaticPressure: started from its
    lattice deck (hydrostatic_start.npz), it appended the probe (0.98, 0.06) pressure and
    the kinetic energy every 5000 steps (PointPressure.dat, TotalKineticEnergy.dat:
    hydrostatic_probe.npz) and its fluid state at step 30000 (FluidParticleInit.dat:
    hydrostatic_init.npz).  The same deck, scheme and kernel here (reference scheme,
    PolyOrderSeven, hsml = eta dx, its springs) for 30000 steps, sampled at the same
    steps.  The reference's run is not reproducible bit for bit (OpenMP reductions in
    thread order), and its probe reads the state inside derivatives() rather than after
    the step, so the comparison is loose."""
    from uph_amd.solver import GpuSolver
    sc = load_golden_scene("hydrostatic_start.npz", mode=scenes.MODE_REF, kernel=scenes.KERNEL_POLY7,
                           alpha_av=0.0)
    z = np.load(os.path.join(GOLDEN, "hydrostatic_probe.npz"), allow_pickle=False)
    rec_p, rec_ke = z["point_pressure"], z["kinetic_energy"]
    probe = (0.98, 0.06)
    g = GpuSolver(sc, use_graph=True)
    got_p, got_ke = [], []
    for k in range(6):
        g.step(5000)
        d = g.diagnostics(probe)
        got_p.append(d["probe_pressure"])
        got_ke.append(d["kinetic_energy"])
    assert g.sync() == 0
    st = g.state()
    ref = load_golden_scene("hydrostatic_init.npz")
    nf = sc.n_fluid
    drho = np.abs(st["rho"][:nf] - ref.rho[:nf]) / ref.rho[:nf]
    dpos = np.linalg.norm(st["x"][:nf] - ref.pos[:nf], axis=1) / sc.dx
    move = np.linalg.norm(ref.pos[:nf] - sc.pos[:nf], axis=1) / sc.dx
    rp = np.asarray(got_p) / rec_p[:6, 1]
    rk = np.asarray(got_ke) / rec_ke[:6, 1]
    _say(f"recorded run, t = 0.05..0.3 s: probe P here {np.round(got_p, 1).tolist()}, recorded "
         f"{np.round(rec_p[:6, 1], 1).tolist()} (ratio {np.round(rp, 4).tolist()}); KE here "
         f"{np.round(got_ke, 5).tolist()}, recorded {np.round(rec_ke[:6, 1], 5).tolist()} "
         f"(ratio {np.round(rk, 3).tolist()})")
    _say(f"state at step 30000 vs the reference's FluidParticleInit.dat: density rel diff mean "
         f"{drho.mean():.2e} max {drho.max():.2e}; position diff mean {dpos.mean():.2e} dx, max "
         f"{dpos.max():.2e} dx (the reference's particles moved {move.mean():.3f} dx on average, max "
         f"{move.max():.3f} dx)")
    # the recorded values carry 10 decimals (positions 1e-10 m = 1e-8 dx)
    assert np.all(np.abs(rp - 1.0) < 1e-3)
    assert np.all(np.abs(rk - 1.0) < 2e-3)
    assert drho.max() < 1e-6
    assert dpos.max() < 1e-3

