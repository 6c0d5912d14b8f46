"""uph_diagnostics (device reductions) against the FP64 statement oracle/diag_ref.py."""
import numpy as np
import pytest

from conftest import load_golden_scene
from oracle import diag_ref
from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-diag] {msg}", flush=True)


def _expect(sc, st, probe):
    nf = sc.n_fluid
    x, v, rho = st["x"][:nf], st["v"][:nf], st["rho"][:nf]
    m = sc.mass[:nf]
    P = sc.B * ((rho / sc.rho0) ** sc.gamma - 1.0)
    ke, pe = diag_ref.total_energy(x, v, m, sc.g)
    p, norm, cnt = diag_ref.pressure_probe(x, P, m, rho, sc.h, *probe)
    one, two, three, unc = diag_ref.density_consistency(rho, rho)
    return dict(kinetic_energy=ke, potential_energy=pe, probe_pressure=p if p is not None else 0.0,
                probe_norm=norm, probe_neighbors=cnt, consistency_one=one, consistency_two=two,
                consistency_three=three, density_uncertainty=unc)


def _close(a, b, rel=1e-11):
    for k in b:
        assert np.isclose(a[k], b[k], rtol=rel, atol=1e-300), (k, a[k], b[k])


def test_diagnostics_match_the_statement():
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(jitter=0.1, seed=4)
    rng = np.random.default_rng(5)
    sc.vel[:sc.n_fluid] = rng.normal(0, 0.5, (sc.n_fluid, 2))
    probe = (0.5, 0.3)
    g = GpuSolver(sc)
    d0 = g.diagnostics(probe)
    e0 = _expect(sc, dict(x=sc.pos, v=sc.vel, rho=sc.rho), probe)
    _say(f"step 0: {d0}")
    assert d0["probe_neighbors"] > 50
    _close(d0, e0)
    g.step(50)
    st = g.state()
    d1 = g.diagnostics(probe)
    _close(d1, _expect(sc, st, probe))
    # a probe with no particle within 3h
    d2 = g.diagnostics((100.0, 100.0))
    assert d2["probe_neighbors"] == 0 and d2["probe_norm"] == 1.0 and d2["probe_pressure"] == 0.0


def test_diagnostics_of_a_batch_scene():
    from uph_amd.solver import GpuSolver
    scs = [scenes.dam_break(width=0.4, height=0.6, dx=0.02), scenes.dam_break(dx=0.02)]
    g = GpuSolver(scs)
    g.step(10)
    for k, sc in enumerate(scs):
        _close(g.diagnostics((0.2, 0.1), scene=k), _expect(sc, g.state(k), (0.2, 0.1)))


def _moving(sc, seed=2, amp=0.3):
    rng = np.random.default_rng(seed)
    nf = sc.n_fluid
    sc.vel[:nf] = rng.normal(0, amp, (nf, 2))
    sc.lvel[:nf] = sc.vel[:nf] + rng.normal(0, 0.1 * amp, (nf, 2))
    return sc


def _vort_case(case):
    if case == "cubic":
        return _moving(scenes.dam_break(jitter=0.1, seed=4))
    if case == "poly7":
        sc = _moving(scenes.dam_break(jitter=0.1, seed=6))
        sc.kernel = scenes.KERNEL_POLY7
        return sc
    if case == "capped":
        return _moving(load_golden_scene("dambreak.npz"), amp=0.1)
    sc = load_golden_scene("dambreak.npz", mode=scenes.MODE_REF, kernel=scenes.KERNEL_POLY7, alpha_av=0.0)
    return _moving(sc, amp=0.1)


@pytest.mark.parametrize("case", ["cubic", "poly7", "capped", "ref_poly7_capped"])
def test_vorticity_matches_the_statement(case):
    """approxVorticity of the first step's evaluation (the initial state: the first step
    predicts nothing) against the FP64 statement over the oracle's pair list (K cap
    included on the reference's dambreak deck).  The device sums FP64 pair terms of
    the FP32 neighbour mirror; tolerance 1e-5 of the largest magnitude."""
    from oracle import oracle as O
    from uph_amd.solver import GpuSolver
    sc = _vort_case(case)
    g = GpuSolver(sc)
    g.step(1)
    got = g.vorticity()
    ip, jp, _ = O.search(sc.pos, sc.n_fluid, sc.n, sc.particle_boundary, sc.K, sc.cutoff2)
    exp = diag_ref.approx_vorticity(sc.pos, sc.vel, sc.rho, sc.mass, sc.h, ip, jp, sc.n_fluid,
                                    1 if sc.kernel == scenes.KERNEL_POLY7 else 0)
    err = float(np.abs(got - exp).max() / np.abs(exp).max())
    _say(f"vorticity {case}: n={sc.n}, max |w| {np.abs(exp).max():.3e}, rel-to-max error {err:.2e}")
    assert err < 1e-5


def test_probe_two_matches_the_statement():
    from uph_amd.solver import GpuSolver
    sc = scenes.dam_break(jitter=0.05, seed=8)
    g = GpuSolver(sc)
    g.step(30)
    st = g.state()
    nf = sc.n_fluid
    P = sc.B * ((st["rho"][:nf] / sc.rho0) ** sc.gamma - 1.0)
    for loc in ((0.5, 0.3), (0.2, 1.0), (50.0, 50.0)):
        p, n = g.probe_two(*loc)
        pe, ne = diag_ref.probe_two(st["x"][:nf], P, sc.h, *loc)
        _say(f"probe two at {loc}: {p:.6f} ({n}) vs {pe:.6f} ({ne})")
        assert n == ne
        assert np.isclose(p, pe, rtol=1e-11, atol=0)
