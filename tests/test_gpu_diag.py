"""uph_diagnostics (device reductions) against the FP64 statement oracle/diag_ref.py."""
import numpy as np
import pytest

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
