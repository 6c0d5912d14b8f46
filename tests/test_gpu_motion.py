"""Moving-boundary kinematics on the GPU (inputbp.c:192-283) against the FP64 oracle:
the lid-driven cavity, the dam break released by a gate, the sloshing tank (also with
a moving row of another itype, past which the reference's loop stops the left wall).

Fluid states to the stepping tolerance of the other GPU tests (density 1e-5 relative,
positions 1e-3 dx), or, where walls sweep through neighbour sets that K = 40 caps
(the sloshing tank and its mirror rows), to four times the oracle's own spread
under perturbations of the fluid positions (1e-12 to 1e-8 dx): a neighbour that enters the
capped set one step earlier changes the rates by a kernel weight, so there the
trajectories are only as close as the problem's conditioning.  Boundary rows: the state after n steps has them where the n-th
derivatives() call put them, at n dt, in the oracle and through uph_get_state and
uph_state_to_device; the kinematics are FP64 on both, so they agree to rounding.
"""
import numpy as np
import pytest

from conftest import oracle_spread
from uph_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-motion] {msg}", flush=True)


def _run(sc, steps, name, conditioned=False):
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    _say(f"{name}: n={sc.n} steps={steps} motion={sc.motion}")
    g = GpuSolver(sc)
    g.step(steps)
    assert g.sync() >= 0
    sg = g.state()
    o = Sim(sc, nthreads=0)
    o.step(steps)
    so = walls = o.state()
    nf = sc.n_fluid
    drho = float(np.max(np.abs(sg["rho"][:nf] - so["rho"][:nf]) / so["rho"][:nf]))
    dx = float(np.abs(sg["x"][:nf] - so["x"][:nf]).max()) / sc.dx
    bx = float(np.abs(sg["x"][nf:] - walls["x"][nf:]).max())
    bv = float(np.abs(sg["v"][nf:] - walls["v"][nf:]).max())
    _say(f"{name}: fluid rho {drho:.2e}, |dx|/dx {dx:.2e}; walls |dx| {bx:.2e} m, |dv| {bv:.2e} m/s")
    tr, tx = 1e-5, 1e-3
    if conditioned:
        sr, sx = oracle_spread(sc, steps, so)
        _say(f"{name}: oracle spread under 1e-12..1e-8 dx perturbations: rho {sr:.2e}, |dx|/dx {sx:.2e}")
        tr, tx = max(tr, 4 * sr), max(tx, 4 * sx)
    assert drho < tr and dx < tx
    assert bx < 1e-12 and bv < 1e-12
    # the device-side getter places them the same way
    from uph_amd.devbuf import DevBuf
    pos, vel, rho = DevBuf(16 * sc.n), DevBuf(16 * sc.n), DevBuf(8 * sc.n)
    g.state_to_device(pos.ptr, vel.ptr, rho.ptr)
    dpos, dvel = pos.download((sc.n, 2), np.float64), vel.download((sc.n, 2), np.float64)
    assert np.array_equal(dpos[:nf], sg["x"][:nf])
    assert float(np.abs(dpos[nf:] - sg["x"][nf:]).max()) < 1e-13
    assert float(np.abs(dvel[nf:] - sg["v"][nf:]).max()) < 1e-13
    return sg, so


def test_lid_driven_cavity():
    sc = scenes.lid_cavity(size=0.2, dx=0.005, lid_u=1.0)
    sg, _ = _run(sc, 300, "lid cavity 0.2 m, dx 5 mm, 300 steps")
    mv = slice(sc.n_fluid + sc.n_fixed, sc.n)
    assert np.allclose(sg["x"][mv, 0] - sc.pos[mv, 0], 300 * sc.dt, rtol=1e-12)


def test_gate_dam_break_through_its_stop():
    sc = scenes.gate_dam_break(width=0.3, height=0.6, dx=0.01, gate_u=2.0)
    # stop the gate after 100.5 steps so that 200 steps cover both phases
    sc.depth = 15 * sc.dx + 2.0 * 100.5 * sc.dt / 1.5
    sg, _ = _run(sc, 200, "gate dam break 0.3 x 0.6, gate stops at step 100, 200 steps")
    mv = slice(sc.n_fluid + sc.n_fixed, sc.n)
    assert np.allclose(sg["x"][mv, 1] - sc.gate_y[mv], 2.0 * 100.5 * sc.dt, rtol=1e-12)
    assert np.all(sg["v"][mv, 1] == 0.0)


def test_sloshing_tank():
    """120 steps (the walls travel half a lattice spacing) before any neighbour of the
    capped sets near the walls changes: the fixed tolerance; 300 steps, through such
    changes: the oracle's own spread."""
    sc = scenes.sloshing_tank(width=0.2, depth=0.06, dx=0.004, wall_height=0.12, stroke=0.01, period=0.1)
    _run(sc, 120, "sloshing tank 0.2 x 0.06, dx 4 mm, 120 steps")
    _run(sc, 300, "sloshing tank 0.2 x 0.06, dx 4 mm, 300 steps", conditioned=True)


def test_sloshing_tank_with_a_row_of_another_itype():
    sc = scenes.sloshing_tank(width=0.2, depth=0.06, dx=0.004, wall_height=0.12, stroke=0.01, period=0.1)
    it = sc.itype
    first = sc.n_fluid + sc.n_fixed + int((it[sc.n_fluid + sc.n_fixed:] == 1).sum())
    sc.itype = it.copy()
    sc.itype[first:first + 3] = 4
    sg, _ = _run(sc, 100, "sloshing tank, itype-4 rows after the bottom, 100 steps", conditioned=True)
    left = np.nonzero(sc.itype == 2)[0]
    assert np.all(sg["x"][left, 0] == 0.0) and np.all(sg["v"][left, 0] == 0.0)


def test_lid_driven_cavity_reference_scheme():
    """The reference's run-phase scheme with the lid: its localVelocity is the lid
    velocity (inputbp.c:221), which the velocity fluctuation and the viscous term see."""
    from oracle.oracle import Sim
    from uph_amd.solver import GpuSolver
    sc = scenes.lid_cavity(size=0.1, dx=0.0025, lid_u=1.0)
    sc.mode = scenes.MODE_REF
    sc.kappa, sc.mu, sc.eps, sc.alpha_av = 0.001, 1e-3, 1.0, 0.0
    steps = 200
    _say(f"ref mode lid cavity: n={sc.n} steps={steps}")
    g = GpuSolver(sc)
    g.step(steps)
    assert g.sync() >= 0
    sg, Pg = g.state(), g.pressure()
    o = Sim(sc, nthreads=0)
    o.step(steps)
    so, Po = o.state(), o.ref()["P"]
    nf = sc.n_fluid
    drho = float(np.max(np.abs(sg["rho"][:nf] - so["rho"][:nf]) / so["rho"][:nf]))
    dP = float(np.abs(Pg[:nf] - Po[:nf]).max() / np.abs(Po[:nf]).max())
    dx = float(np.abs(sg["x"][:nf] - so["x"][:nf]).max()) / sc.dx
    _say(f"ref mode lid cavity: rho {drho:.2e}, P rel-to-max {dP:.2e}, |dx|/dx {dx:.2e}")
    assert drho < 1e-6 and dP < 1e-5 and dx < 1e-4
