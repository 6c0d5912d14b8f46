"""Pins the oracle's stressType 2 path of the reference's run-phase scheme:
ClassicShearStress (derivatives.c:1686-1721) and ViscousForcesTwoStep
(derivatives.c:1724-1768), selected at derivatives.c:4231-4244.

What the mathematics fixes:

* the stress of a linear shear flow v = (c y, 0) with localVelocity = v is
  viscosity0 (grad v + grad v^T), so xy = yx = viscosity0 c and xx = yy = 0 on a
  symmetric lattice; with localVelocity = 0 the same flow gives zero stress inside
  (the pair term takes Velocity of one particle and localVelocity of the other);
* the two-step force of a parabolic flow v = (c y^2, 0) is div(stress) / rho =
  2 viscosity0 c / rho along x, 0 along y (the opposite sign and a factor rho apart
  from the single-step predicting form pinned in test_oracle_ref.py, as in the
  reference);
* the force is linear in viscosity0, and the predicting accelerations do not depend
  on stressType;
* both functions against an independent per-particle gather statement over a pair
  set found by numpy, walls included (a wall's stress enters its fluid neighbours'
  force).
"""
import math

import numpy as np
import pytest

from test_oracle_ref import KERNELS, _eval, _interior, _patch, _pressure_accel
from uph_amd import scenes


def _stress(sc, vloc=None):
    from oracle.oracle import Sim
    s = Sim(sc)
    s.set_local(rholoc=sc.rho, Ploc=sc.pressure, vloc=sc.vel if vloc is None else vloc)
    s.derivatives(sc.dt)
    return s.stress()


@pytest.mark.parametrize("kernel", KERNELS)
def test_shear_stress_of_a_linear_shear_flow(oracle_lib, kernel):
    sc = _patch(kernel=kernel)
    sc.mu, sc.stress_type = 1e-3, 2
    yc = sc.pos[:, 1].mean()
    c = 2.5
    sc.vel = np.stack([c * (sc.pos[:, 1] - yc), np.zeros(sc.n)], 1)
    S = _stress(sc)
    m = _interior(sc, 3 * sc.h)
    assert np.allclose(S[m, 1], sc.mu * c, rtol=2e-2)
    assert np.array_equal(S[:, 1], S[:, 2])
    assert np.abs(S[m][:, [0, 3]]).max() < 1e-9 * sc.mu * c
    # Velocity of the particle against localVelocity of its neighbours: with the local
    # field zero, the sum is v_p times the kernel gradient's zero first moment
    S0 = _stress(sc, vloc=np.zeros_like(sc.vel))
    L = sc.pos[:, 1].max() - yc
    assert np.abs(S0[m]).max() < 1e-9 * sc.mu * c * L / sc.h


@pytest.mark.parametrize("kernel", KERNELS)
def test_two_step_force_of_a_parabolic_flow(oracle_lib, kernel):
    """P = 0, no gravity, no walls: the acceleration is the two-step viscous force only."""
    sc = _patch(kernel=kernel, n=36 if kernel == scenes.KERNEL_CUBIC else 44)
    sc.mu, sc.stress_type = 1e-3, 2
    yc = sc.pos[:, 1].mean()
    c = 30.0
    sc.vel = np.stack([c * (sc.pos[:, 1] - yc) ** 2, np.zeros(sc.n)], 1)
    r = _eval(sc)
    m = _interior(sc, 5 * sc.h)
    assert m.sum() > 20
    want = 2 * sc.mu * c / sc.rho0
    assert np.allclose(r["a"][m, 0], want, rtol=5e-2)
    assert np.abs(r["a"][m, 1]).max() < 1e-6 * want


def test_two_step_force_is_linear_in_viscosity_and_leaves_the_predicting_form(oracle_lib):
    sc = _patch(jitter=0.2)
    rng = np.random.default_rng(4)
    sc.vel = rng.normal(0, 0.5, sc.vel.shape)
    sc.rho = 1000.0 * (1 + rng.uniform(0, 0.01, sc.n))
    sc.pressure = rng.uniform(0, 300, sc.n)
    sc.stress_type = 2
    out = {}
    for mu in (0.0, 1e-3, 2e-3):
        sc.mu = mu
        out[mu] = _eval(sc)
    d1 = out[1e-3]["a"] - out[0.0]["a"]
    d2 = out[2e-3]["a"] - out[0.0]["a"]
    assert np.abs(d1).max() > 0
    assert np.abs(d2 - 2 * d1).max() < 1e-9 * np.abs(d2).max()
    sc.mu = 1e-3
    sc.stress_type = 1
    one = _eval(sc)
    assert np.array_equal(one["at"], out[1e-3]["at"])
    assert np.abs(one["a"] - out[1e-3]["a"]).max() > 1e-6 * np.abs(d1).max()


def _gather_stress(sc, st, vloc, rholoc, ip, jp, r2, kernel):
    """ClassicShearStress per particle p over its pair partners q, written in gather form:
    -mu (m_q rholoc_q / (rho_q rho_p)) (D G^T + G D^T) with D = v_p - vloc_q and
    G = grad W(x_p - x_q), whichever of p, q is the pair's first index; then
    ViscousForcesTwoStep for fluid p: m_q (S_p c_p + S_q c_q) G with
    c_p = rholoc_q / (rho_p^2 rho_q), c_q = rholoc_p / (rho_q^2 rho_p)."""
    from oracle import oracle as O
    x, v, rho, m = st["x"], st["v"], st["rho"], sc.mass
    kf = O.poly7 if kernel == scenes.KERNEL_POLY7 else O.cubic
    S = np.zeros((sc.n, 2, 2))
    G = {}
    for i, j, d2 in zip(ip, jp, r2):
        for p, q in ((i, j), (j, i)):
            rpq = x[p] - x[q]
            _, gx, gy = kf(rpq[0], rpq[1], math.sqrt(d2), sc.h)
            g = np.array([gx, gy])
            G[(p, q)] = g
            D = v[p] - vloc[q]
            S[p] -= sc.mu * (m[q] * rholoc[q] / (rho[q] * rho[p])) * (np.outer(D, g) + np.outer(g, D))
    A = np.zeros((sc.n, 2))
    for (p, q), g in G.items():
        if p >= sc.n_fluid:
            continue
        cp = rholoc[q] / (rho[p] ** 2 * rho[q])
        cq = rholoc[p] / (rho[q] ** 2 * rho[p])
        A[p] += m[q] * ((S[p] * cp + S[q] * cq) @ g)
    return S, A


@pytest.mark.parametrize("kernel", KERNELS)
def test_two_step_gather_statement_agrees_with_scatter(oracle_lib, kernel):
    from oracle.oracle import Sim
    from test_oracle import np_search
    sc = scenes.dam_break(width=0.3, height=0.3, dx=0.02, tank=(0.6, 0.5), jitter=0.1, seed=3)
    sc.kernel = kernel
    sc.mode, sc.stress_type = scenes.MODE_REF, 2
    sc.kappa, sc.mu, sc.eps, sc.alpha_av, sc.gravity = 0.015, 1e-3, 1.0, 0.0, 0
    scenes.hydrostatic_walls(sc, 0.3)
    nf = sc.n_fluid
    rng = np.random.default_rng(11)
    sc.vel[:nf] = rng.normal(0, 0.3, (nf, 2))
    sc.vel[nf:] = rng.normal(0, 0.1, (sc.n - nf, 2))     # walls carry velocities too
    sc.rho[:nf] *= 1 + rng.uniform(-0.005, 0.005, nf)
    sc.pressure[:nf] = rng.uniform(0, 2000, nf)
    sc.mass[:nf] *= 1 + rng.uniform(-0.05, 0.05, nf)
    rl = sc.rho.copy()
    rl[:nf] *= 1 + rng.uniform(-1e-3, 1e-3, nf)
    vl = sc.vel + rng.normal(0, 0.05, sc.vel.shape)
    s = Sim(sc)
    s.set_local(rholoc=rl, Ploc=sc.pressure, vloc=vl)
    s.derivatives(sc.dt)
    st, ref, rr = s.state(), s.ref(), s.rates()
    ip, jp, r2 = np_search(st["x"], nf, sc.n, sc.particle_boundary, 40, sc.cutoff2)
    S, A = _gather_stress(sc, st, st["vloc"], st["rholoc"], ip, jp, r2, kernel)
    got = s.stress().reshape(sc.n, 2, 2)
    scale = np.abs(S).max()
    walls = np.arange(sc.n) >= nf
    assert np.abs(S[walls]).max() > 1e-3 * scale        # walls do get stresses
    assert np.abs(got - S).max() < 1e-10 * scale
    visc = rr["a"][:nf] - _pressure_accel(sc, st, st["rholoc"], ref["P"], ip, jp, r2, kernel,
                                          predicting=False)[:nf]
    scale = np.abs(A[:nf]).max()
    assert scale > 0
    assert np.abs(visc - A[:nf]).max() < 1e-8 * scale
