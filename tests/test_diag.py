"""The diagnostics statement (oracle/diag_ref.py) against closed forms."""
import math

import numpy as np

from oracle import diag_ref


def _lattice(nx, ny, dx):
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    return np.stack([(i.ravel() + 0.5) * dx, (j.ravel() + 0.5) * dx], 1)


def test_energy_closed_forms():
    x = _lattice(20, 30, 0.01)
    m = np.full(len(x), 0.1)
    v = np.tile([0.3, -0.4], (len(x), 1))
    ke, pe = diag_ref.total_energy(x, v, m, 9.81)
    M = m.sum()
    assert math.isclose(ke, 0.5 * M * 0.25, rel_tol=1e-12)
    assert math.isclose(pe, M * 9.81 * x[:, 1].mean(), rel_tol=1e-12)


def test_probe_weights_integrate_to_one_and_average_a_uniform_field():
    """The weights are a normalised Gaussian truncated at 3h: on a lattice their sum
    approximates pi h^2 (1 - 10 e^-9) / (2 pi h^2 coeff1) = 1; a uniform pressure is
    returned exactly; the count is the brute-force count of particles within 3h."""
    dx, h = 0.01, 0.013
    x = _lattice(60, 60, dx)
    rho = np.full(len(x), 1000.0)
    m = rho * dx * dx
    P = np.full(len(x), 1234.5)
    p, norm, n = diag_ref.pressure_probe(x, P, m, rho, h, 0.3, 0.3)
    assert math.isclose(p, 1234.5, rel_tol=1e-13)
    assert abs(norm - 1.0) < 2e-3
    r = np.hypot(x[:, 0] - 0.3, x[:, 1] - 0.3)
    assert n == int((r <= 3 * h).sum())
    # negative pressures are left out; no particle in range gives norm 1
    P2 = P.copy()
    P2[r < h] = -1.0
    _, _, n2 = diag_ref.pressure_probe(x, P2, m, rho, h, 0.3, 0.3)
    assert n2 == n - int((r < h).sum())
    assert diag_ref.pressure_probe(x, P, m, rho, h, 5.0, 5.0) == (None, 1.0, 0)


def test_density_consistency_reduces_to_the_mean():
    rho = np.random.default_rng(1).uniform(990, 1010, 500)
    one, two, three, unc = diag_ref.density_consistency(rho, rho)
    assert one == 0.0
    assert math.isclose(two, rho.mean(), rel_tol=1e-12) and math.isclose(three, rho.mean(), rel_tol=1e-12)
    assert math.isclose(unc, -rho.mean(), rel_tol=1e-12)
