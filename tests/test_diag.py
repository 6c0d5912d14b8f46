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


def test_kernel_w_matches_the_pinned_oracle_kernels():
    """The vectorised W of the statement equals the oracle's cubic spline and
    PolyOrderSeven (pinned by their own closed-form tests) from q = 0 past q = 2."""
    from oracle import oracle as O
    h = 0.013
    for r in np.linspace(0.0, 2.2 * h, 45):
        w0 = O.cubic(r, 0.0, r, h)[0]
        w1 = O.poly7(r, 0.0, r, h)[0]
        assert math.isclose(float(diag_ref.kernel_w(0, r, h)), w0, rel_tol=1e-12, abs_tol=1e-9)
        assert math.isclose(float(diag_ref.kernel_w(1, r, h)), w1, rel_tol=1e-12, abs_tol=1e-9)


def _pairs(x, nf, pb, h):
    """The search contract's pairs without the K cap: fluid i < j, j < pb, d2 < 4 h^2."""
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    i, j = np.nonzero((d2 < 4 * h * h) & (np.arange(len(x))[:, None] < np.arange(len(x))[None, :]))
    keep = (i < nf) & (j < pb)
    return i[keep], j[keep]


def test_vorticity_of_a_rigid_rotation_is_the_lattice_sum():
    """For v = omega z x (x - c) the pair term is omega r^2 W / (r^2 + 0.01 h^2), so a
    particle with full support gets omega times the lattice sum of dx^2 W(r) r^2 / (r^2 +
    0.01 h^2) over its neighbours (positive: counter-clockwise).  A uniform translation
    gives exactly 0."""
    dx, h, om = 0.01, 0.013, 2.5
    x = _lattice(30, 30, dx)
    n = len(x)
    c = x.mean(0)
    v = om * np.stack([-(x[:, 1] - c[1]), x[:, 0] - c[0]], 1)
    rho = np.full(n, 1000.0)
    m = rho * dx * dx
    ip, jp = _pairs(x, n, n, h)
    for kind in (0, 1):
        w = diag_ref.approx_vorticity(x, v, rho, m, h, ip, jp, n, kind)
        k = np.arange(-3, 4)
        off = np.stack(np.meshgrid(k, k, indexing="ij"), -1).reshape(-1, 2) * dx
        r2 = (off ** 2).sum(1)
        r2 = r2[(r2 > 0) & (r2 < 4 * h * h)]
        S = float(np.sum(dx * dx * diag_ref.kernel_w(kind, np.sqrt(r2), h) * r2 / (r2 + 0.01 * h * h)))
        inner = np.all((x > 2 * h) & (x < 30 * dx - 2 * h), axis=1)
        assert inner.sum() > 400 and S > 0.3
        assert np.allclose(w[inner], om * S, rtol=1e-12, atol=0)
        vt = np.tile([0.7, -0.2], (n, 1))
        assert np.all(diag_ref.approx_vorticity(x, vt, rho, m, h, ip, jp, n, kind) == 0.0)


def test_vorticity_scatter_equals_a_per_particle_gather():
    """The pair-list form (added to i, and to j when j is fluid) against a direct sum
    over each fluid particle's admitted partners, with walls (rows >= nf), the pair
    filter j < pb and non-uniform masses, densities and velocities."""
    rng = np.random.default_rng(3)
    dx, h = 0.01, 0.013
    x = _lattice(16, 16, dx) + rng.uniform(-0.1, 0.1, (256, 2)) * dx
    n, nf = len(x), 200
    pb = n - 1
    v = rng.normal(0, 0.4, (n, 2))
    rho = rng.uniform(990, 1010, n)
    m = rng.uniform(0.09, 0.11, n)
    ip, jp = _pairs(x, nf, pb, h)
    got = diag_ref.approx_vorticity(x, v, rho, m, h, ip, jp, nf, 0)
    exp = np.zeros(nf)
    for p in range(nf):
        for q in range(n):
            if q == p:
                continue
            lo, hi = min(p, q), max(p, q)
            if not (lo < nf and hi < pb):
                continue
            e = x[p] - x[q]
            r2 = float(e @ e)
            if r2 >= 4 * h * h:
                continue
            dv = v[p] - v[q]
            a = (dv[1] * e[0] - dv[0] * e[1]) * float(diag_ref.kernel_w(0, math.sqrt(r2), h)) / (r2 + 0.01 * h * h)
            exp[p] += m[q] / rho[q] * a
    assert np.allclose(got, exp, rtol=1e-12, atol=1e-12 * np.abs(exp).max())


def test_probe_two_closed_forms():
    dx, h = 0.01, 0.013
    x = _lattice(40, 40, dx)
    P = np.full(len(x), 777.0)
    p, n = diag_ref.probe_two(x, P, h, 0.2, 0.2)
    r = np.hypot(x[:, 0] - 0.2, x[:, 1] - 0.2)
    assert n == int((r < 2 * h).sum()) and n > 5
    assert math.isclose(p, 777.0, rel_tol=1e-14)
    P2 = P.copy()
    P2[r < h] = -5.0                       # non-positive pressures are left out
    p2, n2 = diag_ref.probe_two(x, P2, h, 0.2, 0.2)
    assert n2 == n - int((r < h).sum()) and math.isclose(p2, 777.0, rel_tol=1e-14)
    assert diag_ref.probe_two(x, P, h, 9.0, 9.0) == (0.0, 0)
