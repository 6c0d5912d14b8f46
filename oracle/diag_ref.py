"""FP64 numpy statement of the diagnostics sph.cc's output reports.

TEST INFRASTRUCTURE ONLY: the checker for uph_diagnostics (the device
reductions in uph_amd/csrc/uph_diag.hip).  Pinned by closed forms in
tests/test_diag.py (uniform fields, the probe weights' normalisation, brute-force
neighbour counts).

  total_energy       TotalKineticEnergy, src/lib/derivatives.c:2254-2263
  pressure_probe     NeighborsToPressurePoint, derivatives.c:2290-2338
  density_consistency DensityConsistencyFunctions, derivatives.c:3090-3117
  probe_two          NeighborsToPressurePointTwo + EnsembleAveragedPointPressureTwo,
                     derivatives.c:2340-2393
  kernel_w           W of the cubic spline (derivatives.c:44-66) and of
                     PolyOrderSevenConvolutionKernel (derivatives.c:767-784), vectorised
  approx_vorticity   ApproximateVortcity, derivatives.c:2840-2885
"""
from __future__ import annotations

import math

import numpy as np


def total_energy(x, v, m, g):
    """(kinetic, potential) = (sum 1/2 m |v|^2, sum m g y) over the given (fluid) rows."""
    ke = float(np.sum(0.5 * m * (v[:, 0] ** 2 + v[:, 1] ** 2)))
    pe = float(np.sum(m * g * x[:, 1]))
    return ke, pe


def pressure_probe(x, P, m, rho, h, px, py):
    """(pressure, norm, count): the Gaussian-weighted average of the positive pressures
    within 3 h of (px, py); norm 1 and pressure None when no particle qualifies."""
    r = np.sqrt((x[:, 0] - px) ** 2 + (x[:, 1] - py) ** 2)
    sel = (r <= 3.0 * h) & (P > 0.0)
    coeff1 = 0.5 - 5.0 * math.exp(-9.0)
    w = (np.exp(-r[sel] ** 2 / (h * h)) - math.exp(-9.0)) * m[sel] / (2.0 * math.pi * h * h * coeff1 * rho[sel])
    n = int(sel.sum())
    if n == 0:
        return None, 1.0, 0
    return float(np.sum(P[sel] * w) / np.sum(w)), float(np.sum(w)), n


def density_consistency(rho, rholoc):
    """(one, two, three, uncertainty).  The reference adds its first sum outside the
    OpenMP reduction and then assigns the reduction's total (0) to it."""
    n = len(rho)
    one = 0.0
    two = float(np.sum(rholoc / n))
    three = float(np.sum(rho / n))
    return one, two, three, one - two * two / three


def probe_two(x, P, h, px, py):
    """(mean, count) of the positive pressures strictly within 2h of (px, py); (0, 0)
    when none (the averaging loop then does not run)."""
    r = np.sqrt((x[:, 0] - px) ** 2 + (x[:, 1] - py) ** 2)
    sel = (r < 2.0 * h) & (P > 0.0)
    n = int(sel.sum())
    return (float(np.sum(P[sel] / n)) if n else 0.0), n


def kernel_w(kind, r, h):
    """W at distances r: kind 0 the cubic spline, alpha = 10 / (7 pi h^2); kind 1
    PolyOrderSeven, alpha (1 - q^2/4)^8 (2048 - 3072 q^2 + 2816 q^4 - 1984 q^6 + 1184 q^8
    - 628 q^10 + 305 q^12), alpha = 2145 / (1347584 pi h^2).  Zero for q >= 2."""
    r = np.asarray(r, np.float64)
    q = r / h
    if kind == 0:
        a = 10.0 / (7.0 * math.pi * h * h)
        w = np.where(q < 1.0, a * (1.0 - 1.5 * q * q + 0.75 * q ** 3), a * 0.25 * (2.0 - q) ** 3)
    else:
        a = 2145.0 / (1347584.0 * math.pi * h * h)
        u = q * q
        poly = 2048 - 3072 * u + 2816 * u ** 2 - 1984 * u ** 3 + 1184 * u ** 4 - 628 * u ** 5 + 305 * u ** 6
        w = a * (1.0 - 0.25 * u) ** 8 * poly
    return np.where(q < 2.0, w, 0.0)


def approx_vorticity(x, v, rho, m, h, ip, jp, nf, kind):
    """approxVorticity of rows < nf from the pair list (ip, jp): for each pair
    a = (dv_y dx_x - dv_x dx_y) W / (r^2 + 0.01 h^2) with dx = x_i - x_j, dv = v_i - v_j,
    added as (m_j / rho_j) a to i and, when j is fluid, as (m_i / rho_i) a to j."""
    ip, jp = np.asarray(ip, np.int64), np.asarray(jp, np.int64)
    dx = x[ip] - x[jp]
    dv = v[ip] - v[jp]
    r2 = dx[:, 0] ** 2 + dx[:, 1] ** 2
    a = (dv[:, 1] * dx[:, 0] - dv[:, 0] * dx[:, 1]) * kernel_w(kind, np.sqrt(r2), h) / (r2 + 0.01 * h * h)
    out = np.zeros(nf)
    np.add.at(out, ip, m[jp] / rho[jp] * a)
    fl = jp < nf
    np.add.at(out, jp[fl], (m[ip] / rho[ip] * a)[fl])
    return out
