"""FP64 numpy statement of the diagnostics sph.cc's output reports.

TEST INFRASTRUCTURE ONLY: the checker for uph_diagnostics (the device
reductions in uph_amd/csrc/uph_gpu.hip).  Pinned by closed forms in
tests/test_diag.py (uniform fields, the probe weights' normalisation, brute-force
neighbour counts).

  total_energy       TotalKineticEnergy, src/lib/derivatives.c:2254-2263
  pressure_probe     NeighborsToPressurePoint, derivatives.c:2290-2338
  density_consistency DensityConsistencyFunctions, derivatives.c:3090-3117
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
