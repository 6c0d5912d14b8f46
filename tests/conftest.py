import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden_scene(name, **over):
    """Scene from a fixture written by tests/golden/make_fixtures.py."""
    from uph_amd.scenes import scene_from_rows
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    prm = dict(zip([str(k) for k in z["param_names"]], z["param_values"].tolist()))
    return scene_from_rows(z["fluid"], z["fixed"], z["moving"], prm, **over)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle


def oracle_spread(sc, steps, so, amplitudes=(1e-12, 1e-10, 1e-8)):
    """Largest relative density and position (in dx) differences, after `steps`, of the
    oracle from its own run `so` when the fluid positions start moved by up to each of
    `amplitudes` (times dx), one run each: the conditioning of a trajectory whose
    capped neighbour sets change (each flip of a neighbour in or out of the capped
    set moves the rates by a kernel weight, whatever the size of the perturbation)."""
    from oracle.oracle import Sim
    nf = sc.n_fluid
    dr, dx = 0.0, 0.0
    for k, a in enumerate(amplitudes):
        sp = sc.copy()
        sp.pos[:nf] += np.random.default_rng(7 + k).uniform(-a, a, (nf, 2)) * sc.dx
        o = Sim(sp, nthreads=0)
        o.step(steps)
        st = o.state()
        dr = max(dr, float(np.max(np.abs(st["rho"][:nf] - so["rho"][:nf]) / so["rho"][:nf])))
        dx = max(dx, float(np.abs(st["x"][:nf] - so["x"][:nf]).max()) / sc.dx)
    return dr, dx
