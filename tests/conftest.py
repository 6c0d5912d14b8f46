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
