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
    config.addinivalue_line("markers", "slow: longer CPU case")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible device (GPU tests only)."""
    import raptr_amd
    from raptr_amd import device
    raptr_amd.lib()
    if device.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return raptr_amd


def oracle_geom(oracle, setup, grid):
    s = oracle.make_setup(setup.pos0x, setup.pos1x, setup.na, setup.n0z, setup.n0y,
                          setup.n1z, setup.n1y, setup.da, setup.segx, setup.segy, setup.segz)
    g = oracle.make_grid(grid.o, grid.d, grid.n)
    return s, g


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
