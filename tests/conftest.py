import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libhmg.so")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def hm():
    """The product package with its default GPU context initialised."""
    import highmap_amd

    highmap_amd.init_gpu(0)
    return highmap_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle


# Parameters of the 1024^2 vpipes workload (seed 2): rain 1e-3, dt 0.5,
# k_capacity 0.1, k_erode 0.05, k_depose 0.05, evaporation 0.01.
VP_WORKLOAD = dict(water_height=1e-3, evap_rate=0.01, dt=0.5, k_capacity=0.1, k_erode=0.05,
                   k_depose=0.05)
