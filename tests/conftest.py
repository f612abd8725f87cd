import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle64():
    from oracle import Oracle
    return Oracle(f64=True)


@pytest.fixture(scope="session")
def manta():
    from mantaflow_amd import manta as m
    return m


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
