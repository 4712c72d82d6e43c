import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def cfg0():
    from help_amd import instances
    return instances.random_lp(2000, 4000, 4, 0)


@pytest.fixture(scope="session")
def small_mcf():
    from help_amd import instances
    return instances.mcf(side=12, commodities=3, pairs=6, seed=3)


def log_uniform(rng, n, lo=-6.0, hi=6.0):
    """d spread over many decades, as in late IPM iterations."""
    return np.exp(rng.uniform(lo, hi, n) * np.log(10.0))
