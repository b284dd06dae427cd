import gzip
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libslx.so")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~10 s")


@pytest.fixture(scope="session")
def prem_text():
    with gzip.open(os.path.join(GOLDEN, "prem_512.l96C.ump3.lm1.gz"), "rt") as fh:
        return fh.read()


@pytest.fixture(scope="session")
def topo_seed():
    return np.load(os.path.join(GOLDEN, "topography_seed_mm.npz"))["topo_mm"] / 1000.0


@pytest.fixture(scope="session")
def tgrid8_change():
    return np.load(os.path.join(GOLDEN, "tgrid8_change.npz"))["change"]
