import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libdft_hip.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def sto3g():
    from dft_amd.basis import load_basis
    return load_basis()


@pytest.fixture(scope="session")
def water_shells(sto3g):
    from dft_amd.basis import build_shells
    from dft_amd.geometry import water
    Z, xyz = water()
    return Z, xyz, build_shells(sto3g, Z, xyz)
