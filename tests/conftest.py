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
    config.addinivalue_line("markers", "slow: large parity case (seconds of CPU oracle time)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    return O


@pytest.fixture(scope="session")
def ddla():
    """The HIP library; a GPU test fails loudly if it was not built."""
    import ddla_amd
    ddla_amd.init(0)
    return ddla_amd


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def rel_fro(a, b):
    d = np.linalg.norm(a - b)
    s = np.linalg.norm(b)
    return d / s if s > 0 else d
