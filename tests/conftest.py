import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    return pyoracle


def read_xyz(path):
    """Reference ASCII xyz fixture: count line, title line, `P x y z m` rows."""
    import numpy as np
    with open(path) as f:
        lines = [ln.split() for ln in f.read().strip().splitlines()]
    n = int(lines[0][0])
    rows = [ln for ln in lines[2:2 + n]]
    pos = np.array([[float(r[1]), float(r[2]), float(r[3])] for r in rows])
    mass = np.array([float(r[4]) for r in rows])
    return pos, mass
