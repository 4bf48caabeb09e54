import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs on an MI355X through the HIP library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def same_bits(a, b):
    """True when two float64 arrays hold identical bit patterns."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_difference(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    bad = np.nonzero(a.view(np.uint64) != b.view(np.uint64))[0]
    if bad.size == 0:
        return None
    i = int(bad[0])
    return f"{bad.size} differ; first at {i}: {a[i]!r} vs {b[i]!r}"


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def hip():
    from neofiz_amd import _native
    return _native.binding()
