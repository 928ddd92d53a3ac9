import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmbxg.so")


@pytest.fixture(scope="session")
def systems():
    return np.load(os.path.join(GOLDEN, "systems.npz"))


def water_case(systems, key):
    """(xyz[9n], box or None, cutoff2b, cutoff3b, published[1B,2B,3B]) of a fixture system."""
    xyz = systems[key + "/xyz"]
    box = systems[key + "/box"]
    c = systems[key + "/cutoffs"]
    return xyz, (box if box.size else None), float(c[0]), float(c[1]), systems[key + "/published"]


def sorted_tuples(a):
    """Canonical form of an n-mer list: sort within each tuple, then the tuples."""
    a = np.sort(np.asarray(a, np.int64), axis=1)
    if len(a) == 0:
        return a
    return a[np.lexsort(a.T[::-1])]
