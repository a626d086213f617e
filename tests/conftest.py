import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "particleincellcodegolf.jl_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU physics check")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def ulp_diff(a, b):
    """Distance in units in the last place between two finite doubles."""
    import numpy as np
    ia = np.asarray(a, dtype=np.float64).view(np.int64).astype(object)
    ib = np.asarray(b, dtype=np.float64).view(np.int64).astype(object)

    def key(i):
        return i if i >= 0 else -(i & 0x7FFFFFFFFFFFFFFF)
    fa = np.vectorize(key, otypes=[object])(ia)
    fb = np.vectorize(key, otypes=[object])(ib)
    return np.abs((fa - fb).astype(np.float64))
