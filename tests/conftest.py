import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: full-size workload")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def random_graph(rng: np.random.Generator, n: int, m: int, wmax: int = 20, zero_w: bool = True):
    """Random directed multigraph edge list with self-loops, duplicates and (optionally) zero weights."""
    src = rng.integers(0, n, m, dtype=np.uint32)
    dst = rng.integers(0, n, m, dtype=np.uint32)
    lo = 0 if zero_w else 1
    w = rng.integers(lo, wmax + 1, m, dtype=np.uint32)
    return src, dst, w
