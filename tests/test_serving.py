"""Serving host logic (CPU): request stream and memory-based slot sizing."""
import numpy as np

from pathfinding_gpu_amd.serving import BYTES_PER_SLOT_VERTEX, poisson_requests, slots_for_memory


def test_poisson_stream():
    reqs = poisson_requests(1000.0, 20000, {"grid": 0.9, "rmat": 0.1}, {"grid": 4_000_000, "rmat": 4_194_304},
                            {"grid": True, "rmat": False}, seed=3)
    t = np.array([r.arrival for r in reqs])
    assert t[0] == 0.0 and (np.diff(t) >= 0).all()
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1e-3) < 5e-5                      # rate
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05           # exponential: cv = 1
    frac = np.mean([r.graph == "grid" for r in reqs])
    assert abs(frac - 0.9) < 0.01
    assert all(0 <= r.src < 4_000_000 and 0 <= r.dst < 4_000_000 for r in reqs if r.graph == "grid")
    assert all(r.dst == -1 and 0 <= r.src < 4_194_304 for r in reqs if r.graph == "rmat")
    again = poisson_requests(1000.0, 100, {"grid": 0.9, "rmat": 0.1}, {"grid": 10, "rmat": 10},
                             {"grid": True, "rmat": False}, seed=3)
    assert [r.arrival for r in again] == [r.arrival for r in reqs[:100]]


def test_slots_for_memory():
    n = 4_000_000
    free = 200 * 2**30
    k = slots_for_memory(n, free, 0.5)
    assert k * BYTES_PER_SLOT_VERTEX * n <= 0.5 * free < (k + 1) * BYTES_PER_SLOT_VERTEX * n or k == 1024
    assert slots_for_memory(n, 1, 0.5) == 1
    assert slots_for_memory(10, 2**40, 1.0, cap=77) == 77
    assert slots_for_memory(2**30, 2**50, 1.0) * 2**30 < 2**32
