"""Request generation and the admission memory model (CPU)."""
import numpy as np

from fricc import serving as sv


def test_poisson_requests_shape():
    sizes = [(10, 90), (20, 180), (32, 256)]
    reqs = sv.poisson_requests(400, rate=4.0, sizes=sizes, mfrac=(0.01, 0.10), seed=3)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.25) < 0.05  # rate 4 / s
    for r in reqs:
        n = r.nocc ** 2 * r.nvir ** 2
        assert (r.nocc, r.nvir) in sizes and 0.01 * n - 1 <= r.m <= 0.10 * n


def test_memory_model_monotone():
    a = sv.estimate_bytes(10, 90)
    b = sv.estimate_bytes(20, 180)
    c = sv.estimate_bytes(32, 256)
    assert a < b < c and c < 240e9
    # the FP64 mode keeps the FP64 ladder and FP64 operands
    assert sv.estimate_bytes(32, 256, "fp64") > c


def test_large_solve_is_admitted_alone():
    """A request needing more than half the budget waits until nothing else
    runs; smaller requests are still checked against the budget.  (Both
    refusals are decided before any device query.)"""
    big = sv.Request(0, 0.0, 48, 352, m=1)
    small = sv.Request(1, 0.0, 10, 90, m=1)
    server = sv.Server(hbm_budget=1.05 * sv.estimate_bytes(48, 352), precision="mixed")
    running = [object()]
    assert sv.estimate_bytes(48, 352) > 0.5 * server.budget
    assert not server._fits(big, used=1.0, active=running, idle=[])
    assert not server._fits(small, used=server.budget, active=running, idle=[])
