"""Serving loop: solves admitted by an HBM budget run side by side (own
streams, scratch arenas and recorded steps); each solve's energies are the
same alone and in the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serving():
    import torch

    from fricc import serving as sv

    assert torch.cuda.is_available()
    return sv


def _requests(sv):
    sizes = [(3, 10), (4, 12), (4, 14), (5, 12)]
    reqs = []
    for i, (o, v) in enumerate(sizes * 2):
        n = o * o * v * v
        reqs.append(sv.Request(i, 0.01 * i, o, v, m=int(n * (0.05 + 0.01 * i)), seed=100 + i, sys_seed=1 + i % 3,
                               max_cycle=6, tol=0.0))
    return reqs


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_alone_equals_in_the_mix(serving, prec):
    sv = serving
    reqs = _requests(sv)
    server = sv.Server(hbm_budget=1e12, precision=prec, max_active=8)
    res = server.run(reqs)
    assert len(res) == len(reqs) and server.max_concurrent > 1
    for r, q in zip(res, reqs):
        assert r.iterations == q.max_cycle
        assert r.energies == sv.solve_alone(q, precision=prec), q.id
    st = sv.latency_stats(res, server.makespan)
    assert st["p99_s"] >= st["p50_s"] > 0 and st["solves"] == len(reqs)


def test_budget_limits_admission(serving):
    sv = serving
    reqs = _requests(sv)[:4]
    one = max(sv.estimate_bytes(q.nocc, q.nvir, "fp64") for q in reqs)
    server = sv.Server(hbm_budget=1.5 * one, precision="fp64", max_active=8)
    res = server.run(reqs)
    assert server.max_concurrent == 1 and len(res) == 4
    # served one after another in arrival order
    fin = [r.finished for r in res]
    assert fin == sorted(fin)
