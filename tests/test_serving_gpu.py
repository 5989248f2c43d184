"""Serving loop: solves admitted by an HBM budget interleave their
iterations (own streams, scratch arenas and recorded steps), one solve's work
at a time on the device or side by side; each solve's energies are the same
alone and in the mix."""
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


@pytest.mark.parametrize("overlap,policy", [(False, "srpt"), (False, "rr"), (True, "rr")])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_alone_equals_in_the_mix(serving, prec, overlap, policy):
    sv = serving
    reqs = _requests(sv)
    server = sv.Server(hbm_budget=1e12, precision=prec, max_active=8, overlap=overlap, policy=policy)
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


def test_dropped_solve_frees_its_memory(serving):
    """A retired solve's device buffers go back as soon as it is dropped (no
    reference cycle holds them until a cyclic collection), and a second solve
    on the same stream leaves nothing behind either (per-stream library
    workspaces are created once)."""
    import gc

    import torch

    sv = serving
    req = sv.Request(0, 0.0, 6, 40, m=5000, seed=3, sys_seed=2, max_cycle=2, tol=0.0)
    st = torch.cuda.Stream()

    def run_once():
        s = sv.Solve(req, 1, "mixed", stream=st)
        s.issue()
        s.stream.synchronize()
        s.complete()
        return s

    del_first = run_once()
    del del_first
    gc.collect()
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    gc.disable()
    try:
        s = run_once()
        held = torch.cuda.memory_allocated() - base
        assert held > (8 << 20), held
        del s
        torch.cuda.synchronize()
        assert torch.cuda.memory_allocated() - base <= (64 << 10), (held, torch.cuda.memory_allocated() - base)
    finally:
        gc.enable()


def test_refused_request_does_not_stop_the_others(serving):
    """A request the solver refuses at set-up (m > |T2|, fricc/friccsd.py:68-72)
    is answered with its error; the solves around it run as alone."""
    sv = serving
    reqs = _requests(sv)[:3]
    o, v = reqs[1].nocc, reqs[1].nvir
    reqs[1].m = o * o * v * v + 1
    server = sv.Server(hbm_budget=1e12, precision="mixed", max_active=8)
    res = server.run(reqs)
    assert len(res) == 3
    assert "m_keep" in res[1].error and res[1].iterations == 0
    for r, q in ((res[0], reqs[0]), (res[2], reqs[2])):
        assert not r.error and r.energies == sv.solve_alone(q, precision="mixed")
    st = sv.latency_stats(res, server.makespan)
    assert st["solves"] == 2 and st["failed"] == 1


def test_one_solves_work_at_a_time(serving):
    """With overlap off, the iterations of different solves occupy disjoint
    intervals of device time (each waits for the end of the work queued
    before it), while the solves still interleave."""
    sv = serving
    server = sv.Server(hbm_budget=1e12, precision="mixed", max_active=8, record_timeline=True)
    res = server.run(_requests(sv))
    assert len(res) == 8 and server.max_concurrent > 1
    t0 = server.timeline[0][1]  # the first queued iteration starts the chain
    iv = sorted((t0.elapsed_time(a), t0.elapsed_time(b), rid) for rid, a, b in server.timeline)
    assert len(iv) == sum(r.iterations for r in res)
    for (s1, e1, _), (s2, _, _) in zip(iv, iv[1:]):
        assert s2 >= e1 - 1e-3, (s1, e1, s2)
    # interleaved: some solve's iterations come in more than one run
    ids = [rid for _, _, rid in iv]
    runs = 1 + sum(a != b for a, b in zip(ids, ids[1:]))
    assert runs > len(set(ids))


def test_c5_sizes_side_by_side(serving):
    """Three c5 sizes (10x90, 20x180, 32x256) replayed side by side on the
    device (overlap), several solves at once: every solve completes and its
    energies equal those it has alone."""
    import torch

    sv = serving
    sizes = [(10, 90), (20, 180), (32, 256), (20, 180), (10, 90), (32, 256)]
    reqs = [sv.Request(i, 0.0, o, v, m=int(0.03 * o * o * v * v), seed=700 + i, sys_seed=1, max_cycle=4, tol=0.0)
            for i, (o, v) in enumerate(sizes)]
    server = sv.Server(hbm_budget=240e9, precision="mixed", max_active=8, overlap=True)
    res = server.run(reqs)
    assert [r.iterations for r in res] == [4] * len(reqs) and not any(r.error for r in res)
    assert server.max_concurrent >= 3
    torch.cuda.empty_cache()
    for r, q in zip(res, reqs):
        assert r.energies == sv.solve_alone(q, precision="mixed"), (q.nocc, q.nvir)
        torch.cuda.empty_cache()
    print(f"c5 side by side: makespan {server.makespan:.2f} s, max concurrent {server.max_concurrent}")


def test_c5_sizes_alone_equal_in_the_mix(serving):
    """The four c5 sizes (10x90, 20x180, 32x256, 48x352; m/|T2| in [1%, 10%],
    mixed precision, 240 GB budget) served together, a few iterations each:
    every solve's energies are the same as when it runs alone."""
    import torch

    sv = serving
    sizes = [(10, 90), (20, 180), (32, 256), (48, 352)]
    fr = [0.01, 0.10, 0.04, 0.02]
    reqs = [sv.Request(i, 0.02 * i, o, v, m=int(f * o * o * v * v), seed=500 + i, sys_seed=1, max_cycle=3, tol=0.0)
            for i, ((o, v), f) in enumerate(zip(sizes, fr))]
    server = sv.Server(hbm_budget=240e9, precision="mixed", max_active=8)
    res = server.run(reqs)
    assert [r.iterations for r in res] == [3] * 4 and not any(r.error for r in res)
    torch.cuda.empty_cache()
    for r, q in zip(res, reqs):
        print(f"c5 {q.nocc}x{q.nvir}: in the mix {r.energies}", flush=True)
        assert r.energies == sv.solve_alone(q, precision="mixed"), (q.nocc, q.nvir)
        torch.cuda.empty_cache()
    st = sv.latency_stats(res, server.makespan)
    print(f"c5 sizes: makespan {server.makespan:.2f} s, p50 {st['p50_s']:.2f} s, p99 {st['p99_s']:.2f} s")
