"""Request generation, the admission memory model and the replica batch's argument checks (CPU)."""
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


def test_srpt_priority():
    """Shortest remaining work first: an unmeasured solve (its first
    iteration) before any measured one, then iterations left x measured
    iteration time; round robin keeps the order in which solves became ready."""
    from types import SimpleNamespace as NS

    def solve(max_cycle, done, t_iter, ready_at):
        return NS(req=NS(max_cycle=max_cycle), energies=[0.0] * done, t_iter=t_iter, ready_at=ready_at)

    big_late = solve(50, 45, 30.0, 0.1)     # 5 left x 30 ms = 150
    small = solve(50, 10, 2.0, 0.3)         # 40 left x 2 ms = 80
    fresh = solve(50, 0, None, 0.4)
    big = solve(50, 0, 30.0, 0.2)           # 1500
    server = sv.Server(policy="srpt")
    order = sorted([big_late, small, fresh, big], key=server._priority)
    assert order == [fresh, small, big_late, big]
    rr = sv.Server(policy="rr")
    assert sorted([big_late, small, fresh, big], key=rr._priority) == [big_late, big, small, fresh]
    import pytest

    with pytest.raises(ValueError):
        sv.Server(policy="lifo")


def test_system_cache_keeps_recent_systems():
    """Requests of one system share its host factors (generated once, on the
    worker thread); beyond the byte bound the least recently requested
    system is dropped."""
    import numpy as np

    from fricc.systems import SyntheticSystem

    one = 39 * 13 * 13 * 8  # L of a 3x10 system, naux 39
    cache = sv.SystemCache(max_bytes=2.5 * one)
    reqs = [sv.Request(i, 0.0, 3, 10, m=1, sys_seed=s) for i, s in enumerate([1, 2, 1, 3])]
    f0 = cache.request(reqs[0])
    np.testing.assert_array_equal(f0.result().L, SyntheticSystem(3, 10, 39, 1).L)
    cache.request(reqs[1]).result()
    assert cache.request(reqs[2]) is f0
    cache.request(reqs[3]).result()
    cache.request(reqs[3])
    assert list(cache.entries) == [(3, 10, 39, 1), (3, 10, 39, 3)]


def test_srpt_second_slot_never_delays_the_shorter_solve():
    """With a short solve in flight, a long solve's iteration is not queued
    behind it (it would delay the short solve's next iteration by a long
    one); the other way round it is, and a solve in flight on its last
    iteration holds nothing back."""
    from types import SimpleNamespace as NS

    def solve(done, t_iter):
        return NS(req=NS(max_cycle=50), energies=[0.0] * done, t_iter=t_iter, ready_at=0.0)

    server = sv.Server(policy="srpt")
    short, long_ = solve(10, 2.0), solve(10, 250.0)
    assert not server._may_follow(long_, [short])
    assert server._may_follow(short, [long_])
    assert server._may_follow(long_, [solve(49, 2.0)])
    assert server._may_follow(long_, [solve(0, None)])
    assert sv.Server(policy="rr")._may_follow(long_, [short])


def test_overlap_refused_while_unsafe(monkeypatch):
    """Side-by-side replay is refused whenever a step's kernels may wait for
    workgroups of their own grid (OVERLAP_SAFE off), and allowed otherwise."""
    import pytest

    monkeypatch.setattr(sv, "OVERLAP_SAFE", False)
    with pytest.raises(ValueError, match="overlap"):
        sv.Server(overlap=True)
    monkeypatch.setattr(sv, "OVERLAP_SAFE", True)
    assert sv.Server(overlap=True).overlap


def test_step_uses_no_vendor_gemm():
    """The amplitude updates call no PyTorch matrix product (which would run a
    vendor BLAS kernel): every dense product goes through fricc.dense."""
    import inspect
    import re

    from fricc import friccsd, frilccsd

    for mod in (friccsd, frilccsd):
        src = inspect.getsource(mod)
        for pat in (r"torch\.(matmul|bmm|baddbmm|addmm|einsum|mm|mv)\(", r"\.(matmul|bmm|baddbmm_?|addmm_?|mv)\(",
                    r"\w\)? @ \w"):
            bad = [line for line in src.splitlines() if re.search(pat, line.split("#")[0])
                   and not line.strip().startswith(('"', "'"))]
            assert not bad, (mod.__name__, pat, bad[:3])


def test_solve_failure_classification():
    """Only a refused request or a step's own data error is answered as one
    solve's failure; device errors and other runtime errors are not."""
    from fricc import _capi

    assert sv._solve_failure(ValueError("m_keep"))
    assert sv._solve_failure(_capi.FriccValueError(_capi.FRICC_EINVAL, "bad"))
    assert sv._solve_failure(_capi.FriccError(_capi.FRICC_EPROB, "p > 1"))
    assert sv._solve_failure(_capi.FriccError(_capi.FRICC_ENONFINITE, "nan"))
    assert not sv._solve_failure(_capi.FriccError(_capi.FRICC_EDEVICE, "illegal address"))
    assert not sv._solve_failure(_capi.FriccError(_capi.FRICC_ENOMEM, "oom"))
    assert not sv._solve_failure(RuntimeError("HIP error: an illegal memory access was encountered"))


def _fake_device(monkeypatch, solve_cls):
    """Replace the device layer the scheduler touches (torch.cuda, streams)
    with host stand-ins, and the solve with ``solve_cls``."""
    from types import SimpleNamespace as NS

    class Event:
        def __init__(self, **kw):
            self.t = 0.0

        def record(self, stream=None):
            pass

        def elapsed_time(self, other):
            return 1.0

    class Stream:
        def wait_event(self, ev):
            pass

        def synchronize(self):
            pass

        def query(self):
            return True

    cuda = NS(synchronize=lambda: None, mem_get_info=lambda: (1 << 40, 1 << 40), memory_reserved=lambda: 0,
              memory_allocated=lambda: 0, empty_cache=lambda: None, Event=Event, Stream=Stream,
              OutOfMemoryError=MemoryError)
    fake = NS(cuda=cuda, OutOfMemoryError=MemoryError)
    monkeypatch.setattr(sv, "_torch", lambda: fake)
    monkeypatch.setattr(sv, "arena_stream", lambda arena: Stream())
    monkeypatch.setattr(sv, "Solve", solve_cls)


def _fake_solve(fail_status):
    from fricc import _capi

    class FakeSolve:
        def __init__(self, req, arena, precision, stream=None, system=None):
            self.req, self.stream, self.energies, self.converged = req, stream, [], False
            self.cc = type("CC", (), {"arena": arena})()
            self.t_iter = None

        def issue(self):
            pass

        def ready(self):
            return True

        def complete(self, order=None):
            raise _capi.FriccError(fail_status, "mocked status")

        def close(self):
            pass

    return FakeSolve


def test_device_error_reaches_the_caller(monkeypatch):
    """A device error while a solve runs (mocked FRICC_EDEVICE status) stops
    the server and reaches the caller instead of being recorded as one
    solve's error."""
    import pytest

    from fricc import _capi

    _fake_device(monkeypatch, _fake_solve(_capi.FRICC_EDEVICE))
    server = sv.Server(hbm_budget=1e12, precision="mixed", max_active=2)
    monkeypatch.setattr(server.systems, "request", lambda req: type("F", (), {"done": lambda s: True,
                                                                            "result": lambda s: None})())
    with pytest.raises(_capi.FriccError) as ei:
        server.run([sv.Request(0, 0.0, 3, 10, m=50, max_cycle=2)])
    assert ei.value.status == _capi.FRICC_EDEVICE


def test_data_error_is_one_solves_failure(monkeypatch):
    """A step's own data error (mocked FRICC_ENONFINITE) retires that solve
    with the error recorded; the server goes on."""
    from fricc import _capi

    _fake_device(monkeypatch, _fake_solve(_capi.FRICC_ENONFINITE))
    server = sv.Server(hbm_budget=1e12, precision="mixed", max_active=2)
    monkeypatch.setattr(server.systems, "request", lambda req: type("F", (), {"done": lambda s: True,
                                                                            "result": lambda s: None})())
    monkeypatch.setattr(sv, "_release_arenas", lambda n: None)
    res = server.run([sv.Request(0, 0.0, 3, 10, m=50, max_cycle=2), sv.Request(1, 0.0, 3, 10, m=50, max_cycle=2)])
    assert len(res) == 2 and all("mocked status" in r.error for r in res)


def test_replica_batch_rejects_bad_arenas():
    """Branch arenas must be distinct library arenas 0..31 (checked before
    any device work)."""
    import pytest

    from fricc.replicas import ReplicaBatch

    for arenas in ([], [1, 1], [32], [-1]):
        with pytest.raises(ValueError, match="arenas"):
            ReplicaBatch(None, {}, [10], [0], arenas=arenas)
    with pytest.raises(ValueError, match="one m and one seed"):
        ReplicaBatch(None, {}, [10, 20], [0])
