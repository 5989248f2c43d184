"""Continuous-batching FRI-CCSD solve server on one GPU.

Requests -- a synthetic density-fitted system (nocc, nvir), m and a uniform
seed -- arrive over time.  The server admits them in arrival order while the
estimated HBM of all admitted solves stays within a budget.  Each admitted
solve gets its own HIP stream, its own library scratch arena, device-resident
state and its own recorded step (``friccsd.GraphStep``: the whole iteration is
one graph launch).  The iterations of the admitted solves interleave on the
device, one solve's work at a time (by default the solve with the least
remaining work goes next); whenever a solve's iteration has finished the
scheduler reads that solve's energy (one small transfer), retires it when
converged or out of iterations, and otherwise lines up its next iteration.
Requests that arrived in the meantime are admitted as soon as memory frees
up; their systems are generated on the host while the device works.

A solve's arithmetic does not depend on what else runs (its own buffers,
scratch and graph; no shared accumulators), so its energies are bit-identical
alone and in the mix.

The reference has no serving layer (one solve per Python process,
fricc/friccsd.py:379-456); this is the device-resident loop of that kernel,
scheduled many at a time.
"""
import math
import time
from collections import OrderedDict
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .friccsd import FRICCSD, GraphStep, arena_stream
from .uniforms import PhiloxUniforms


def _torch():
    import torch

    return torch


@dataclass
class Request:
    id: int
    arrival: float  # seconds after the start of the run
    nocc: int
    nvir: int
    m: int
    seed: int = 0  # uniform stream (Philox)
    sys_seed: int = 1  # synthetic system
    naux: int = 0  # 0: 3 * (nocc + nvir)
    max_cycle: int = 50
    tol: float = 1e-6  # |dE| to stop
    level_shift: float = 2.0


@dataclass
class Result:
    id: int
    nocc: int
    nvir: int
    m: int
    arrival: float
    admitted: float
    finished: float
    iterations: int
    converged: bool
    energies: list = field(default_factory=list)
    error: str = ""  # why the solve was refused or stopped, if it was

    @property
    def latency(self):
        return self.finished - self.arrival


def poisson_requests(n, rate, sizes, mfrac=(0.01, 0.10), seed=0, max_cycle=50, tol=1e-6):
    """n requests with exponential inter-arrival times (rate per second),
    (nocc, nvir) drawn uniformly from ``sizes`` and m / |T2| uniform in mfrac."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        o, v = sizes[int(rng.integers(len(sizes)))]
        m = max(1, int(rng.uniform(*mfrac) * o * o * v * v))
        out.append(Request(i, float(t[i]), o, v, m, seed=int(rng.integers(1 << 31)), sys_seed=1 + (i % 3),
                           max_cycle=max_cycle, tol=tol))
    return out


def estimate_bytes(o, v, precision="mixed"):
    """HBM of one admitted solve in steady state: the FP64 integral blocks it
    keeps, the operands prepared from them, the ladder (FP64 in the FP64 mode,
    else only its BF16 planes, built slice by slice from the factors),
    amplitudes and graph temporaries, views and scratch."""
    ov3, o2v2, v4 = float(o) * v ** 3, float(o * v) ** 2, float(v) ** 4
    if precision == "fp64":
        return 8.0 * v4 + 64.0 * ov3 + 340.0 * o2v2 + 2.0e9
    return 4.0 * v4 + 32.0 * ov3 + 300.0 * o2v2 + 2.0e9


def transient_bytes(o, v, precision="mixed"):
    """Extra HBM while a solve is set up: one FP64 ladder slice and its
    converted copy, and the FP64 source of one reduced-precision operand."""
    if precision == "fp64":
        return 8.0 * float(o) * v ** 3 + 0.5e9
    return 2.0 * (1 << 31) + 8.0 * float(o) * v ** 3 + 0.5e9


def device_available():
    """Bytes held by nothing: free device memory plus what the caching
    allocator holds unused (a footprint is the drop of this figure)."""
    torch = _torch()
    free, _ = torch.cuda.mem_get_info()
    return free + torch.cuda.memory_reserved() - torch.cuda.memory_allocated()


def device_free(need):
    """Device memory a request of ``need`` bytes can count on: what the device
    reports free, plus what the caching allocator holds unused when the two
    together cover the request (a retired solve's buffers, which the next
    solve of its shape reuses; the allocator returns unused segments to the
    device by itself when an allocation does not fit them).  Otherwise the
    unused segments are returned to the device first and its free memory
    counts."""
    torch = _torch()
    free, _ = torch.cuda.mem_get_info()
    if free >= need:
        return free
    avail = device_available()
    if avail >= need:
        return avail
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    return free


def _solve_failure(err):
    """True when ``err`` is a failure of one solve that the server answers and
    goes on from: a refused request (ValueError, e.g. m > |T2|) or a status a
    step reported about its own data (non-finite amplitudes, p > 1).  A device
    error or any other runtime error is not: the context may be unusable, and
    it must reach the caller."""
    from . import _capi

    if isinstance(err, _capi.FriccError):
        return isinstance(err, ValueError) or err.status in (_capi.FRICC_EPROB, _capi.FRICC_ENONFINITE)
    return isinstance(err, ValueError)


def _release_arenas(n):
    """Free the library scratch of arenas 1..n and of their side arenas
    (waits for the device)."""
    from . import _capi
    from .friccsd import side_arena

    for a in range(1, n + 1):
        _capi.check(_capi.lib().fricc_workspace_release(a))
        _capi.check(_capi.lib().fricc_workspace_release(side_arena(a)))


def system_key(req):
    """(nocc, nvir, naux, sys_seed): the synthetic system a request names."""
    return (req.nocc, req.nvir, req.naux or 3 * (req.nocc + req.nvir), req.sys_seed)


class SystemCache:
    """Host factors of the systems requests name.  Each is generated on a
    worker thread as soon as a request for it arrives (numpy releases the
    interpreter lock while it fills arrays, so the scheduler keeps queueing
    iterations meanwhile) and kept for later requests of the same system,
    least recently requested first out beyond ``max_bytes``."""

    def __init__(self, max_bytes=16e9):
        self.max_bytes = float(max_bytes)
        self.entries = OrderedDict()  # system_key -> Future of SyntheticSystem
        self._pool = None

    def request(self, req):
        """The future of the request's system (submitted on first request)."""
        from .systems import SyntheticSystem

        k = system_key(req)
        fut = self.entries.get(k)
        if fut is None:
            if self._pool is None:
                self._pool = ThreadPoolExecutor(max_workers=1)
            fut = self.entries[k] = self._pool.submit(SyntheticSystem, *k)
        self.entries.move_to_end(k)
        total = sum(f.result().L.nbytes for f in self.entries.values() if f.done() and f.exception() is None)
        for key in list(self.entries)[:-1]:
            if total <= self.max_bytes:
                break
            f = self.entries[key]
            if f.done():
                if f.exception() is None:
                    total -= f.result().L.nbytes
                del self.entries[key]
        return fut


class Solve:
    """One admitted request: its system, solver, recorded step and stream.
    ``stream``: the stream to run on (a new one when None; the server reuses
    one per arena, since the BLAS library keeps a workspace per stream it has
    seen).  ``system``: the request's SyntheticSystem when already generated."""

    def __init__(self, req, arena, precision, stream=None, system=None):
        torch = _torch()
        from .systems import DeviceERIs, SyntheticSystem

        self.req = req
        self.stream = stream if stream is not None else torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            sysm = system if system is not None else SyntheticSystem(*system_key(req))
            # reduced-precision solves never hold the FP64 ladder: its operand
            # is built slice by slice from the factors
            self.eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc,
                                                build_ladder=(precision == "fp64"))
            cc = FRICCSD(self.eris, {"m_keep": int(req.m), "compression": "fri", "sampling_method": "systematic",
                                     "precision": precision}, uniforms=PhiloxUniforms(req.seed))
            cc.level_shift = req.level_shift
            cc.arena = arena
            self.cc = cc
            t1, t2 = cc.get_init_guess(self.eris)
            self.e = cc.energy(t1, t2, self.eris)
            self.runner = GraphStep(cc, self.eris, t1, t2)
        self.energies = []
        self.done = False
        self.converged = False
        self.t_iter = None  # device time of the last iteration (ms), when measured
        self.ready_at = 0.0

    def issue(self):
        torch = _torch()
        with torch.cuda.stream(self.stream):
            self.cc._iter = len(self.energies)
            self.runner.issue()

    def close(self):
        """Drop the recorded step (which unpins the solve's scratch arena)."""
        self.runner = None

    def ready(self):
        return self.stream.query()

    def complete(self, order=None):
        """Read the finished iteration; True when the solve is done.  With an
        ``order`` (DeviceOrder), the copies that follow the read are queued
        after the work queued before them."""
        torch = _torch()
        with torch.cuda.stream(self.stream):
            self.runner.complete(None if order is None else (lambda: order.begin(self.stream)))
        e = self.cc._last_step[2]
        de = e - self.e
        self.e = e
        self.energies.append(e)
        if abs(de) < self.req.tol:
            self.converged = True
        self.done = self.converged or len(self.energies) >= self.req.max_cycle or not math.isfinite(e)
        return self.done


# Whether solves' device work may run side by side: every kernel a step
# launches must finish without waiting for other workgroups of its grid.  The
# step's kernels are the library's own (compression, gathered and dense
# contractions, finalize) and PyTorch's elementwise and reduction kernels;
# none of them waits for another workgroup.  (Vendor stream-K GEMMs, whose
# workgroups wait for partial tiles of their own grid, stalled side-by-side
# replays when the dense remainder used them.)
OVERLAP_SAFE = True


class DeviceOrder:
    """Runs the device work of all solves one after another, in the order it
    is queued, whatever stream it is on: work queued on a stream first waits
    for the end of the work queued before it (``begin``), and its own end is
    what the next work waits for (``end``).  The host never waits for it."""

    def __init__(self):
        self.tail = None

    def begin(self, stream):
        if self.tail is not None:
            stream.wait_event(self.tail)

    def end(self, stream):
        ev = _torch().cuda.Event(enable_timing=True)
        ev.record(stream)
        self.tail = ev
        return ev


class Server:
    """Admission by HBM budget, per-solve streams and arenas, continuous batching.

    Up to ``max_active`` solves are admitted at once and their iterations
    interleave.  With ``overlap`` False (the default) the device runs one
    solve's work at a time (DeviceOrder), a solve's set-up starts once the
    device is idle, and ``policy`` picks the next iteration: "srpt" gives the
    device to the solve with the least remaining work (shortest remaining
    processing time first, which minimises the mean latency of a single
    server), with up to ``depth`` iterations queued so the device does not
    wait for the host -- but never one that would hold back the next
    iteration of a solve in flight that comes first; "rr" queues every
    solve's next iteration as soon as its previous one has finished (round
    robin).  The BLAS library's GEMM
    kernels (stream-K: workgroups of one grid wait for partial tiles from
    other workgroups of the same grid) assume that their whole grid can be
    resident at once; work of other solves running beside them -- other such
    grids, or the sparse GEMM that holds a whole CU per workgroup -- breaks
    that assumption, and two grids each waiting for their own unscheduled
    workgroups block each other.  With ``overlap`` True, solves' iterations
    run side by side on the device; that is validated only on small systems
    (a run of several production-size solves side by side stopped making
    progress).

    Side by side (``overlap``) every kernel of a step must make progress with
    other solves' kernels resident on the device: all of them do (no kernel
    waits for workgroups of its own grid; ``OVERLAP_SAFE``).

    A request is admitted while (a) the admitted solves' footprints plus its
    own stay within ``hbm_budget``, (b) the device still has room for its
    footprint and set-up transients and (c) it needs at most half the budget
    or nothing else is running.  A footprint is the larger of
    ``estimate_bytes`` and the most the device lost to an earlier set-up of
    the same shape (measured around ``Solve``)."""

    margin = 2.0e9  # device memory left free beyond each admission's need

    depth = 2  # iterations queued on the device at once ("srpt"): one runs while the host reads the other

    def __init__(self, hbm_budget=240e9, precision="mixed", max_active=8, progress=None, overlap=False,
                 policy="srpt", record_timeline=False):
        if policy not in ("srpt", "rr"):
            raise ValueError(f"unknown policy {policy!r} (srpt, rr)")
        if overlap and not OVERLAP_SAFE:
            raise ValueError("Server(overlap=True) is disabled: a step's kernels may wait for other workgroups "
                             "of their grid, which side-by-side replays can starve")
        self.progress = progress  # called with each admission / retirement / deferral record
        self.budget = float(hbm_budget)
        self.precision = precision
        self.max_active = min(int(max_active), 31)  # arenas 1..31; 33..63 are their side arenas
        self.overlap = bool(overlap)
        self.policy = policy
        # (request id, start event, end event) of every queued iteration when
        # asked for (one solve's work at a time only)
        self.timeline = [] if record_timeline else None
        self.systems = SystemCache()
        self.measured = {}
        self.log = []  # (time, event, id, nocc, nvir, need, used, device_free)

    def _note(self, rec):
        self.log.append(rec)
        if self.progress is not None:
            self.progress(rec)

    def _need(self, req):
        key = (req.nocc, req.nvir)
        return max(estimate_bytes(req.nocc, req.nvir, self.precision), self.measured.get(key, 0.0))

    def _reclaim(self, idle):
        """Free the library scratch of idle arenas (grow-only buffers a retired
        solve left behind).  Waits for the device once."""
        from . import _capi
        from .friccsd import side_arena

        torch = _torch()
        rel = sorted(self._dirty & set(idle))
        for a in rel:
            _capi.check(_capi.lib().fricc_workspace_release(a))
            _capi.check(_capi.lib().fricc_workspace_release(side_arena(a)))
            self._dirty.discard(a)
        torch.cuda.empty_cache()
        return bool(rel)

    def _fits(self, req, used, active, idle):
        need = self._need(req)
        if active and used + need > self.budget:
            return False
        if active and need > 0.5 * self.budget:
            # a solve that takes most of the budget is admitted alone: its
            # set-up beside running solves is not validated on the device
            return False
        want = need + transient_bytes(req.nocc, req.nvir, self.precision) + self.margin
        if device_free(want) >= want:
            return True
        if self._reclaim(idle) and _torch().cuda.mem_get_info()[0] >= want:
            return True
        return not active

    def _stream(self, arena):
        # the arena's stream: the solve's step is recorded on it (its own BLAS
        # workspace) and replayed on it
        return arena_stream(arena)

    def _try_solve(self, req, arena, system=None):
        """Set up a solve; None when the device ran out of memory doing so
        (its partial allocations are released)."""
        torch = _torch()
        try:
            return Solve(req, arena, self.precision, stream=self._stream(arena), system=system)
        except (torch.OutOfMemoryError, MemoryError, RuntimeError) as e:
            if isinstance(e, RuntimeError) and "out of memory" not in str(e) and "hipMalloc" not in str(e):
                raise
        import gc

        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        return None

    def _priority(self, s):
        """Order in which waiting solves get the device (smallest first): a
        solve not yet measured first (its first iteration measures it), then
        the least remaining work, iterations left times the solve's measured
        iteration time; "rr": the order in which they became ready."""
        if self.policy == "rr":
            return (1, s.ready_at)
        if s.t_iter is None:
            return (0, s.ready_at)
        return (1, (s.req.max_cycle - len(s.energies)) * s.t_iter)

    def _may_follow(self, s, queued):
        """Whether s may be queued behind the iterations in flight: not when a
        solve in flight would be picked before s for its next iteration (s
        would hold that iteration back by one of its own)."""
        if self.policy != "srpt":
            return True
        for q in queued:
            left = q.req.max_cycle - len(q.energies) - 1
            if q.t_iter is None or left <= 0:
                continue
            if self._priority(s) > (1, left * q.t_iter):
                return False
        return True

    def _queue(self, s, order):
        """Queue the solve's next iteration behind the work queued so far and
        time it on the device."""
        torch = _torch()
        order.begin(s.stream)
        s.ev0 = torch.cuda.Event(enable_timing=True)
        s.ev0.record(s.stream)
        s.issue()
        s.ev1 = order.end(s.stream)
        if self.timeline is not None:
            self.timeline.append((s.req.id, s.ev0, s.ev1))

    def run(self, requests, clock=time.perf_counter):
        torch = _torch()
        from . import _capi

        by_arrival = sorted(requests, key=lambda r: r.arrival)
        pending = list(by_arrival)
        arrived = 0  # requests (in arrival order) whose system has been asked for
        free_arenas = list(range(1, self.max_active + 1))
        active, results, used = [], [], 0.0
        queued, waiting = [], []  # iteration on the device / ready for its next one
        self.max_concurrent = 0
        self.deferred = 0
        self._dirty = set()  # arenas whose scratch may hold memory
        wait_retire = False  # a set-up ran out of memory: admit again after a retirement
        order = None if self.overlap else DeviceOrder()
        depth = self.max_active if (order is None or self.policy == "rr") else self.depth
        t0 = clock()
        while pending or active:
            now = clock() - t0
            # a request's system is generated on the host as soon as it arrives
            while arrived < len(by_arrival) and by_arrival[arrived].arrival <= now:
                self.systems.request(by_arrival[arrived])
                arrived += 1
            while pending and pending[0].arrival <= now and free_arenas and not wait_retire:
                sysf = self.systems.request(pending[0])
                if not sysf.done() and active:
                    break  # keep the device fed until the system is ready
                if not self._fits(pending[0], used, active, free_arenas):
                    break
                req = pending[0]
                system = sysf.result()
                if order is not None:
                    torch.cuda.synchronize()  # the set-up runs with nothing else on the device
                before = device_available()
                arena = free_arenas.pop(0)
                self._dirty.add(arena)
                try:
                    s = self._try_solve(req, arena, system)
                    if s is None and not active and self._reclaim(free_arenas + [arena]):
                        before = device_available()
                        self._dirty.add(arena)
                        s = self._try_solve(req, arena, system)
                except (ValueError, RuntimeError) as err:
                    if not _solve_failure(err):
                        raise
                    # the solver refused the request (e.g. m > |T2|): answer it, go on
                    pending.pop(0)
                    free_arenas.insert(0, arena)
                    fin = clock() - t0
                    results.append(Result(req.id, req.nocc, req.nvir, req.m, req.arrival, fin, fin, 0, False,
                                          error=f"{type(err).__name__}: {err}"))
                    self._note((round(fin, 3), "refused", req.id, req.nocc, req.nvir, None, used, None))
                    continue
                if s is None:
                    free_arenas.insert(0, arena)
                    if not active:
                        raise MemoryError(f"solve {req.id} ({req.nocc}x{req.nvir}) does not fit on the device alone")
                    key = (req.nocc, req.nvir)
                    self.measured[key] = max(self.measured.get(key, 0.0), 1.25 * self._need(req))
                    self.deferred += 1
                    self._note((round(clock() - t0, 3), "deferred", req.id, req.nocc, req.nvir, None, used,
                                torch.cuda.mem_get_info()[0]))
                    wait_retire = True
                    break
                pending.pop(0)
                s.admitted = s.ready_at = clock() - t0
                key = (req.nocc, req.nvir)
                self.measured[key] = max(self.measured.get(key, 0.0), float(before - device_available()))
                s.need = self._need(req)
                used += s.need
                self._note((round(s.admitted, 3), "admit", req.id, req.nocc, req.nvir, s.need, used,
                            torch.cuda.mem_get_info()[0]))
                active.append(s)
                waiting.append(s)
                self.max_concurrent = max(self.max_concurrent, len(active))
            if not active:
                time.sleep(max(0.0, min(0.05, pending[0].arrival - (clock() - t0))))
                continue
            progressed = False
            for s in list(queued):
                if not s.ready():
                    continue
                progressed = True
                queued.remove(s)
                if order is not None:
                    s.t_iter = s.ev0.elapsed_time(s.ev1)
                error = ""
                try:
                    done = s.complete(order)
                except (ValueError, RuntimeError) as err:
                    if not _solve_failure(err):
                        raise
                    # a status the step reported about its own data (non-finite amplitudes, p > 1)
                    done, error = True, f"{type(err).__name__}: {err}"
                if order is not None:
                    order.end(s.stream)  # the copies of the outputs into the inputs
                if done:
                    fin = clock() - t0
                    results.append(Result(s.req.id, s.req.nocc, s.req.nvir, s.req.m, s.req.arrival, s.admitted, fin,
                                          len(s.energies), s.converged, list(s.energies), error=error))
                    active.remove(s)
                    s.close()
                    used -= s.need
                    self._note((round(fin, 3), "retire", s.req.id, s.req.nocc, s.req.nvir, s.need, used, None))
                    free_arenas.append(s.cc.arena)
                    free_arenas.sort()
                    wait_retire = False
                else:
                    s.ready_at = clock() - t0
                    waiting.append(s)
            while waiting and len(queued) < depth:
                if order is None:
                    s = waiting.pop(0)
                    s.issue()
                else:
                    s = min(waiting, key=self._priority)
                    if queued and not self._may_follow(s, queued):
                        break
                    waiting.remove(s)
                    self._queue(s, order)
                queued.append(s)
                progressed = True
            s = None
            # a retired solve's buffers stay in the allocator's cache, where the
            # next solve of that shape finds them; an admission that finds the
            # device short returns the cache first (device_free)
            if not progressed:
                time.sleep(5e-5)
        torch.cuda.synchronize()
        _release_arenas(self.max_active)
        self.makespan = clock() - t0
        return sorted(results, key=lambda r: r.id)


def latency_stats(results, makespan):
    """Throughput and latency of the solves that ran (refused or failed ones
    are counted apart)."""
    ok = [r for r in results if not r.error]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    return {"solves": len(ok), "solves_per_s": len(ok) / makespan if makespan > 0 else None,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_iterations": float(np.mean([r.iterations for r in ok])) if ok else 0.0,
            "converged": int(sum(r.converged for r in ok)), "failed": len(results) - len(ok),
            "makespan_s": makespan}


def solve_alone(req, precision="mixed"):
    """The request run by itself (same recorded step), for comparison."""
    torch = _torch()
    s = Solve(req, 0, precision, stream=arena_stream(0))
    s.issue()
    while True:
        s.stream.synchronize()
        if s.complete():
            return list(s.energies)
        s.issue()
