"""Independent FRI-CCSD replica trajectories advanced together.

Replica r is an ordinary ``FRICCSD`` solve of one system with its own
``m_keep`` and its own uniform stream (``PhiloxUniforms``: counter-based,
independent per seed; glibc's global generator cannot serve several
trajectories and its seeds 0 and 1 coincide).  The replicas share the
system's integrals and prepared operands.

``ReplicaBatch`` records one iteration of *every* replica -- each replica's
whole device step (``friccsd._update_amps_device``) followed by the copy of its
outputs into its inputs -- as a single graph, so one launch advances all R
trajectories by one iteration.  The graph has K branches (streams): replica r
runs on branch r mod K with scratch arena ``arenas[r mod K]``, so K replicas'
kernels share the GPU at any time (one replica's grids at the small sizes
replicas run at leave most CUs idle), and replicas on one branch run one
after another.  Per iteration the host writes the R x k uniforms with one
copy from pinned memory, replays, and reads every replica's energy and norm
in one transfer.  Each replica's arithmetic is exactly that of the replica
run alone (same kernels, same order, no kernel shared between replicas), so a
replica's trajectory in the batch is bit-identical to its trajectory alone.

The reference runs one trajectory per process and averages afterwards
(examples/03_friccsd_unbiased.py; no batched form exists there).
"""
import math

import numpy as np

from .friccsd import FRICCSD, _check_m_keep, _draw_uniforms, _pin_arena, arena_stream
from .uniforms import PhiloxUniforms


def _torch():
    import torch

    return torch


def log_uniform_m(n, lo, hi, seed, cap=None):
    """n values of m drawn log-uniformly in [lo, hi] (integers), optionally capped."""
    rng = np.random.default_rng(seed)
    m = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)
    if cap is not None:
        m = np.minimum(m, cap)
    return [int(x) for x in m]


class ReplicaBatch:
    """R replicas of one system (``eris``) stepped by one graph launch per iteration.

    settings: solver settings shared by all replicas (precision, compression,
    ...); ``m_list`` and ``seeds`` give each replica's m_keep and Philox seed;
    ``solver``: FRICCSD (default) or FRILCCSD; ``arenas``: the library scratch
    arenas (0..31) of the graph's branches, one branch per arena (default
    0..min(R, 8)-1); no other solve may use them while the batch lives."""

    def __init__(self, eris, settings, m_list, seeds, level_shift=0.0, solver=FRICCSD, arenas=None):
        torch = _torch()
        if len(m_list) != len(seeds):
            raise ValueError("one m and one seed per replica")
        self.eris = eris
        self.R = len(seeds)
        if arenas is None:
            arenas = range(min(self.R, 8))
        self.arenas = [int(a) for a in arenas]
        if not self.arenas or len(set(self.arenas)) != len(self.arenas) or not all(0 <= a < 32 for a in self.arenas):
            raise ValueError("arenas: distinct library arenas in 0..31")
        self.reps = []
        for r, (m, seed) in enumerate(zip(m_list, seeds)):
            cc = solver(eris, dict(settings, m_keep=int(m)), uniforms=PhiloxUniforms(seed))
            cc.level_shift = level_shift
            cc.arena = self.arenas[r % len(self.arenas)]
            if cc._iteration_dependent():
                raise ValueError("replica batches need the unbiased averaging off")
            self.reps.append(cc)
        # all uniforms of one iteration: device R x k, written from pinned host memory
        nu = max(2, max(cc._n_compressions() for cc in self.reps))
        self.u_dev = torch.zeros(self.R, nu, dtype=torch.float64, device="cuda")
        self.u_host = torch.zeros(self.R, nu, dtype=torch.float64, pin_memory=True)
        for r, cc in enumerate(self.reps):
            cc._u_dev = self.u_dev[r]
        t1, t2 = self.reps[0].get_init_guess(eris)
        for cc in self.reps:
            _check_m_keep(cc, t2)
        self.t1 = [t1.clone() for _ in range(self.R)]
        self.t2 = [t2.clone() for _ in range(self.R)]
        self.energies = []  # per iteration: R energies
        self.graph = None

    def _branches(self, st, body):
        """body(r, cc) for every replica, replica r on branch r mod K; the
        branches fork from and join back into stream st."""
        torch = _torch()
        streams = [arena_stream(a) for a in self.arenas]
        for b in streams:
            if b is not st:
                b.wait_stream(st)
        out = []
        for r, cc in enumerate(self.reps):
            with torch.cuda.stream(streams[r % len(streams)]):
                out.append(body(r, cc))
        for b in streams:
            if b is not st:
                st.wait_stream(b)
        return out

    def _record(self):
        torch = _torch()
        for cc in self.reps:
            cc._single_stream = True  # each step on its branch's one stream (see friccsd._update_amps_device)
        st = arena_stream(self.arenas[0])  # warm-up and capture on the first branch's stream
        st.wait_stream(torch.cuda.current_stream())
        # warm-up: one replica alone first, which builds the operands every
        # replica shares (prepared integrals, BF16 splits) before any other
        # branch reads them; then every replica on its branch, which allocates
        # each arena's workspaces and each replica's views
        with torch.cuda.stream(st):
            self.u_dev.fill_(0.5)
            self.reps[0]._device_step(self.t1[0], self.t2[0], self.eris)
        torch.cuda.synchronize()
        with torch.cuda.stream(st):
            self._branches(st, lambda r, cc: cc._device_step(self.t1[r], self.t2[r], self.eris))
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()

        def rec(r, cc):
            t1n, t2n, red = cc._device_step(self.t1[r], self.t2[r], self.eris)
            # the branch's stream orders the copies after the step that read t1[r], t2[r]
            self.t1[r].copy_(t1n)
            self.t2[r].copy_(t2n)
            return (t1n, t2n, red)

        with torch.cuda.graph(self.graph, stream=st):
            self.out = self._branches(st, rec)
            self.red = torch.stack([o[2] for o in self.out])
        torch.cuda.synchronize()
        # the graph holds the branches' arenas' slots (one stream each: no side arena)
        for a in self.arenas:
            _pin_arena(a, self)

    def step(self):
        """One iteration of every replica; returns the R energies."""
        if self.graph is None:
            self._record()
        for r, cc in enumerate(self.reps):
            us = _draw_uniforms(cc)
            for k, u in enumerate(us):
                self.u_host[r, k] = u
        self.u_dev.copy_(self.u_host, non_blocking=True)
        self.graph.replay()
        red = self.red.cpu().numpy()  # one transfer: {E2, |dt2|^2, E1, |dt1|^2} per replica
        en = []
        for r, cc in enumerate(self.reps):
            for k in sorted(cc._used_slots):
                cc._slots[k].sync_status()
            e = float(red[r, 0] + red[r, 2])
            cc._last_step = (self.t1[r], self.t2[r], e, math.sqrt(float(red[r, 1] + red[r, 3])))
            cc._iter += 1
            en.append(e)
        self.energies.append(en)
        return en

    def run(self, n_iter):
        for _ in range(n_iter):
            self.step()
        return self.summary()

    def summary(self, last=None):
        """Replica mean and standard error of E_corr at the last iteration (or
        averaged over the last ``last`` iterations per replica first)."""
        e = np.asarray(self.energies)
        x = e[-1] if not last else e[-last:].mean(axis=0)
        se = x.std(ddof=1) / math.sqrt(len(x)) if len(x) > 1 else float("nan")
        return {"mean": float(x.mean()), "stderr": float(se), "n_replicas": len(x), "n_iter": len(e)}


def run_alone(eris, settings, m, seed, n_iter, level_shift=0.0, solver=FRICCSD):
    """The same replica run by itself, eagerly (reference for the batch)."""
    cc = solver(eris, dict(settings, m_keep=int(m)), uniforms=PhiloxUniforms(seed))
    cc.level_shift = level_shift
    t1, t2 = cc.get_init_guess(eris)
    en = []
    for it in range(n_iter):
        cc._iter = it
        t1, t2 = cc.update_amps(t1, t2, eris)
        en.append(cc._last_step[2])
    return en, t2


__all__ = ["ReplicaBatch", "log_uniform_m", "run_alone"]
