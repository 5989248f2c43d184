"""Batched replica trajectories (one graph launch per iteration for all
replicas) against each replica run alone, and against the CPU oracle's serial
replicas (same system, same uniforms)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from fricc.friccsd import FRICCSD
    from fricc.replicas import ReplicaBatch, log_uniform_m, run_alone
    from fricc.systems import SyntheticSystem
    from oracle import rccsd

    assert torch.cuda.is_available()
    return torch, FRICCSD, ReplicaBatch, run_alone, SyntheticSystem, rccsd, log_uniform_m


@pytest.mark.parametrize("prec,arenas", [("fp64", None), ("mixed", None), ("mixed", (3, 5)), ("fp64", (7,))])
def test_batch_equals_each_replica_alone(env, prec, arenas):
    """One branch per replica (default), two replicas per branch on arenas 3
    and 5, and all four on one branch: each replica bit for bit its run alone."""
    torch, FRICCSD, ReplicaBatch, run_alone, SyntheticSystem, rccsd, log_uniform_m = env
    eris = FRICCSD(SyntheticSystem(4, 14, 40, 8), {}).ao2mo()
    settings = {"compression": "fri", "sampling_method": "systematic", "precision": prec}
    m_list, seeds = [700, 1500, 2600, 1100], [11, 12, 13, 14]
    batch = ReplicaBatch(eris, settings, m_list, seeds, level_shift=2.0, arenas=arenas)
    batch.run(4)
    for r in range(len(seeds)):
        en, t2 = run_alone(eris, settings, m_list[r], seeds[r], 4, level_shift=2.0)
        assert [batch.energies[i][r] for i in range(4)] == en, r
        assert torch.equal(batch.t2[r], t2), r
    s = batch.summary()
    assert s["n_replicas"] == 4 and np.isfinite(s["stderr"])


def test_batch_average_matches_cpu_serial_average(env):
    """Replica-averaged E_corr and its standard error: the GPU batch against
    the oracle's replicas run one after another on the CPU (FP64, same Philox
    uniforms).  Iteration-1 energies agree to 1e-10; later ones can differ at
    a rounding boundary of the index set, so the averages are compared to
    1e-9 relative and the standard errors to 1e-6."""
    torch, FRICCSD, ReplicaBatch, run_alone, SyntheticSystem, rccsd, log_uniform_m = env
    from fricc.uniforms import PhiloxUniforms

    o, v, naux, seed = 4, 12, 30, 6
    eris = FRICCSD(SyntheticSystem(o, v, naux, seed), {}).ao2mo()
    oeris, _ = rccsd.synthetic_eris(o, v, naux, seed)
    n = o * o * v * v
    R = 6
    m_list = log_uniform_m(R, n // 20, n // 4, seed=1)
    seeds = list(range(R))
    settings = {"compression": "fri", "sampling_method": "systematic", "precision": "fp64"}
    batch = ReplicaBatch(eris, settings, m_list, seeds, level_shift=2.0)
    batch.run(3)
    cpu = []
    for m, sd in zip(m_list, seeds):
        s = rccsd.default_settings(m_keep=m, compression="fri", level_shift=2.0)
        *_, en = rccsd.kernel(oeris, s, PhiloxUniforms(sd), max_cycle=3, tol=0, tolnormt=0)
        cpu.append(en)
    cpu = np.asarray(cpu)
    gpu = np.asarray(batch.energies).T
    np.testing.assert_allclose(gpu[:, 0], cpu[:, 0], rtol=1e-10)
    g, c = gpu[:, -1], cpu[:, -1]
    assert abs(g.mean() - c.mean()) <= 1e-9 * abs(c.mean())
    assert abs(g.std(ddof=1) - c.std(ddof=1)) <= 1e-6 * abs(c.std(ddof=1)) + 1e-12
    s = batch.summary()
    assert abs(s["mean"] - g.mean()) < 1e-15 and s["n_replicas"] == R


def test_c4_64_replicas_on_the_c2_system(env):
    """c4 at its size: 64 replicas of the c2 system (nocc=20, nvir=180,
    naux=600; the c1 system cannot hold m up to 2e6), m log-uniform in
    [2e5, 2e6], Philox seeds 0..63, mixed precision, one graph launch per
    iteration for all.  Replicas sampled across the batch equal their runs
    alone bit for bit; the replica mean and standard error are finite and
    reported."""
    torch, FRICCSD, ReplicaBatch, run_alone, SyntheticSystem, rccsd, log_uniform_m = env
    from fricc.systems import DeviceERIs

    sysm = SyntheticSystem.from_config("c2")
    eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc)
    n = sysm.nocc ** 2 * sysm.nvir ** 2
    R, K = 64, 3
    m_list = log_uniform_m(R, 2e5, 2e6, seed=1, cap=n)
    assert min(m_list) >= 200_000 and max(m_list) <= 2_000_000
    settings = {"compression": "fri", "sampling_method": "systematic", "precision": "mixed"}
    batch = ReplicaBatch(eris, settings, m_list, list(range(R)), level_shift=2.0)
    print("c4: 64 replicas recorded", flush=True)
    batch.run(K)
    for r in (0, 29, 63):
        en, t2 = run_alone(eris, settings, m_list[r], r, K, level_shift=2.0)
        assert [batch.energies[i][r] for i in range(K)] == en, r
        assert torch.equal(batch.t2[r], t2), r
    s = batch.summary()
    print(f"c4 replica E_corr after {K} iterations: {s['mean']:.10f} +- {s['stderr']:.3g} (n={s['n_replicas']})")
    assert s["n_replicas"] == R and np.isfinite(s["mean"]) and 0 < s["stderr"] < 1.0
    del batch, eris
    torch.cuda.empty_cache()
