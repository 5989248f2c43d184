"""Device-resident FRILCCSD vs the oracle's NumPy restatement of
fricc/frilccsd.py on the same integrals (bit-identical fma-chain ERIs), the
same MP2 guess and the same uniforms.

Bar: sampled index sets identical; t1/t2 within 1e-10 relative (FP64);
precision modes within their stated tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from fricc.frilccsd import FRILCCSD
    from fricc.systems import SyntheticSystem
    from oracle import lccsd, rccsd

    assert torch.cuda.is_available()
    return torch, FRILCCSD, SyntheticSystem, lccsd, rccsd


def _pair(env, o, v, naux, seed, settings, us):
    torch, FRILCCSD, SyntheticSystem, lccsd, rccsd = env
    sysm = SyntheticSystem(o, v, naux, seed)
    cc = FRILCCSD(sysm, settings, uniforms=iter(list(us)))
    cc.level_shift = settings.get("level_shift", 0.0)
    eris = cc.ao2mo()
    oeris, _ = rccsd.synthetic_eris(o, v, naux, seed)
    keys = ("m_keep", "compression", "compressed_contractions", "LCCD", "Independent Compressions", "level_shift")
    s = lccsd.default_settings(**{k: settings[k] for k in settings if k in keys})
    return cc, eris, oeris, s


def _perturbed(t1, t2, seed):
    rng = np.random.default_rng(seed)
    t1 = t1 + 0.02 * rng.standard_normal(t1.shape)
    t2 = t2 + 0.005 * rng.standard_normal(t2.shape)
    return t1, t2


@pytest.mark.parametrize("ccs", [["O^2V^4", "O^3V^3", "O^4V^2"], ["O^2V^4"], ["O^3V^3"], ["O^4V^2"], []])
@pytest.mark.parametrize("variant", ["fri", "largest", "lccd", "independent"])
def test_update_amps_fp64_matches_oracle(env, ccs, variant):
    torch, lccsd, rccsd = env[0], env[3], env[4]
    o, v, naux = 4, 13, 30
    settings = {"m_keep": 900, "compression": "largest" if variant == "largest" else "fri",
                "sampling_method": "systematic", "compressed_contractions": ccs,
                "LCCD": variant == "lccd", "Independent Compressions": variant == "independent"}
    us = [0.31, 0.77, 0.12, 0.55, 0.93, 0.08]
    cc, eris, oeris, s = _pair(env, o, v, naux, 5, settings, us)
    _, t1, t2 = rccsd.init_amps(oeris)
    t1, t2 = _perturbed(t1, t2, 1)
    rt1, rt2, sp = lccsd.update_amps(t1, t2, oeris, s, iter(us), return_sparse=True)
    dt1, dt2 = cc.update_amps(torch.tensor(t1, device="cuda"), torch.tensor(t2, device="cuda"), eris)
    for k, ts in enumerate(sp):
        assert np.array_equal(cc._slots[k].indices(), ts.idx), k
    e2 = np.abs(dt2.cpu().numpy() - rt2).max() / np.abs(rt2).max()
    if variant == "lccd":
        assert not dt1.cpu().numpy().any()
        e1 = 0.0
    else:
        e1 = np.abs(dt1.cpu().numpy() - rt1).max() / np.abs(rt1).max()
    assert e2 < 1e-10 and e1 < 1e-10, (e1, e2)
    # the fused finalize's energy is lin_cc_energy of the new amplitudes
    assert abs(cc._last_step[2] - lccsd.energy(rt1, rt2, oeris)) < 1e-10 * abs(lccsd.energy(rt1, rt2, oeris))


@pytest.mark.parametrize("prec,tol,etol", [("fp32", 5e-6, 1e-6), ("mixed", 2e-5, 5e-6)])
def test_precision_modes(env, prec, tol, etol):
    torch = env[0]
    o, v, naux = 6, 24, 40
    us = [0.42]
    base = {"m_keep": 6000, "compression": "fri", "sampling_method": "systematic"}
    c64, eris, _, _ = _pair(env, o, v, naux, 3, base, us)
    cm, erism, _, _ = _pair(env, o, v, naux, 3, dict(base, precision=prec), us)
    t1, t2 = c64.get_init_guess(eris)
    t1 = t1 + 0.01 * torch.randn_like(t1)
    a1, a2 = c64.update_amps(t1, t2, eris)
    b1, b2 = cm.update_amps(t1, t2, erism)
    rel = (b2 - a2).abs().max().item() / a2.abs().max().item()
    assert rel < tol, rel
    assert np.array_equal(c64._slots[0].indices(), cm._slots[0].indices())
    de = abs(c64.energy(a1, a2, eris) - cm.energy(b1, b2, erism))
    assert de < etol, de


def test_kernel_trajectory_matches_oracle(env):
    torch, FRILCCSD, SyntheticSystem, lccsd, rccsd = env
    o, v, naux = 4, 14, 40
    us = list(np.random.default_rng(9).random(20))
    settings = {"m_keep": 1500, "compression": "fri", "sampling_method": "systematic", "level_shift": 2.0}
    cc, eris, oeris, s = _pair(env, o, v, naux, 8, settings, us)
    cc.max_cycle = 5
    cc.conv_tol = 0
    cc.conv_tol_normt = 0
    cc.kernel()
    _, _, _, _, en = lccsd.kernel(oeris, s, iter(us), max_cycle=5, tol=0, tolnormt=0)
    assert len(cc.energies) == 5
    np.testing.assert_allclose(cc.energies, en, rtol=1e-9)


@pytest.mark.parametrize("independent", [False, True])
def test_graph_replay_matches_eager(env, independent):
    torch, FRILCCSD, SyntheticSystem, lccsd, rccsd = env
    from fricc.friccsd import GraphStep

    us = list(np.random.default_rng(4).random(60))
    base = {"m_keep": 1500, "compression": "fri", "sampling_method": "systematic", "precision": "mixed",
            "Independent Compressions": independent}
    runs = {}
    for graph in (False, True):
        cc = FRILCCSD(SyntheticSystem(4, 14, 40, 8), base, uniforms=iter(us))
        cc.level_shift = 2.0
        eris = cc.ao2mo()
        t1, t2 = cc.get_init_guess(eris)
        runner = GraphStep(cc, eris, t1, t2) if graph else None
        energies, idx = [], []
        for it in range(4):
            t1, t2 = runner.step(t1, t2) if graph else cc.update_amps(t1, t2, eris)
            energies.append(cc._last_step[2])
            idx.append([cc._slots[k].indices().copy() for k in sorted(cc._used_slots)])
        runs[graph] = (energies, idx, t2.cpu().numpy().copy())
    assert runs[True][0] == runs[False][0]
    for a, b in zip(runs[True][1], runs[False][1]):
        assert len(a) == (6 if independent else 1)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    assert np.array_equal(runs[True][2], runs[False][2])
