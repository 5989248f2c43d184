"""Device-resident FRICCSD.update_amps vs the oracle's NumPy restatement of
fricc/friccsd.py:29-301 on the same integrals (bit-identical fma-chain ERIs),
the same MP2 guess (bit-identical) and the same uniforms.

Bar: sampled index sets identical; t1/t2 within 1e-10 relative (FP64 plain);
"mixed" precision within its stated tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from fricc.friccsd import FRICCSD
    from fricc.systems import SyntheticSystem
    from oracle import rccsd

    assert torch.cuda.is_available()
    return torch, FRICCSD, SyntheticSystem, rccsd


def _pair(env, o, v, naux, seed, settings, us):
    torch, FRICCSD, SyntheticSystem, rccsd = env
    sysm = SyntheticSystem(o, v, naux, seed)
    cc = FRICCSD(sysm, settings, uniforms=iter(list(us)))
    cc.level_shift = settings.get("level_shift", 0.0)
    eris = cc.ao2mo()
    oeris, _ = rccsd.synthetic_eris(o, v, naux, seed)
    s = rccsd.default_settings(**{k: settings[k] for k in settings if k in ("m_keep", "compression",
                                                                            "compressed_contractions",
                                                                            "vvvv_t1t1", "level_shift",
                                                                            "sampling_method")})
    return cc, eris, oeris, s


def _perturbed(t1, t2, seed):
    rng = np.random.default_rng(seed)
    t1 = t1 + 0.02 * rng.standard_normal(t1.shape)
    t2 = t2 + 0.005 * rng.standard_normal(t2.shape)
    return t1, t2


@pytest.mark.parametrize("ccs", [["O^2V^4", "O^3V^3", "O^4V^2"], ["O^2V^4"], ["O^3V^3"], ["O^4V^2"], []])
@pytest.mark.parametrize("comp", ["fri", "largest"])
def test_update_amps_fp64_matches_oracle(env, ccs, comp):
    torch = env[0]
    o, v, naux = 4, 13, 30
    settings = {"m_keep": 900, "compression": comp, "sampling_method": "systematic",
                "compressed_contractions": ccs}
    us = [0.31, 0.77]
    cc, eris, oeris, s = _pair(env, o, v, naux, 5, settings, us)
    _, t1, t2 = env[3].init_amps(oeris)
    gt1, gt2 = cc.get_init_guess(eris)
    assert np.array_equal(gt2.cpu().numpy(), t2), "MP2 guess must be bit-identical"
    t1, t2 = _perturbed(t1, t2, 1)  # nonzero t1 exercises every t1 term
    rt1, rt2, (ts, ts_alt) = env[3].update_amps(t1, t2, oeris, s, iter(us), return_sparse=True)
    dt1, dt2 = cc.update_amps(torch.tensor(t1, device="cuda"), torch.tensor(t2, device="cuda"), eris)
    if ccs:
        assert np.array_equal(cc._slots[0].indices(), ts.idx)
    if ts_alt is not None:
        assert np.array_equal(cc._slots[1].indices(), ts_alt.idx)
    e2 = np.abs(dt2.cpu().numpy() - rt2).max() / np.abs(rt2).max()
    e1 = np.abs(dt1.cpu().numpy() - rt1).max() / np.abs(rt1).max()
    assert e2 < 1e-10 and e1 < 1e-10, (e1, e2)


def test_update_amps_pivotal_matches_oracle(env):
    """sampling_method="pivotal" (the binding's default, src/py_rccsd.cpp:30):
    both samples' index sets and values bit for bit against the oracle's
    canonical pivotal chain, the update within 1e-10."""
    torch = env[0]
    settings = {"m_keep": 700, "compression": "fri", "sampling_method": "pivotal",
                "compressed_contractions": ["O^2V^4", "O^3V^3", "O^4V^2"]}
    us = [0.42, 0.58]
    cc, eris, oeris, s = _pair(env, 4, 13, 30, 7, settings, us)
    _, t1, t2 = env[3].init_amps(oeris)
    t1, t2 = _perturbed(t1, t2, 2)
    rt1, rt2, (ts, ts_alt) = env[3].update_amps(t1, t2, oeris, s, iter(us), return_sparse=True)
    dt1, dt2 = cc.update_amps(torch.tensor(t1, device="cuda"), torch.tensor(t2, device="cuda"), eris)
    for slot, ref in ((cc._slots[0], ts), (cc._slots[1], ts_alt)):
        assert np.array_equal(slot.indices(), ref.idx)
        assert np.array_equal(slot.values().view(np.int64), ref.val.view(np.int64))
    assert len(set(cc._slots[0].indices().tolist())) == 700
    e2 = np.abs(dt2.cpu().numpy() - rt2).max() / np.abs(rt2).max()
    e1 = np.abs(dt1.cpu().numpy() - rt1).max() / np.abs(rt1).max()
    assert e2 < 1e-10 and e1 < 1e-10, (e1, e2)


def test_update_amps_rccsd_t1t1_mode(env):
    torch = env[0]
    settings = {"m_keep": 900, "compression": "largest", "vvvv_t1t1": "rccsd"}
    cc, eris, oeris, s = _pair(env, 3, 10, 25, 2, settings, [])
    _, t1, t2 = env[3].init_amps(oeris)
    t1, t2 = _perturbed(t1, t2, 2)
    rt1, rt2 = env[3].update_amps(t1, t2, oeris, s, iter([]))
    dt1, dt2 = cc.update_amps(torch.tensor(t1, device="cuda"), torch.tensor(t2, device="cuda"), eris)
    assert np.abs(dt2.cpu().numpy() - rt2).max() / np.abs(rt2).max() < 1e-10


@pytest.mark.parametrize("prec,tol,etol", [("fp32", 5e-6, 1e-6), ("mixed", 2e-5, 1e-5), ("bf16-ladder", 5e-3, 1e-3)])
def test_update_amps_precision_modes(env, prec, tol, etol):
    """Same T2, same uniforms: the precision modes stay within their stated
    relative error of the FP64 step (FP32: FP32 operands and accumulation;
    mixed: BF16 hi/lo MFMA for the sparse terms, about 16 significant bits
    per product, FP32 dense GEMMs; bf16-ladder: the ladder's products in BF16,
    8 significant bits, everything else FP32).  On this 6 x 24 system the few
    terms per element do not average the product errors out, so the bounds
    are loose; bench.py reports each mode's per-term and energy error at
    o=32, v=256."""
    torch = env[0]
    o, v, naux = 6, 24, 40
    us = [0.42, 0.58]
    base = {"m_keep": 6000, "compression": "fri", "sampling_method": "systematic"}
    cc64, eris, oeris, s = _pair(env, o, v, naux, 3, base, us)
    ccm, erism, _, _ = _pair(env, o, v, naux, 3, dict(base, precision=prec), us)
    t1, t2 = cc64.get_init_guess(eris)
    t1 = t1 + 0.01 * torch.randn_like(t1)
    a1, a2 = cc64.update_amps(t1, t2, eris)
    b1, b2 = ccm.update_amps(t1, t2, erism)
    rel = (b2 - a2).abs().max().item() / a2.abs().max().item()
    print(f"{prec}: t2 max rel error {rel:.3g}", flush=True)
    assert rel < tol, rel
    assert np.array_equal(cc64._slots[0].indices(), ccm._slots[0].indices())
    de = abs(cc64.energy(a1, a2, eris) - ccm.energy(b1, b2, erism))
    assert de < etol, de  # stated tolerance of each mode on this E_corr ~ -1 Ha system


def test_kernel_trajectory_matches_oracle(env):
    """Five FRI-CCSD iterations (level shift 2.0): energies agree with the oracle
    loop on the same uniforms.  Compressions after the first act on GPU and CPU
    t2 that differ in the last bits, so index sets may differ at a rounding
    boundary; energies are compared to 1e-9 relative."""
    torch, FRICCSD, SyntheticSystem, rccsd = env
    o, v, naux = 4, 14, 40
    us = list(np.random.default_rng(9).random(20))
    settings = {"m_keep": 1500, "compression": "fri", "sampling_method": "systematic", "level_shift": 2.0}
    cc, eris, oeris, s = _pair(env, o, v, naux, 8, settings, us)
    cc.max_cycle = 5
    cc.conv_tol = 0
    cc.conv_tol_normt = 0
    cc.kernel()
    _, _, _, _, en = rccsd.kernel(oeris, s, iter(us), max_cycle=5, tol=0, tolnormt=0)
    assert len(cc.energies) == 5
    np.testing.assert_allclose(cc.energies, en, rtol=1e-9)


def test_unbiased_branch_matches_oracle(env):
    """The unbiased estimator (fricc/friccsd.py:108-136): after the burn-in,
    t2_sparse_alt samples the running average of t2 instead of t2.  Seven
    iterations with burn_in = 2 (so the average runs over five, past the
    it == 0 and it == 1 special cases) agree with the oracle loop on the same
    uniforms: energies to 1e-9 relative, final t2 likewise."""
    torch, FRICCSD, SyntheticSystem, rccsd = env
    o, v, naux = 4, 14, 40
    us = list(np.random.default_rng(13).random(40))
    settings = {"m_keep": 1500, "compression": "fri", "sampling_method": "systematic", "level_shift": 2.0,
                "unbiased": True, "burn_in": 2}
    cc, eris, oeris, s = _pair(env, o, v, naux, 8, settings, us)
    s.update(unbiased=True, burn_in=2)
    cc.max_cycle = 7
    cc.conv_tol = cc.conv_tol_normt = 0
    cc.kernel()
    _, _, _, ot2, en = rccsd.kernel(oeris, s, iter(us), max_cycle=7, tol=0, tolnormt=0)
    assert len(cc.energies) == 7
    np.testing.assert_allclose(cc.energies, en, rtol=1e-9)
    gt2 = cc.t2.cpu().numpy()
    assert np.abs(gt2 - ot2).max() <= 1e-9 * np.abs(ot2).max()
    # the averaged samples were taken: the branch ran past the burn-in
    assert getattr(cc, "_t2_avg", None) is not None


def test_nan_stops_the_loop(env):
    torch, FRICCSD, SyntheticSystem, rccsd = env
    cc = FRICCSD(SyntheticSystem(3, 8, 20, 1), {"m_keep": 300, "compressed_contractions": []},
                 uniforms=iter(np.random.default_rng(0).random(100)))
    eris = cc.ao2mo()
    t1, t2 = cc.get_init_guess(eris)
    t2 = t2.clone()
    t2[0, 0, 0, 0] = float("nan")
    cc.max_cycle = 3
    with pytest.raises(Exception):
        # compression rejects a non-finite amplitude vector (the reference would
        # propagate NaN into the energy and stop at fricc/friccsd.py:444-447)
        cc.kernel(t1=t1, t2=t2)


def test_bad_settings(env):
    torch, FRICCSD, SyntheticSystem, rccsd = env
    sysm = SyntheticSystem(2, 4, 8, 1)
    with pytest.raises(ValueError):
        FRICCSD(sysm, {"compressed_contractions": ["O^5V^5"]})
    cc = FRICCSD(sysm, {"m_keep": 0})
    eris = cc.ao2mo()
    t1, t2 = cc.get_init_guess(eris)
    with pytest.raises(ValueError):
        cc.update_amps(t1, t2, eris)
    cc = FRICCSD(sysm, {"m_keep": 10 ** 9})
    with pytest.raises(ValueError):
        cc.update_amps(t1, t2, eris)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_graph_replay_matches_eager(env, prec):
    """The step recorded once and replayed each iteration -- with host reads,
    copies and an unrelated eager step on the same device between replays --
    gives the eager trajectory exactly: energies, index sets and amplitudes,
    iteration by iteration."""
    torch, FRICCSD, SyntheticSystem, rccsd = env
    from fricc.friccsd import GraphStep

    o, v, naux = 4, 14, 40
    us = list(np.random.default_rng(4).random(40))
    base = {"m_keep": 1500, "compression": "fri", "sampling_method": "systematic", "precision": prec}
    other = FRICCSD(SyntheticSystem(3, 9, 20, 2), {"m_keep": 200}, uniforms=iter(np.linspace(0.1, 0.9, 50)))
    oe = other.ao2mo()
    ot1, ot2 = other.get_init_guess(oe)
    runs = {}
    for graph in (False, True):
        cc = FRICCSD(SyntheticSystem(o, v, naux, 8), base, uniforms=iter(us))
        cc.level_shift = 2.0
        eris = cc.ao2mo()
        t1, t2 = cc.get_init_guess(eris)
        runner = GraphStep(cc, eris, t1, t2) if graph else None
        energies, idx = [], []
        for it in range(5):
            cc._iter = it
            t1, t2 = runner.step(t1, t2) if graph else cc.update_amps(t1, t2, eris)
            energies.append(cc._last_step[2])
            idx.append(cc._slots[0].indices().copy())
            other.update_amps(ot1, ot2, oe)  # eager work between replays
        runs[graph] = (energies, idx, t2.cpu().numpy().copy())
    assert runs[True][0] == runs[False][0]
    for a, b in zip(runs[True][1], runs[False][1]):
        assert np.array_equal(a, b)
    assert np.array_equal(runs[True][2], runs[False][2])


def test_graph_mode_kernel_loop(env):
    """FRICCSD(..., {"graph": True}).kernel() runs the recorded step."""
    torch, FRICCSD, SyntheticSystem, rccsd = env
    us = list(np.random.default_rng(5).random(40))
    res = {}
    for graph in (False, True):
        cc = FRICCSD(SyntheticSystem(4, 12, 30, 3), {"m_keep": 900, "graph": graph}, uniforms=iter(us))
        cc.level_shift = 2.0
        cc.max_cycle = 4
        cc.conv_tol = cc.conv_tol_normt = 0
        cc.kernel()
        res[graph] = list(cc.energies)
    assert res[True] == res[False] and len(res[True]) == 4


def test_mixed_step_reuses_the_finalize_fp32_copy(env):
    """In the reduced-precision modes the finalize also writes t2new rounded to
    FP32, and the next step reads that copy when it is handed back the same,
    unmodified tensor.  A second step from the returned T2 (copy reused) must
    match a second step from a clone of it (converted afresh), and a T2
    changed in place after the first step must not be served the stale copy:
    it must match a step from an equal, freshly built tensor.  The tolerance
    (1e-6 relative) is far below the change a stale copy would make (about
    1e-2 relative here)."""
    from fricc import friccsd

    torch = env[0]
    o, v, naux = 6, 24, 40
    us = [0.42, 0.58, 0.31, 0.77]
    base = {"m_keep": 6000, "compression": "fri", "sampling_method": "systematic", "precision": "mixed"}
    runs = [_pair(env, o, v, naux, 3, base, us) for _ in range(4)]
    t1, t2 = runs[0][0].get_init_guess(runs[0][1])
    t1 = t1 + 0.01 * torch.randn_like(t1)
    firsts = [cc.update_amps(t1, t2, eris) for cc, eris, _, _ in runs]
    for cc, (_, t2n) in zip(runs, firsts):
        assert friccsd._t2_reduced(cc[0], t2n, torch.float32) is cc[0]._t2_low[2]
        assert torch.equal(cc[0]._t2_low[2], t2n.float())
    g = torch.Generator(device=t2.device).manual_seed(5)
    bump = 0.01 * torch.randn(t2.shape, generator=g, dtype=t2.dtype, device=t2.device) * t2.abs().max()
    (c0, e0, _, _), (c1, e1, _, _), (c2, e2, _, _), (c3, e3, _, _) = runs
    reused = c0.update_amps(firsts[0][0], firsts[0][1], e0)[1]
    fresh = c1.update_amps(firsts[1][0], firsts[1][1].clone(), e1)[1]
    t2x = firsts[2][1]
    t2x.add_(bump)  # in place: the version counter moves, the FP32 copy is stale
    assert friccsd._t2_reduced(c2, t2x, torch.float32) is not c2._t2_low[2]
    changed = c2.update_amps(firsts[2][0], t2x, e2)[1]
    rebuilt = c3.update_amps(firsts[3][0], firsts[3][1].clone() + bump, e3)[1]
    scale = reused.abs().max().item()
    assert (reused - fresh).abs().max().item() <= 1e-6 * scale
    assert (changed - rebuilt).abs().max().item() <= 1e-6 * scale
    assert (changed - reused).abs().max().item() > 1e-4 * scale  # the bump is visible
