"""Parity at the benchmark configurations' real sizes.

* c2 (nocc=20, nvir=180, naux=600, seed 1, m = 1e6): the MP2 amplitudes bit
  for bit; three FP64 FRI-CCSD iterations against the oracle's digests
  (tests/golden/c2.npz, written by tests/golden/make_c2_golden.py): identical
  kept index sets of both samples of every iteration (their values bit for
  bit at iteration 1, within 1e-10 after it), energies,
  norms, t1 and t2 (at 20000 fixed positions and three whole-tensor
  checksums) within 1e-10 relative; one FP32 and one mixed-precision
  iteration within their stated tolerances; compression of a later
  iteration's T2 bit for bit against the oracle run on the same T2 here.
* c3 (nocc=32, nvir=256, naux=864, m = 4e6): both samples of the MP2
  amplitudes (N = 67,108,864) bit for bit against the oracle's digests
  (tests/golden/c3_compress.json), and "largest".
"""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def c2():
    import torch

    from fricc.systems import DeviceERIs, SyntheticSystem

    assert torch.cuda.is_available()
    g = dict(np.load(os.path.join(GOLD, "c2.npz")))
    sysm = SyntheticSystem(int(g["nocc"]), int(g["nvir"]), int(g["naux"]), int(g["seed"]))
    eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc)
    yield g, sysm, eris
    del eris
    torch.cuda.empty_cache()


def _solver(g, eris, precision):
    from fricc.friccsd import FRICCSD
    from fricc.uniforms import FixedUniforms

    cc = FRICCSD(eris, {"m_keep": int(g["m_keep"]), "compression": "fri", "sampling_method": "systematic",
                        "precision": precision}, uniforms=FixedUniforms(list(g["uniforms"])))
    cc.level_shift = float(g["level_shift"])
    return cc


def _digests(g, t2):
    n = t2.numel()
    pos = np.sort(np.random.default_rng(int(g["pos_seed"])).choice(n, int(g["n_pos"]), replace=False))
    w = np.random.default_rng(int(g["weight_seed"])).standard_normal(n)
    flat = t2.reshape(-1).cpu().numpy()
    return flat[pos], flat.sum(), (flat * flat).sum(), (flat * w).sum()


def test_c2_uniforms_are_the_libc_stream(c2):
    """The golden uniforms are rand()/RAND_MAX after srand(1) (the reference's
    draws, src/fri_utils.cpp:47) as this machine's C library gives them."""
    from fricc.uniforms import GlibcUniforms

    g = c2[0]
    u = GlibcUniforms(int(g["seed"]))
    assert [next(u) for _ in range(len(g["uniforms"]))] == list(g["uniforms"])


def test_c2_mp2_guess_bit_identical(c2):
    g, _, eris = c2
    cc = _solver(g, eris, "fp64")
    _, t2 = cc.get_init_guess(eris)
    assert _sha(t2.cpu().numpy().astype("<f8")) == str(g["mp2_t2_sha"])
    assert abs(cc.emp2 - float(g["emp2"])) <= 1e-12 * abs(float(g["emp2"]))


def test_c2_fp64_trajectory_matches_oracle(c2):
    """Three FP64 iterations from the MP2 guess: both samples of every
    iteration bit for bit, amplitudes and energies within 1e-10 relative."""
    g, _, eris = c2
    cc = _solver(g, eris, "fp64")
    t1, t2 = cc.get_init_guess(eris)
    for k in range(len(g["E"])):
        cc._iter = k
        print(f"c2 fp64 iteration {k + 1}", flush=True)
        t1, t2 = cc.update_amps(t1, t2, eris)
        vpos = np.sort(np.random.default_rng(int(g["val_seed"])).choice(int(g["m_keep"]), int(g["n_val"]),
                                                                         replace=False))
        for s in (0, 1):
            sl = cc._slots[s]
            assert _sha(sl.indices().astype("<u4")) == str(g["idx_sha"][k][s]), (k, s)
            vals = sl.values()
            if k == 0:  # identical inputs (the MP2 guess): values bit for bit
                assert _sha(vals.astype("<f8")) == str(g["val_sha"][k][s]), (k, s)
            # later inputs carry the last-bit differences of the FP64 update
            ref = g["val_at"][k][s]
            assert np.abs(vals[vpos] - ref).max() <= 1e-10 * np.abs(ref).max(), (k, s)
            assert sl.info["n_D"] == int(g["nD"][k][s])
        e, normt = cc._last_step[2], cc._last_step[3]
        assert abs(e - g["E"][k]) <= 1e-10 * abs(g["E"][k]), (k, e, g["E"][k])
        assert abs(normt - g["normt"][k]) <= 1e-10 * g["normt"][k]
        ref_t1 = g["t1"][k]
        assert np.abs(t1.cpu().numpy() - ref_t1).max() <= 1e-10 * np.abs(ref_t1).max()
        at, s1, s2, sw = _digests(g, t2)
        assert np.abs(at - g["t2_at"][k]).max() <= 1e-10 * np.abs(g["t2_at"][k]).max()
        assert abs(s2 - g["t2_sumsq"][k]) <= 1e-10 * g["t2_sumsq"][k]
        scale = np.sqrt(g["t2_sumsq"][k] * t2.numel())  # |sum| <= sqrt(N * sumsq)
        assert abs(s1 - g["t2_sum"][k]) <= 1e-10 * scale
        assert abs(sw - g["t2_wsum"][k]) <= 1e-10 * scale


@pytest.mark.parametrize("prec,t2tol,etol", [("fp32", 5e-6, 1e-6), ("mixed", 2e-5, 1e-6)])
def test_c2_reduced_precision_step(c2, prec, t2tol, etol):
    """One iteration from the MP2 guess in a reduced-precision mode: the
    compressions (FP64 in every mode) keep the oracle's index sets; t2 within
    the mode's stated tolerance (relative to max |t2|), E_corr within etol Ha."""
    g, _, eris = c2
    cc = _solver(g, eris, prec)
    t1, t2 = cc.get_init_guess(eris)
    cc._iter = 0
    t1, t2 = cc.update_amps(t1, t2, eris)
    for s in (0, 1):
        assert _sha(cc._slots[s].indices().astype("<u4")) == str(g["idx_sha"][0][s])
    at = _digests(g, t2)[0]
    assert np.abs(at - g["t2_at"][0]).max() <= t2tol * np.abs(g["t2_at"][0]).max()
    assert abs(cc._last_step[2] - g["E"][0]) <= etol


def test_c2_later_iteration_compression_bit_exact(c2):
    """The T2 of the fifth FP64 iteration, compressed on the GPU and by the
    oracle (canonical mode) on the same bits with the same uniforms: identical
    kept indices and values, both samples."""
    import torch

    from fricc import py_rccsd
    from oracle import liboracle

    g, _, eris = c2
    cc = _solver(g, eris, "fp64")
    cc.uniforms = iter(np.random.default_rng(55).random(20))
    t1, t2 = cc.get_init_guess(eris)
    for k in range(5):
        cc._iter = k
        t1, t2 = cc.update_amps(t1, t2, eris)
    m = int(g["m_keep"])
    x = t2.reshape(-1).cpu().numpy()
    for u in (0.13, 0.87):
        ts = py_rccsd.SparseTensor4d(t2, tuple(t2.shape), m, "fri", "systematic", u=u)
        oi, ov, info = liboracle.compress(x, m, liboracle.FRI_SYSTEMATIC, u)
        assert info["status"] == 0
        assert np.array_equal(ts.indices(), oi)
        assert np.array_equal(ts.values().view(np.uint64), ov.view(np.uint64))
        assert ts.info["n_D"] == info["n_D"]
    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def c3_mp2():
    import torch

    from fricc.friccsd import init_amps
    from fricc.systems import DeviceERIs, SyntheticSystem

    g = json.load(open(os.path.join(GOLD, "c3_compress.json")))
    sysm = SyntheticSystem(g["nocc"], g["nvir"], g["naux"], g["seed"])
    eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc, build_ladder=False)
    _, _, t2 = init_amps(eris)
    del eris
    torch.cuda.empty_cache()
    yield g, t2
    del t2
    torch.cuda.empty_cache()


def test_c3_mp2_amplitudes_bit_identical(c3_mp2):
    g, t2 = c3_mp2
    assert t2.numel() == 67_108_864
    assert _sha(t2.cpu().numpy().astype("<f8")) == g["mp2_t2_sha"]


def test_c3_full_size_compression_bit_exact(c3_mp2):
    """Both FRI samples of the c3 MP2 amplitudes (the pair entry point, one
    selection and two emissions) and "largest", against the oracle's digests."""
    import torch

    from fricc import py_rccsd

    g, t2 = c3_mp2
    m = g["m_keep"]
    a = py_rccsd.SparseTensor4d.empty(tuple(t2.shape))
    b = py_rccsd.SparseTensor4d.empty(tuple(t2.shape))
    u = torch.tensor([s["u"] for s in g["samples"]], dtype=torch.float64, device="cuda")
    a.recompress_pair_async(b, t2, m, "fri", u[0:1], u[1:2])
    for sp, ref in zip((a, b), g["samples"]):
        info = sp.sync_status()
        assert _sha(sp.indices().astype("<u4")) == ref["idx_sha"]
        assert _sha(sp.values().astype("<f8")) == ref["val_sha"]
        assert info["n_D"] == ref["n_D"] and info["n_sampled"] == ref["n_sampled"]
        assert info["n_notfound"] == ref["n_notfound"]
        assert info["norm0"] == ref["norm0"] and info["remaining_norm"] == ref["remaining_norm"]
    lg = py_rccsd.SparseTensor4d(t2, tuple(t2.shape), m, "largest")
    assert _sha(lg.indices().astype("<u4")) == g["largest"]["idx_sha"]
    assert _sha(lg.values().astype("<f8")) == g["largest"]["val_sha"]
