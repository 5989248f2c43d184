"""GPU FRI compression vs the CPU oracle (canonical mode): bit-exact indices
and values on the same input and uniform.

Covers random vectors from 10^3 to 1.3*10^7 entries, many uniforms including
u = 0 and u = 1 (the "not found -> index 0" path), exact ties, zeros, m = N,
m = 1, the "largest" mode, the c2-shaped MP2 guess, the free functions of the
binding and the error paths (p > 1, non-finite input)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import torch

    from fricc import _capi, py_rccsd
    from oracle import liboracle

    assert torch.cuda.is_available()
    return py_rccsd, liboracle, torch


def _vec(n, seed, spread=10.0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) * np.exp(-rng.uniform(0, spread, n))


def _gpu(py_rccsd, x, m, comp, u):
    T = py_rccsd.SparseTensor4d(x.reshape(1, 1, 1, -1), (1, 1, 1, x.size), m, comp, "systematic", u=u)
    return T.indices(), T.values(), T.info


def _check(mods, x, m, u, comp="fri"):
    py_rccsd, lo, _ = mods
    gi, gv, ginfo = _gpu(py_rccsd, x, m, comp, u)
    oi, ov, oinfo = lo.compress(x, m, lo.FRI_SYSTEMATIC if comp == "fri" else lo.LARGEST, u)
    assert oinfo["status"] == 0
    assert np.array_equal(gi, oi), f"indices differ at {np.flatnonzero(gi != oi)[:10]}"
    assert np.array_equal(gv.view(np.int64), ov.view(np.int64)), "values differ bitwise"
    if comp == "fri":
        assert ginfo["n_D"] == oinfo["n_D"]
        assert ginfo["remaining_norm"] == oinfo["remaining_norm"]
        assert ginfo["norm0"] == oinfo["norm0"]
        assert ginfo["n_notfound"] == oinfo["n_notfound"]
    return ginfo


@pytest.mark.parametrize("n,m", [(1000, 100), (1000, 999), (5000, 5000), (40_000, 4_000), (300_000, 20_000),
                                 (1_000_000, 250_000)])
def test_bitexact_random(mods, n, m):
    x = _vec(n, n + m)
    rng = np.random.default_rng(n)
    for u in [0.0, 1.0, 0.5] + list(rng.random(7)):
        _check(mods, x, m, u)


def test_bitexact_many_uniforms(mods):
    x = _vec(20_000, 77)
    rng = np.random.default_rng(1)
    us = np.concatenate([[0.0, 1.0, 1e-300, 1 - 2 ** -53], rng.random(100)])
    for u in us:
        _check(mods, x, 1500, float(u))


@pytest.mark.slow
def test_bitexact_13M(mods):
    x = _vec(13_000_000, 5)
    for u in (0.0, 0.61, 1.0):
        _check(mods, x, 1_000_000, u)


def test_ties_and_zeros(mods):
    rng = np.random.default_rng(3)
    x = np.zeros(50_000)
    x[rng.choice(50_000, 8000, replace=False)] = rng.choice([1.0, -1.0, 0.5, 0.25, 2.0 ** -30], 8000)
    for m in (10, 3000, 8000):
        for u in (0.0, 0.3, 1.0):
            _check(mods, x, m, u)
    for m in (10, 3000, 8000, 9000, 50_000):
        _check(mods, x, m, 0.0, comp="largest")
    # m beyond the nonzeros: the remaining norm reaches 0 and p is undefined
    # (quirk A.6); the reference emits garbage, here both paths report it
    py_rccsd, lo, _ = mods
    for m in (9000, 50_000):
        assert lo.compress(x, m, lo.FRI_SYSTEMATIC, 0.3)[2]["status"] == 1
        with pytest.raises(RuntimeError):
            _gpu(py_rccsd, x, m, "fri", 0.3)


def test_largest_mode(mods):
    x = _vec(123_457, 9)
    x[1000:2000] = x[5]
    for m in (1, 77, 4096, 4097, 123_457):
        _check(mods, x, m, 0.0, comp="largest")


def test_m_equals_n_keeps_vector(mods):
    py_rccsd, lo, _ = mods
    x = _vec(70_000, 12)
    gi, gv, info = _gpu(py_rccsd, x, x.size, "fri", 0.4)
    d = np.zeros_like(x)
    np.add.at(d, gi, gv)
    assert np.array_equal(d, x)


def test_mp2_guess_tensor(mods):
    """A real MP2 T2 (exactly pair-tied: t_ijab == t_jiba bitwise)."""
    from fricc.systems import SyntheticSystem
    from fricc.friccsd import init_amps
    from fricc.systems import DeviceERIs

    sysm = SyntheticSystem(8, 48, 96, 1)
    eris = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc, build_ladder=False)
    _, _, t2 = init_amps(eris)
    x = t2.cpu().numpy().ravel()
    assert np.array_equal(t2.cpu().numpy(), t2.cpu().numpy().transpose(1, 0, 3, 2))
    for u in (0.0, 0.123, 0.999, 1.0):
        _check(mods, x, x.size // 12, u)


def test_free_functions(mods):
    py_rccsd, lo, _ = mods
    x = _vec(10_000, 21)
    idx, val = py_rccsd.fri_compression(x, 900, "systematic", u=0.2)
    oi, ov, info = lo.compress(x, 900, lo.FRI_SYSTEMATIC, 0.2)
    assert idx == oi.tolist() and val == ov.tolist()
    assert py_rccsd.argsort(x, 50) == np.argsort(-np.abs(x), kind="stable")[:50].tolist()
    D, R = py_rccsd.get_d_largest(x, 900)
    assert D == oi[: info["n_D"]].tolist() and R == info["remaining_norm"]
    p = py_rccsd.make_probability_vector(x, 900 - len(D), D, R)
    p_ref = np.abs(x) / R * float(900 - len(D))
    p_ref[D] = 0
    assert np.array_equal(np.array(p), p_ref)
    S = py_rccsd.sample_systematic(900 - len(D), np.array(p), u=0.2)
    assert S == oi[info["n_D"]:].tolist()
    idx, val = py_rccsd.fri_compression(x, 900, u=0.2)  # the binding's default: pivotal
    oi, ov, info = lo.compress(x, 900, lo.FRI_PIVOTAL, 0.2)
    assert idx == oi.tolist() and val == ov.tolist()
    S = py_rccsd.sample_pivotal(900 - len(D), np.array(p), u=0.2)
    assert S == oi[info["n_D"]:].tolist()


def test_uniform_drawn_from_libc_rand(mods):
    """Without an explicit u, one rand()/RAND_MAX is consumed per FRI call."""
    py_rccsd, lo, _ = mods
    from fricc.uniforms import GlibcUniforms

    x = _vec(5000, 2)
    GlibcUniforms(42)
    T = py_rccsd.SparseTensor4d(x.reshape(1, 1, 50, 100), (1, 1, 50, 100), 500, "fri", "systematic")
    u_expected = next(iter(GlibcUniforms(42)))
    assert T.u == u_expected
    oi, _, _ = lo.compress(x, 500, lo.FRI_SYSTEMATIC, u_expected)
    assert np.array_equal(T.indices(), oi)


def test_error_paths(mods):
    py_rccsd, lo, _ = mods
    x = _vec(1000, 3)
    y = x.copy()
    y[17] = np.inf
    with pytest.raises(RuntimeError):
        _gpu(py_rccsd, y, 100, "fri", 0.5)
    # m larger than the nonzeros: remaining norm hits zero, p blows up (quirk A.6)
    z = np.zeros(1000)
    z[:5] = 1.0
    z[5] = 1e-300
    with pytest.raises((RuntimeError, ValueError)):
        _gpu(py_rccsd, z, 50, "fri", 0.5)
    with pytest.raises(ValueError):
        _gpu(py_rccsd, x, 1001, "fri", 0.5)


@pytest.mark.parametrize("comp", ["fri", "largest"])
def test_pair_equals_two_compressions(mods, comp):
    """recompress_pair_async (one selection, two samples) leaves each tensor
    bit for bit as its own recompress_async does, and as the oracle's
    compression with that uniform, including u = 1 (samples past the last
    interval take index 0) and u = 0."""
    py_rccsd, lo, torch = mods
    x = _vec(300_000, 5)
    m = 20_000
    xd = torch.from_numpy(x).cuda()
    dims = (1, 1, 1, x.size)
    method = lo.FRI_SYSTEMATIC if comp == "fri" else lo.LARGEST
    keys = ("n_D", "n_sampled", "n_notfound", "norm0", "remaining_norm", "p_total", "n_candidates")
    for u1, u2 in ((0.25, 0.75), (0.5, 1.0), (1.0, 0.0)):
        ud = torch.tensor([u1, u2], dtype=torch.float64, device="cuda")
        A, B = py_rccsd.SparseTensor4d.empty(dims), py_rccsd.SparseTensor4d.empty(dims)
        A.recompress_pair_async(B, xd, m, comp, ud[0:1], ud[1:2])
        infos = (A.sync_status(), B.sync_status())
        for T, info, k, u in ((A, infos[0], 0, u1), (B, infos[1], 1, u2)):
            S = py_rccsd.SparseTensor4d.empty(dims)
            S.recompress_async(xd, m, comp, ud[k:k + 1])
            sinfo = S.sync_status()
            assert np.array_equal(T.indices(), S.indices()), (u1, u2, k)
            assert np.array_equal(T.values().view(np.int64), S.values().view(np.int64)), (u1, u2, k)
            assert all(info[q] == sinfo[q] for q in keys), (info, sinfo)
            oi, ov, _ = lo.compress(x, m, method, u)
            assert np.array_equal(T.indices(), oi) and np.array_equal(T.values().view(np.int64), ov.view(np.int64))


def test_pair_edge_cases(mods):
    """The pair entry point when the kept part takes all m entries (no sample
    to emit) and when the second uniform lies outside [0, 1] (that tensor's
    status reports it; the first sample is unaffected)."""
    py_rccsd, lo, torch = mods
    rng = np.random.default_rng(4)
    x = np.zeros(50_000)
    x[rng.choice(x.size, 1000, replace=False)] = rng.standard_normal(1000)
    xd = torch.from_numpy(x).cuda()
    dims = (1, 1, 1, x.size)
    ud = torch.tensor([0.3, 0.6], dtype=torch.float64, device="cuda")
    A, B = py_rccsd.SparseTensor4d.empty(dims), py_rccsd.SparseTensor4d.empty(dims)
    A.recompress_pair_async(B, xd, 1000, "fri", ud[0:1], ud[1:2])
    for T, u in ((A, 0.3), (B, 0.6)):
        info = T.sync_status()
        assert info["n_sampled"] == 0 and info["n_D"] == 1000
        oi, ov, _ = lo.compress(x, 1000, lo.FRI_SYSTEMATIC, u)
        assert np.array_equal(T.indices(), oi) and np.array_equal(T.values().view(np.int64), ov.view(np.int64))
    x = _vec(100_000, 6)
    xd = torch.from_numpy(x).cuda()
    dims = (1, 1, 1, x.size)
    ud = torch.tensor([0.4, 1.5], dtype=torch.float64, device="cuda")
    A, B = py_rccsd.SparseTensor4d.empty(dims), py_rccsd.SparseTensor4d.empty(dims)
    A.recompress_pair_async(B, xd, 5000, "fri", ud[0:1], ud[1:2])
    A.sync_status()
    oi, ov, _ = lo.compress(x, 5000, lo.FRI_SYSTEMATIC, 0.4)
    assert np.array_equal(A.indices(), oi) and np.array_equal(A.values().view(np.int64), ov.view(np.int64))
    with pytest.raises((RuntimeError, ValueError)):
        B.sync_status()


@pytest.mark.parametrize("count", [6, 5, 3, 1])
def test_multi_equals_separate_compressions(mods, count):
    """compress_multi_async: up to six samples of one vector (uniforms in
    order, u = 1 among them; two samples per emission pass, an odd last one
    alone), each bit for bit its own compression and the oracle's."""
    py_rccsd, lo, torch = mods
    x = _vec(200_000, 9)
    m = 15_000
    xd = torch.from_numpy(x).cuda()
    dims = (1, 1, 1, x.size)
    us = [0.31, 1.0, 0.0, 0.77, 0.5, 0.123][:count]
    ud = torch.tensor(us, dtype=torch.float64, device="cuda")
    ts = [py_rccsd.SparseTensor4d.empty(dims) for _ in us]
    py_rccsd.SparseTensor4d.compress_multi_async(ts, xd, m, "fri", ud)
    for k, (T, u) in enumerate(zip(ts, us)):
        info = T.sync_status()
        S = py_rccsd.SparseTensor4d.empty(dims)
        S.recompress_async(xd, m, "fri", ud[k:k + 1])
        sinfo = S.sync_status()
        assert np.array_equal(T.indices(), S.indices()), k
        assert np.array_equal(T.values().view(np.int64), S.values().view(np.int64)), k
        assert info["n_notfound"] == sinfo["n_notfound"] and info["p_total"] == sinfo["p_total"]
        oi, ov, _ = lo.compress(x, m, lo.FRI_SYSTEMATIC, u)
        assert np.array_equal(T.indices(), oi) and np.array_equal(T.values().view(np.int64), ov.view(np.int64))
