"""GPU pivotal sampling (sample_pivotal, src/fri_utils.cpp:63-127, and
fri_compression with "pivotal", :486-489) against the CPU oracle's canonical
chain (oracle/fricc_oracle.cpp, orc_pivotal, pinned by
tests/test_oracle_pivotal.py): identical indices, in emission order, and
values bit for bit, for explicit step uniforms and for the SplitMix64 stream
keyed by one uniform.  On dyadic probabilities the GPU also equals the
reference's own FP64 chain (the oracle's faithful mode).

Covers 2 .. 3*10^6 probabilities (one to hundreds of 4096-element tiles, stop
steps inside a thread's range and at tile edges), p_0 = 1 starts, zeros, all
ones, n' = 0, the reference's exits (wrong sample count -> FRICC_ESAMPLE),
p outside [0, 1] (FRICC_EPROB), the FRI compression (bit for bit, distinct
indices, the pair entry point) and the allocation-free entry point."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import torch

    from fricc import _capi, py_rccsd
    from oracle import liboracle

    assert torch.cuda.is_available()
    return py_rccsd, liboracle, torch, _capi


def _probs(n, ns, seed, spread=6.0, zeros=0.1):
    """Probabilities summing to ns (up to rounding), none above 1."""
    rng = np.random.default_rng(seed)
    y = rng.random(n) * np.exp(-rng.uniform(0, spread if 2 * ns < n else 1.0, n)) + 1e-300
    if 2 * ns < n:
        y[rng.random(n) < zeros] = 0.0
    p = y / y.sum() * ns
    while p.max() > 1.0:  # cap and renormalise the rest
        big = p >= 1.0
        p[big] = 1.0
        rest = ~big
        p[rest] *= (ns - big.sum()) / p[rest].sum()
    return p


def _gpu_pivotal(mods, p, ns, U=None, u=None):
    py_rccsd, _, torch, _capi = mods
    return np.array(py_rccsd.sample_pivotal(ns, p, u=u, uniforms=U), dtype=np.int64)


@pytest.mark.parametrize("n,ns", [(2, 1), (3, 2), (17, 5), (4096, 300), (4097, 4000), (70_000, 9_000),
                                  (1_000_003, 120_000), (3_000_000, 2_500_000)])
def test_explicit_uniforms_bitexact(mods, n, ns):
    _, lo, _, _ = mods
    p = _probs(n, ns, n + ns)
    U = np.random.default_rng(n).random(n)
    print(f"pivotal n={n} ns={ns}", flush=True)
    S, cnt, st = lo.pivotal(p, ns, U=U)
    assert st == 0 and cnt == ns
    G = _gpu_pivotal(mods, p, ns, U=U)
    assert np.array_equal(G, S), np.flatnonzero(G != S)[:10]
    assert len(np.unique(G)) == ns


@pytest.mark.parametrize("n,ns,u", [(5000, 700, 0.0), (5000, 700, 1.0), (250_000, 40_000, 0.37),
                                    (2_000_000, 100_000, 0.9)])
def test_keyed_stream_bitexact(mods, n, ns, u):
    _, lo, _, _ = mods
    p = _probs(n, ns, 7 * n)
    S, cnt, st = lo.pivotal(p, ns, key=lo.key_of(u))
    assert st == 0
    G = _gpu_pivotal(mods, p, ns, u=u)
    assert np.array_equal(G, S)


@pytest.mark.parametrize("seed", range(3))
def test_dyadic_equals_reference_chain(mods, seed):
    """On dyadic p every FP64 running sum of the reference is exact: the GPU's
    canonical chain is then the reference's FP64 chain (faithful mode)."""
    _, lo, _, _ = mods
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 257, 20_000)
    r, j = int(k.sum()) % 256, 0
    while r:  # trim to an integer total (every running sum is then exact in FP64)
        d = min(r, int(k[j]))
        k[j] -= d
        r -= d
        j += 1
    p = k / 256.0
    ns = int(k.sum()) // 256
    U = rng.random(p.size)
    S, cnt, st = lo.pivotal(p, ns, U=U, mode=lo.FAITHFUL)
    assert st == 0
    assert np.array_equal(_gpu_pivotal(mods, p, ns, U=U), S)


def test_edge_chains(mods):
    _, lo, _, _ = mods
    cases = [
        (np.array([1.0, 1.0, 0.5, 0.5]), 3),
        (np.array([1.0, 0.25, 0.75]), 2),
        (np.ones(9), 9),
        (np.concatenate([np.ones(5000), np.zeros(3000), np.full(4000, 0.5)]), 7000),
        (np.concatenate([np.zeros(6000), np.full(8192, 0.25)]), 2048),
        (np.zeros(100), 0),
    ]
    for p, ns in cases:
        for U in (np.zeros(p.size), np.full(p.size, 0.999999), np.random.default_rng(ns).random(p.size)):
            S, cnt, st = lo.pivotal(p, ns, U=U)
            assert st == 0, (p.size, ns)
            assert np.array_equal(_gpu_pivotal(mods, p, ns, U=U), S), (p.size, ns)


def test_reference_exits_raise(mods):
    py_rccsd, lo, _, _capi = mods
    bad = [
        (np.array([0.5, 0.3]), 1, _capi.FRICC_ESAMPLE),   # mass short of n' at the last step
        (np.array([0.5, 0.5]), 0, _capi.FRICC_ESAMPLE),   # a unit completed with n' = 0
        (np.array([1.0]), 1, _capi.FRICC_ESAMPLE),        # one element: the loop never runs
        (np.full(9000, 0.5), 5000, _capi.FRICC_ESAMPLE),  # too many samples asked for
        (np.array([0.5, 1.5]), 2, _capi.FRICC_EPROB),
        (np.array([0.5, -0.5, 1.0]), 1, _capi.FRICC_EPROB),
    ]
    for p, ns, status in bad:
        with pytest.raises(_capi.FriccError) as e:
            py_rccsd.sample_pivotal(ns, p, uniforms=np.zeros(p.size))
        assert e.value.status == status, (p[:4], ns)
    assert py_rccsd.sample_pivotal(0, np.array([0.0])) == []
    assert py_rccsd.sample_pivotal(0, np.zeros(50), u=0.5) == []


def _vec(n, seed, spread=10.0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) * np.exp(-rng.uniform(0, spread, n))
    x[rng.choice(n, n // 40, replace=False)] = 0.0
    x[:50] = x[60]  # exact ties
    return x


@pytest.mark.parametrize("n,m,u", [(1000, 100, 0.2), (1000, 900, 0.5), (40_000, 4_000, 0.0),
                                   (300_000, 20_000, 0.999), (2_000_000, 250_000, 0.61), (12_960_000, 1_000_000,
                                                                                          0.13)])
def test_fri_pivotal_compression_bitexact(mods, n, m, u):
    py_rccsd, lo, _, _ = mods
    x = _vec(n, n + m)
    print(f"fri pivotal n={n} m={m}", flush=True)
    T = py_rccsd.SparseTensor4d(x.reshape(1, 1, 1, -1), (1, 1, 1, n), m, "fri", "pivotal", u=u)
    oi, ov, info = lo.compress(x, m, lo.FRI_PIVOTAL, u)
    assert info["status"] == 0
    assert np.array_equal(T.indices(), oi)
    assert np.array_equal(T.values().view(np.int64), ov.view(np.int64))
    assert T.info["n_D"] == info["n_D"] and T.info["remaining_norm"] == info["remaining_norm"]
    assert len(np.unique(T.indices())) == m


def test_pair_equals_two_pivotal_compressions(mods):
    py_rccsd, _, torch, _ = mods
    x = _vec(500_000, 3)
    xd = torch.tensor(x, device="cuda")
    m = 60_000
    dims = (1, 1, 1, x.size)
    ud = torch.tensor([0.25, 0.8], dtype=torch.float64, device="cuda")
    A, B = py_rccsd.SparseTensor4d.empty(dims), py_rccsd.SparseTensor4d.empty(dims)
    A.recompress_pair_async(B, xd, m, "fri", ud[0:1], ud[1:2], sampling_method="pivotal")
    A.sync_status()
    B.sync_status()
    for T, u in ((A, 0.25), (B, 0.8)):
        R = py_rccsd.SparseTensor4d(x.reshape(dims), dims, m, "fri", "pivotal", u=u)
        assert np.array_equal(T.indices(), R.indices())
        assert np.array_equal(T.values().view(np.int64), R.values().view(np.int64))


def test_noalloc_pivotal_matches_oracle(mods):
    _, lo, torch, _capi = mods
    x = _vec(600_000, 9)
    m = 50_000
    xd = torch.tensor(x, device="cuda")
    idx = torch.zeros(m, dtype=torch.int32, device="cuda")
    val = torch.zeros(m, dtype=torch.float64, device="cuda")
    info = _capi.CompressInfo()
    _capi.check(_capi.lib().fricc_compress_noalloc(ctypes.c_void_p(xd.data_ptr()), x.size, m, _capi.FRI_PIVOTAL,
                                                   0.45, ctypes.c_void_p(idx.data_ptr()),
                                                   ctypes.c_void_p(val.data_ptr()), ctypes.byref(info),
                                                   _capi.stream_ptr()))
    oi, ov, oinfo = lo.compress(x, m, lo.FRI_PIVOTAL, 0.45)
    assert np.array_equal(idx.cpu().numpy().astype(np.int64), oi)
    assert np.array_equal(val.cpu().numpy().view(np.int64), ov.view(np.int64))


# ---------------------------------------------------------------- parallel_sample
def _ps_gpu(mods, q, ns, T, U=None, u=None):
    py_rccsd = mods[0]
    return np.array(py_rccsd.parallel_sample(ns, q, n_chunks=T, u=u, uniforms=U), dtype=np.int64)


def _ps_check(mods, q, ns, T, U=None, u=None):
    _, lo, _, _capi = mods
    key = lo.key_of(u) if U is None else 0
    S, st = lo.parallel_sample(q, ns, T, U=U, key=key)
    if st == 0:
        G = _ps_gpu(mods, q, ns, T, U=U, u=u)
        assert np.array_equal(G, S), (q.size, ns, T, np.flatnonzero(G != S)[:10])
    else:
        want = {1: _capi.FRICC_EPROB, 5: _capi.FRICC_ESAMPLE}[st]
        with pytest.raises(_capi.FriccError) as e:
            _ps_gpu(mods, q, ns, T, U=U, u=u)
        assert e.value.status == want
    return st


@pytest.mark.parametrize("n,ns,T", [(100, 20, 3), (5000, 700, 7), (4096 * 3 + 5, 3000, 2), (70_000, 9_000, 64),
                                    (1_000_003, 120_000, 256), (3_000_000, 400_000, 7), (200_000, 30_000, 4096)])
def test_parallel_sample_bitexact(mods, n, ns, T):
    p = _probs(n, ns, 3 * n + T)
    print(f"parallel_sample n={n} ns={ns} T={T}", flush=True)
    assert _ps_check(mods, p, ns, T, U=np.random.default_rng(n).random(n)) == 0
    assert _ps_check(mods, p, ns, T, u=0.77) == 0


def test_parallel_sample_edges(mods):
    py_rccsd, lo, _, _capi = mods
    rng = np.random.default_rng(4)
    p = _probs(3000, 500, 9)
    U = rng.random(p.size)
    # one chunk (its expected count exactly n' here): the plain chain
    assert lo.parallel_sample(p, 500, 1, U=U)[1] == 0
    assert _ps_gpu(mods, p, 500, 1, U=U).tolist() == py_rccsd.sample_pivotal(500, p, uniforms=U)
    # integer chunk sums: no extra samples, no move
    q = np.zeros(8192)
    q[:4096] = 0.25
    q[4096:] = 0.125
    assert _ps_check(mods, q, 1536, 2, U=U[:1].repeat(8192)) == 0
    # one element per chunk (a chunk asked for one sample exits, as in the reference)
    _ps_check(mods, np.full(64, 0.5), 32, 64, U=np.zeros(64))
    # no mass
    assert py_rccsd.parallel_sample(0, np.zeros(100), n_chunks=4, u=0.5) == []
    with pytest.raises(_capi.FriccError) as e:
        py_rccsd.parallel_sample(3, np.zeros(100), n_chunks=4, u=0.5)
    assert e.value.status == _capi.FRICC_ESAMPLE
    with pytest.raises(_capi.FriccError) as e:
        py_rccsd.parallel_sample(1, np.array([0.5, 1.5]), n_chunks=1, u=0.5)
    assert e.value.status == _capi.FRICC_EPROB
    for T in (0, 101):
        with pytest.raises(ValueError):
            py_rccsd.parallel_sample(3, np.full(100, 0.03), n_chunks=T, u=0.5)
