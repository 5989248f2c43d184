"""fricc_compress_noalloc (scratch in a fixed device array, no allocation)
against fricc_compress (scratch from the library's workspace) through the C
ABI: indices, values and statistics bit for bit, on two calls in a row with
the second on a smaller input, for FRI systematic and pivotal sampling and
"largest"."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vec(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) * np.exp(-rng.uniform(0, 9, n))
    x[rng.choice(n, n // 50, replace=False)] = 0.0  # zeros and exact ties
    x[:64] = 0.125
    return x


def _run(fn, xd, m, method, u):
    import torch

    from fricc import _capi

    n = xd.numel()
    idx = torch.full((m,), 0xFFFFFFFF, dtype=torch.int64, device="cuda").to(torch.int32)
    val = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
    info = _capi.CompressInfo()
    st = fn(ctypes.c_void_p(xd.data_ptr()), n, m, method, u, ctypes.c_void_p(idx.data_ptr()),
            ctypes.c_void_p(val.data_ptr()), ctypes.byref(info), _capi.stream_ptr())
    _capi.check(st)
    return idx.cpu().numpy().view(np.uint32).copy(), val.cpu().numpy().copy(), info.as_dict()


@pytest.mark.parametrize("method", [1, 0, 2])  # FRICC_FRI_SYSTEMATIC, FRICC_LARGEST, FRICC_FRI_PIVOTAL
def test_noalloc_equals_compress_twice(method):
    import torch

    from fricc import _capi

    L = _capi.lib()
    assert torch.cuda.is_available()
    nmax = L.fricc_compress_noalloc_max_n()
    assert nmax == 1 << 22
    cases = [(nmax, 400_000, 0.6180339887498949, 11), (300_001, 20_000, 0.25, 12)]
    for n, m, u, seed in cases:  # the second call on a smaller input
        x = torch.tensor(_vec(n, seed), device="cuda")
        print(f"noalloc case n={n} m={m} method={method}", flush=True)
        a = _run(L.fricc_compress_noalloc, x, m, method, u)
        b = _run(L.fricc_compress, x, m, method, u)
        assert np.array_equal(a[0], b[0]), (n, m)
        assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64)), (n, m)
        assert a[2] == b[2]
        if method == 1:
            assert a[2]["n_D"] + a[2]["n_sampled"] == m


def test_noalloc_rejects_oversize_and_bad_arguments():
    import torch

    from fricc import _capi

    L = _capi.lib()
    n = L.fricc_compress_noalloc_max_n() + 1
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    idx = torch.zeros(8, dtype=torch.int32, device="cuda")
    val = torch.zeros(8, dtype=torch.float64, device="cuda")
    info = _capi.CompressInfo()
    st = L.fricc_compress_noalloc(ctypes.c_void_p(x.data_ptr()), n, 8, 1, 0.5, ctypes.c_void_p(idx.data_ptr()),
                                  ctypes.c_void_p(val.data_ptr()), ctypes.byref(info), _capi.stream_ptr())
    assert st == _capi.FRICC_EINVAL and b"fixed scratch" in L.fricc_last_error()
    st = L.fricc_compress_noalloc(ctypes.c_void_p(x.data_ptr()), 1000, 0, 1, 0.5, ctypes.c_void_p(idx.data_ptr()),
                                  ctypes.c_void_p(val.data_ptr()), ctypes.byref(info), _capi.stream_ptr())
    assert st == _capi.FRICC_EINVAL
    st = L.fricc_compress_noalloc(ctypes.c_void_p(x.data_ptr()), 1000, 8, 1, 1.5, ctypes.c_void_p(idx.data_ptr()),
                                  ctypes.c_void_p(val.data_ptr()), ctypes.byref(info), _capi.stream_ptr())
    assert st == _capi.FRICC_EINVAL


def test_noalloc_matches_oracle_small():
    """One small call also against the CPU oracle (canonical mode)."""
    import torch

    from fricc import _capi
    from oracle import liboracle

    x = _vec(50_000, 3)
    gi, gv, _ = _run(_capi.lib().fricc_compress_noalloc, torch.tensor(x, device="cuda"), 4000, 1, 0.37)
    oi, ov, _ = liboracle.compress(x, 4000, liboracle.FRI_SYSTEMATIC, 0.37)
    assert np.array_equal(gi.astype(np.int64), np.asarray(oi, dtype=np.int64))
    assert np.array_equal(gv.view(np.uint64), np.asarray(ov, dtype=np.float64).view(np.uint64))
