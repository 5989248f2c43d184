"""The fused T2 finalize (fricc_t2_finalize) called through the C-ABI on
random inputs, against the same formula restated in FP64 on the CPU:

  t2new = (ovovT + acc + S + S[ji,ba] + t1[i,a] t1[j,b] wsum[a,b]) / (eia[i,a] + eia[j,b])
  E2    = sum (t2new + t1new[i,a] t1new[j,b]) * Lovov,  Lovov = 2 ovovT[ijab] - ovovT[ijba]
  dnorm = sum over ia >= jb of (t2new - t2old)^2

with Lovov passed in and with Lovov NULL (formed in the kernel), FP32 and FP64
S, with and without wsum, at sizes whose 32-wide tiles are ragged (v = 70, 33),
odd v (the 8-byte path) and a single occupied orbital, with and without the
FP32 copy of t2new (which must equal t2new rounded to FP32).  The whole update is
checked against the oracle in the update tests; this one pins the kernel's
quad and diagonal bookkeeping element by element.  Tolerance: 1e-12 relative
per element, 1e-11 relative on the sums (FP64 throughout; an FP32 S enters
exactly, widened to FP64)."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from fricc import _capi

    assert torch.cuda.is_available()
    return torch, _capi


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


@pytest.mark.parametrize("o,v", [(3, 70), (2, 33), (1, 64), (4, 96), (2, 5)])
@pytest.mark.parametrize("s_f32", [True, False])
@pytest.mark.parametrize("read_lovov", [True, False])
@pytest.mark.parametrize("with_wsum", [True, False])
@pytest.mark.parametrize("f32_copy", [True, False])
def test_finalize_matches_formula(env, o, v, s_f32, read_lovov, with_wsum, f32_copy):
    torch, _capi = env
    g = torch.Generator().manual_seed(o * 1000 + v)
    f64 = torch.float64

    def rn(*shape):
        return torch.randn(*shape, generator=g, dtype=f64)

    X, S, acc, t2o = rn(o, o, v, v), rn(o, o, v, v), rn(o, o, v, v), rn(o, o, v, v)
    if s_f32:
        S = S.float().double()  # values the FP32 buffer holds exactly
    t1, t1n = rn(o, v), rn(o, v)
    eia = rn(o, v) - 4.0
    W = rn(v, v) if with_wsum else None
    L = 2 * X - X.transpose(2, 3)

    ref = X + acc + S + S.permute(1, 0, 3, 2)
    if with_wsum:
        ref = ref + t1[:, None, :, None] * t1[None, :, None, :] * W
    ref = ref / (eia[:, None, :, None] + eia[None, :, None, :])
    e2 = ((ref + t1n[:, None, :, None] * t1n[None, :, None, :]) * L).sum()
    d = (ref - t2o).permute(0, 2, 1, 3).reshape(o * v, o * v)
    dn = (torch.tril(d) ** 2).sum()

    dev = "cuda"
    Xd, accd, t2od = X.to(dev), acc.to(dev), t2o.to(dev)
    Sd = (S.float() if s_f32 else S).to(dev)
    Ld = L.to(dev) if read_lovov else None
    Wd = W.to(dev) if with_wsum else None
    t1d, t1nd, eiad = t1.to(dev), t1n.to(dev), eia.to(dev)
    sums = torch.full((2,), float("nan"), dtype=f64, device=dev)
    o32 = torch.full((o, o, v, v), float("nan"), dtype=torch.float32, device=dev) if f32_copy else None
    _capi.check(_capi.lib().fricc_t2_finalize(
        o, v, _p(Xd), _p(accd), _p(Sd), 1 if s_f32 else 0, _p(t1d), _p(Wd), _p(eiad), _p(t2od), _p(Ld),
        _p(t1nd), _p(o32), _p(sums), _capi.stream_ptr()))
    torch.cuda.synchronize()
    got = accd.cpu()
    assert torch.all((got - ref).abs() <= 1e-12 * (ref.abs() + 1.0)), float((got - ref).abs().max())
    if f32_copy:  # the FP32 copy is the FP64 result rounded once
        assert torch.equal(o32.cpu(), got.float())
    s = sums.cpu()
    assert abs(float(s[0]) - float(e2)) <= 1e-11 * max(1.0, abs(float(e2))), (float(s[0]), float(e2))
    assert abs(float(s[1]) - float(dn)) <= 1e-11 * max(1.0, float(dn)), (float(s[1]), float(dn))


def test_finalize_lovov_null_is_bitwise_equal(env):
    """Lovov formed in the kernel equals Lovov read from the buffer, bit for
    bit: 2x - y is exact in the product either way."""
    torch, _capi = env
    o, v = 3, 64
    g = torch.Generator().manual_seed(7)
    X, S, acc, t2o = (torch.randn(o, o, v, v, generator=g, dtype=torch.float64).cuda() for _ in range(4))
    t1, t1n = (torch.randn(o, v, generator=g, dtype=torch.float64).cuda() for _ in range(2))
    eia = torch.randn(o, v, generator=g, dtype=torch.float64).cuda() - 4.0
    L = (2 * X - X.transpose(2, 3)).contiguous()
    outs = []
    for lov in (L, None):
        a = acc.clone()
        sums = torch.empty(2, dtype=torch.float64, device="cuda")
        _capi.check(_capi.lib().fricc_t2_finalize(
            o, v, _p(X), _p(a), _p(S), 0, _p(t1), None, _p(eia), _p(t2o), _p(lov), _p(t1n), None,
            _p(sums), _capi.stream_ptr()))
        outs.append((a, sums))
    torch.cuda.synchronize()
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])
