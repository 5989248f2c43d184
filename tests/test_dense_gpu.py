"""The dense MFMA contraction (fricc_dense_contract) against plain PyTorch
einsum in FP64 on the CPU: permuted views, broadcasts (stride 0), ragged
sizes, split-K shapes (tiny output, long sum), mixed FP32/FP64 operands,
beta accumulation (onto C or from a separate source), every tile shape, and
the products of the amplitude update at small sizes.  Tolerance per element:
FP64 compute 1e-12 and FP32 compute 2e-6 of sum_k |a||b| (the magnitude the
sum's rounding scales with)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    from fricc import dense

    assert torch.cuda.is_available()
    return torch, dense


def _ref(torch, spec, A, B):
    a, b = A.detach().cpu().double(), B.detach().cpu().double()
    return torch.einsum(spec, a, b), torch.einsum(spec, a.abs(), b.abs())


def _check(torch, dense, spec, A, B, compute, beta=0.0, out=None):
    ref, mag = _ref(torch, spec, A, B)
    C0 = None
    if out is not None:
        C0 = out.detach().cpu().double().clone()
    got = dense.einsum(spec, A, B, out=out, alpha=0.75, beta=beta, compute=compute)
    want = 0.75 * ref + (beta * C0 if beta else 0.0)
    tol = (1e-12 if compute == torch.float64 else 2e-6) * (0.75 * mag + (abs(beta) * C0.abs() if beta else 0.0)) + 1e-300
    err = (got.detach().cpu().double() - want).abs()
    assert torch.all(err <= tol), (spec, float((err / tol).max()))


SPECS = [
    # (spec, A shape, B shape)
    ("mk,kn->mn", (130, 70), (70, 33)),        # ragged edges
    ("mk,kn->mn", (5, 20000), (20000, 7)),     # split K
    ("bmk,bkn->bmn", (9, 40, 17), (9, 17, 65)),  # batch
    ("klcd,ilcd->ki", (6, 5, 11, 13), (4, 5, 11, 13)),
    ("klad,kldc->ac", (5, 4, 19, 23), (5, 4, 23, 21)),
    ("kd,kcda->ac", (7, 9), (7, 12, 9, 10)),   # GEMV over a 4-index operand
    ("jc,icab->ijab", (6, 9), (5, 9, 10, 11)),
    ("iakj,kb->ijab", (5, 7, 6, 4), (6, 9)),   # three levels in m
    ("ka,kijb->ijab", (6, 8), (6, 5, 4, 9)),
    ("ijka,kb->ijab", (4, 5, 6, 7), (6, 8)),
    ("ac,ijcb->ijab", (9, 10), (4, 5, 10, 9)),
    ("ki,kjab->ijab", (6, 5), (6, 4, 7, 8)),
    ("lkic,jc->klij", (5, 6, 7, 9), (8, 9)),
    ("kcli,la->kcia", (4, 6, 5, 7), (5, 9)),
]


@pytest.mark.parametrize("spec,sa,sb", SPECS)
@pytest.mark.parametrize("compute", ["fp64", "fp32"])
def test_dense_matches_torch(env, spec, sa, sb, compute):
    torch, dense = env
    g = torch.Generator().manual_seed(hash(spec) % 1000)
    cdt = torch.float64 if compute == "fp64" else torch.float32
    A = torch.randn(sa, generator=g, dtype=torch.float64).cuda()
    B = torch.randn(sb, generator=g, dtype=torch.float64).cuda()
    _check(torch, dense, spec, A, B, cdt)
    if compute == "fp32":
        _check(torch, dense, spec, A.float(), B, cdt)  # mixed element types


@pytest.mark.parametrize("compute", ["fp64", "fp32"])
def test_dense_strided_views_and_broadcast(env, compute):
    torch, dense = env
    cdt = torch.float64 if compute == "fp64" else torch.float32
    g = torch.Generator().manual_seed(3)
    base = torch.randn(7, 9, 6, 5, generator=g, dtype=torch.float64).cuda()
    A = base.permute(2, 0, 3, 1)  # (6, 7, 5, 9), no unit stride in the last axis
    B = torch.randn(9, 11, generator=g, dtype=torch.float64).cuda()
    _check(torch, dense, "pqrs,st->pqrt", A, B, cdt)
    v = torch.randn(11, generator=g, dtype=torch.float64).cuda()
    Bb = v.expand(13, 11)  # stride 0 rows
    M = torch.randn(8, 13, generator=g, dtype=torch.float64).cuda()
    _check(torch, dense, "mk,kn->mn", M, Bb, cdt)
    out = torch.randn(8, 11, generator=g, dtype=cdt).cuda()
    _check(torch, dense, "mk,kn->mn", M, Bb, cdt, beta=-1.5, out=out)
    # output into a permuted view, accumulated
    big = torch.randn(11, 8, generator=g, dtype=torch.float64).cuda()
    _check(torch, dense, "mk,kn->mn", M, Bb, cdt, beta=1.0, out=big.t())


def test_dense_every_tile_shape_and_split(env):
    torch, dense = env
    g = torch.Generator().manual_seed(5)
    for m in (3, 40, 100, 300):
        for n in (2, 50, 120, 260):
            for k in (1, 33, 5000):
                A = torch.randn(m, k, generator=g, dtype=torch.float64).cuda()
                B = torch.randn(k, n, generator=g, dtype=torch.float64).cuda()
                _check(torch, dense, "mk,kn->mn", A, B, torch.float64)


@pytest.mark.parametrize("compute", ["fp64", "fp32"])
@pytest.mark.parametrize("spec,sa,sb,cshape", [
    ("id,kcda->kcia", (6, 9), (5, 7, 9, 12), (5, 7, 6, 12)),   # the W builds' shape: 16-byte epilogue
    ("mk,kn->mn", (5, 20000), (20000, 7), (5, 7)),            # split K: the reduce reads the source
    ("mk,kn->mn", (130, 70), (70, 33), (130, 33)),            # ragged: the scalar epilogue
])
def test_dense_beta_term_from_a_source(env, compute, spec, sa, sb, cshape):
    """out = alpha A.B + beta Csrc equals Csrc copied into out and then
    accumulated (bit for bit), and the einsum of the CPU reference; Csrc is
    left unchanged."""
    torch, dense = env
    cdt = torch.float64 if compute == "fp64" else torch.float32
    g = torch.Generator().manual_seed(23)
    A = torch.randn(*sa, generator=g, dtype=torch.float64).to(cdt).cuda()
    B = torch.randn(*sb, generator=g, dtype=torch.float64).to(cdt).cuda()
    src = torch.randn(*cshape, generator=g, dtype=torch.float64).to(cdt).cuda()
    keep = src.clone()
    via_copy = src.clone()
    dense.einsum(spec, A, B, out=via_copy, alpha=0.75, beta=1.0, compute=cdt)
    out = torch.full_like(src, float("nan"))
    dense.einsum(spec, A, B, out=out, alpha=0.75, beta=1.0, compute=cdt, c_src=src)
    assert torch.equal(out, via_copy)
    assert torch.equal(src, keep)
    _check(torch, dense, spec, A, B, cdt, beta=1.0, out=src.clone())
    with pytest.raises(ValueError):
        dense.einsum(spec, A, B, out=out, beta=1.0, compute=cdt, c_src=src.double() if compute == "fp32" else src.float())


def test_dense_rejects_bad_specs(env):
    torch, dense = env
    from fricc import _capi

    A = torch.zeros(3, 4, dtype=torch.float64, device="cuda")
    B = torch.zeros(4, 5, dtype=torch.float64, device="cuda")
    out = torch.zeros(3, 5, dtype=torch.float64, device="cuda")
    with pytest.raises(_capi.FriccValueError):
        dense.einsum("mk,kn->mq", A, B, out=out)
    with pytest.raises(_capi.FriccValueError):
        dense.einsum("mk,jn->mn", A, B, out=out)  # j only in B
    with pytest.raises(_capi.FriccValueError):
        dense.einsum("mk,kn->mn", A, torch.zeros(5, 5, dtype=torch.float64, device="cuda"), out=out)


def _vp(t):
    import ctypes

    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("o,v", [(32, 256), (20, 128), (1, 128), (8, 384), (7, 256)])
def test_dense_zpair_matches_torch(env, o, v, prec):
    """fricc_dense_zpair (both products of the o*v^3 operand, Z read once)
    against FP64 einsums on the CPU: out_ia += t2.Z (accumulated onto a given
    out_ia) and out_ac = t1.Z, FP32 or FP64 operands; tolerance per element
    2e-6 (FP32) or 1e-12 (FP64) of sum |a||b|, as for the dense contraction."""
    torch, _ = env
    from fricc import _capi

    dt = torch.float32 if prec == "fp32" else torch.float64
    tol = 2e-6 if prec == "fp32" else 1e-12
    g = torch.Generator().manual_seed(o * 1000 + v)
    t2 = (torch.randn(o, o, v, v, generator=g, dtype=torch.float64) * 0.1).to(dt)
    t1 = (torch.randn(o, v, generator=g, dtype=torch.float64) * 0.1).to(dt)
    Z = torch.randn(o, v, v, v, generator=g, dtype=torch.float64).to(dt)
    out_ia0 = torch.randn(o, v, generator=g, dtype=torch.float64)
    # device copies held in names: a temporary freed before the launch could
    # be reused by the next allocation
    t2g, t1g, Zg = t2.cuda(), t1.cuda(), Z.cuda()
    out_ia = out_ia0.cuda()
    out_ac = torch.full((v, v), float("nan"), dtype=torch.float64, device="cuda")
    pid = _capi.FP32 if prec == "fp32" else _capi.FP64
    _capi.check(_capi.lib().fricc_dense_zpair(o, v, pid, _vp(t2g), _vp(t1g), _vp(Zg), _vp(out_ia), _vp(out_ac),
                                              _capi.stream_ptr()))
    torch.cuda.synchronize()
    t2d, t1d, Zd = t2.double(), t1.double(), Z.double()
    want_ia = out_ia0 + torch.einsum("ikcd,kcda->ia", t2d, Zd)
    mag_ia = out_ia0.abs() + torch.einsum("ikcd,kcda->ia", t2d.abs(), Zd.abs())
    want_ac = torch.einsum("kd,kcda->ac", t1d, Zd)
    mag_ac = torch.einsum("kd,kcda->ac", t1d.abs(), Zd.abs())
    err_ia = (out_ia.cpu() - want_ia).abs()
    err_ac = (out_ac.cpu() - want_ac).abs()
    assert torch.all(err_ia <= tol * mag_ia), float((err_ia / mag_ia).max())
    assert torch.all(err_ac <= tol * mag_ac), float((err_ac / mag_ac).max())


def test_dense_zpair_rejects_unsupported_shapes(env):
    torch, _ = env
    from fricc import _capi

    buf = torch.zeros(1 << 16, dtype=torch.float64, device="cuda")
    for o, v, p in ((33, 128, 1), (4, 90, 1), (4, 0, 0), (0, 128, 0), (4, 128, 2)):
        st = _capi.lib().fricc_dense_zpair(o, v, p, _vp(buf), _vp(buf), _vp(buf), _vp(buf), _vp(buf),
                                           _capi.stream_ptr())
        assert st == _capi.FRICC_EINVAL, (o, v, p, st)
