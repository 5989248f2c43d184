"""Integral provider: the reduced-precision ladder built slice by slice from
the density-fitting factors equals the one converted from the whole FP64
ladder element for element, and the column sums of the ladder taken from the
factors agree with the sums of the ladder."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eris_pair():
    import torch

    from fricc.systems import DeviceERIs, SyntheticSystem

    assert torch.cuda.is_available()
    out = {}
    for o, v in ((4, 20), (5, 24), (3, 13)):
        sysm = SyntheticSystem(o, v, 3 * (o + v), 2)
        full = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc, build_ladder=True)
        bare = DeviceERIs.from_factors(sysm.L, sysm.mo_energy, sysm.nocc, build_ladder=False)
        out[(o, v)] = (full, bare)
    return out


@pytest.mark.parametrize("ov", [(4, 20), (5, 24), (3, 13)])
@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_sliced_ladder_equals_whole(eris_pair, ov, prec):
    import torch

    full, bare = eris_pair[ov]
    v = ov[1]
    ref = full.sparse_operand("ladder", prec)
    # slices of 1, 2, 3 ... rows of c, a ragged last slice, and one slice
    for chunk in (8 * v ** 3, 8 * v ** 3 * 3, 8 * v ** 4):
        got = bare._reduced_ladder(prec, chunk_bytes=chunk)
        assert got.dtype == ref.dtype and got.shape == ref.shape
        assert torch.equal(got, ref), (ov, prec, chunk)
    assert bare.ladder is None


@pytest.mark.parametrize("ov", [(4, 20), (3, 13)])
def test_s_ab_from_factors(eris_pair, ov):
    import torch

    full, bare = eris_pair[ov]
    a = full._build("s_ab", torch.float64).cpu().numpy()
    b = bare._build("s_ab", torch.float64).cpu().numpy()
    assert np.allclose(a, a.T, rtol=0, atol=1e-12 * np.abs(a).max())
    assert np.allclose(b, a, rtol=0, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("prec", ["mixed", "fp32"])
def test_step_without_fp64_ladder(eris_pair, prec):
    """A reduced-precision step on a system whose FP64 ladder was never built
    matches the step on the system that has it (the operands are identical;
    only the ladder's column sums are summed in another order)."""
    import torch

    from fricc.friccsd import FRICCSD

    full, bare = eris_pair[(5, 24)]
    res = []
    for e in (full, bare):
        cc = FRICCSD(e, {"m_keep": 4000, "compression": "fri", "sampling_method": "systematic",
                         "precision": prec}, uniforms=iter([0.31, 0.77]))
        t1, t2 = cc.get_init_guess(e)
        a1, a2 = cc.update_amps(t1, t2, e)
        res.append((a2, cc.energy(a1, a2, e)))
    (a2, ea), (b2, eb) = res
    assert (a2 - b2).abs().max().item() <= 1e-12 * a2.abs().max().item()
    assert abs(ea - eb) < 1e-12


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_split_of_a_combination_equals_split_of_the_sum(dtype):
    """fricc_split_bf16_lincomb(A, B, 2, -1) gives the planes of 2A - B formed
    first (both products exact, one rounding of the sum), with a column count
    that is not a multiple of the 256-column record."""
    import torch

    from fricc.systems import bf16_planes, bf16_planes_lincomb

    g = torch.Generator(device="cuda").manual_seed(4)
    dt = getattr(torch, dtype)
    R, X = 37, 300
    A = torch.randn(R, X, generator=g, device="cuda", dtype=dt)
    B = torch.randn(R, X, generator=g, device="cuda", dtype=dt)
    got = bf16_planes_lincomb(A, B, 2.0, -1.0, R, X)
    want = bf16_planes(2.0 * A - B, R, X)
    torch.cuda.synchronize()
    assert torch.equal(got, want)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_three_splits_from_one_read(dtype):
    """fricc_split_bf16_lincomb3 gives, bit for bit, the planes of 2A - B, of
    A and of B that three separate splits give (ragged column count)."""
    import torch

    from fricc.systems import bf16_planes, bf16_planes_lincomb, bf16_planes_lincomb3

    g = torch.Generator(device="cuda").manual_seed(5)
    dt = getattr(torch, dtype)
    R, X = 41, 520
    A = torch.randn(R, X, generator=g, device="cuda", dtype=dt)
    B = torch.randn(R, X, generator=g, device="cuda", dtype=dt)
    ab, a, b = bf16_planes_lincomb3(A, B, 2.0, -1.0, R, X)
    torch.cuda.synchronize()
    assert torch.equal(ab, bf16_planes_lincomb(A, B, 2.0, -1.0, R, X))
    assert torch.equal(a, bf16_planes(A, R, X))
    assert torch.equal(b, bf16_planes(B, R, X))
