"""Pins of the NumPy restatement of FRI-LCCSD's update (oracle/lccsd.py),
without PySCF.

* One Jacobi step ("largest", m = N: every compressed contraction exact)
  equals that of an independent spin-orbital linearised CCSD
  (oracle/spinorb.update_linear), with and without off-diagonal Fock blocks,
  and in the LCCD variant; the linearised energies agree.
* A dropped term planted in the restatement fails that pin.
* Compressed and uncompressed branches agree at m = N for every subset.
* "Independent Compressions" consumes one uniform per compressed term, and
  with six equal uniforms reproduces the single-compression step.
"""
import numpy as np
import pytest

from oracle import lccsd, liboracle, rccsd, spinorb


def _system(o=3, v=5, naux=14, seed=7, offdiag=0.0):
    eris, L = rccsd.synthetic_eris(o, v, naux, seed)
    if offdiag:
        rng = np.random.default_rng(seed + 100)
        n = o + v
        f = eris.fock.copy()
        d = offdiag * rng.standard_normal((n, n))
        f += d + d.T
        eris.fock = f
        eris.mo_energy = f.diagonal().copy()
    nmo = o + v
    eri = liboracle.eri_block(L, (0, nmo), (0, nmo), (0, nmo), (0, nmo))
    return eris, eri


def _amps(eris, seed=3):
    o, v = eris.nocc, eris.ovov.shape[1]
    rng = np.random.default_rng(seed)
    _, t1, t2 = rccsd.init_amps(eris)
    t1 = t1 + 0.05 * rng.standard_normal((o, v))
    t2 = t2 + 0.01 * rng.standard_normal(t2.shape)
    return t1, 0.5 * (t2 + t2.transpose(1, 0, 3, 2))


@pytest.mark.parametrize("offdiag", [0.0, 0.02])
@pytest.mark.parametrize("lccd", [False, True])
def test_lccsd_step_equals_spin_orbital(offdiag, lccd):
    eris, eri = _system(offdiag=offdiag)
    o = eris.nocc
    t1, t2 = _amps(eris)
    s = lccsd.default_settings(m_keep=t2.size, compression="largest", LCCD=lccd)
    t1n, t2n = lccsd.update_amps(t1, t2, eris, s, iter([]))
    asym, f = spinorb.spin_integrals(eri, eris.fock, o)
    T1, T2 = spinorb.spin_amplitudes(t1, t2)
    T1n, T2n = spinorb.update_linear(T1, T2, asym, f, o, lccd=lccd)
    np.testing.assert_allclose(T1n[0::2, 0::2], t1n, rtol=0, atol=1e-13)
    np.testing.assert_allclose(T2n[0::2, 1::2, 0::2, 1::2], t2n, rtol=0, atol=1e-13)
    if lccd:
        assert not t1n.any()
    assert abs(lccsd.energy(t1, t2, eris) - spinorb.energy_linear(T1, T2, asym, f, o)) < 1e-13


def test_linear_update_differs_from_ccsd():
    """The linearised step is not the CCSD step (the pin distinguishes them)."""
    eris, eri = _system()
    o = eris.nocc
    t1, t2 = _amps(eris)
    asym, f = spinorb.spin_integrals(eri, eris.fock, o)
    T1, T2 = spinorb.spin_amplitudes(t1, t2)
    _, a = spinorb.update(T1, T2, asym, f, o)
    _, b = spinorb.update_linear(T1, T2, asym, f, o)
    assert np.abs(a - b).max() > 1e-4


@pytest.mark.parametrize("drop", ["1303", "0101", "2323"])
def test_pin_catches_a_dropped_term(drop):
    eris, eri = _system()
    o = eris.nocc
    t1, t2 = _amps(eris)
    s = lccsd.default_settings(m_keep=t2.size, compression="largest")
    orig = lccsd.contract
    try:
        lccsd.contract = lambda W, T, out, term: out if term == drop else orig(W, T, out, term)
        _, t2n = lccsd.update_amps(t1, t2, eris, s, iter([]))
    finally:
        lccsd.contract = orig
    asym, f = spinorb.spin_integrals(eri, eris.fock, o)
    T1, T2 = spinorb.spin_amplitudes(t1, t2)
    _, T2n = spinorb.update_linear(T1, T2, asym, f, o)
    assert np.abs(T2n[0::2, 1::2, 0::2, 1::2] - t2n).max() > 1e-6


@pytest.mark.parametrize("ccs", [["O^2V^4"], ["O^3V^3"], ["O^4V^2"], ["O^2V^4", "O^3V^3", "O^4V^2"]])
def test_largest_m_equals_n_is_dense(ccs):
    eris, _ = _system(o=3, v=6, naux=16, seed=2)
    t1, t2 = _amps(eris, seed=5)
    dense = lccsd.default_settings(m_keep=t2.size, compression="largest", compressed_contractions=[])
    sparse = lccsd.default_settings(m_keep=t2.size, compression="largest", compressed_contractions=ccs)
    a1, a2 = lccsd.update_amps(t1, t2, eris, dense, iter([]))
    b1, b2 = lccsd.update_amps(t1, t2, eris, sparse, iter([]))
    np.testing.assert_allclose(b1, a1, rtol=0, atol=1e-13)
    np.testing.assert_allclose(b2, a2, rtol=0, atol=1e-13)


def test_independent_compressions_draws():
    eris, _ = _system(o=3, v=6, naux=16, seed=2)
    t1, t2 = _amps(eris, seed=5)
    m = t2.size // 5
    ind = lccsd.default_settings(m_keep=m, **{"Independent Compressions": True})
    one = lccsd.default_settings(m_keep=m)
    us = iter([0.3] * 6 + [0.9])
    a1, a2, sp = lccsd.update_amps(t1, t2, eris, ind, us, return_sparse=True)
    assert next(us) == 0.9 and len(sp) == 6  # six uniforms consumed
    b1, b2 = lccsd.update_amps(t1, t2, eris, one, iter([0.3]))
    np.testing.assert_array_equal(a2, b2)
    np.testing.assert_array_equal(a1, b1)
    # different uniforms give different sampled sets for the six terms
    _, _, sp = lccsd.update_amps(t1, t2, eris, ind, iter(np.linspace(0.05, 0.95, 6)), return_sparse=True)
    assert len({tuple(s.idx.tolist()) for s in sp}) > 1


def test_lccsd_kernel_converges_to_the_linear_solution():
    """With "largest", m = N, the loop converges to the solution of the
    linear equations, which for LCCD can be solved directly:
    D t2 = ovov^T + L(t2)  ->  fixed point of the Jacobi map."""
    eris, _ = _system(o=2, v=4, naux=20, seed=3)
    s = lccsd.default_settings(m_keep=4 * 16, compression="largest", LCCD=True, level_shift=1.0)
    conv, e, t1, t2, _ = lccsd.kernel(eris, s, iter([]), max_cycle=400, tol=1e-13, tolnormt=1e-11)
    assert conv
    # the converged t2 is a fixed point of the unshifted map as well
    s0 = dict(s, level_shift=0.0)
    _, t2b = lccsd.update_amps(t1, t2, eris, s0, iter([]))
    np.testing.assert_allclose(t2b, t2, rtol=0, atol=1e-9)
    # and the solution of the dense linear system
    o, v = t1.shape
    n = o * o * v * v
    s_lin = dict(s0)
    b = lccsd.update_amps(np.zeros((o, v)), np.zeros((o, o, v, v)), eris, s_lin, iter([]))[1].ravel()
    cols = []
    for k in range(n):
        e_k = np.zeros(n)
        e_k[k] = 1.0
        cols.append(lccsd.update_amps(np.zeros((o, v)), e_k.reshape(o, o, v, v), eris, s_lin, iter([]))[1].ravel()
                    - b)
    J = np.array(cols).T  # t2 -> map(t2) - map(0)
    x = np.linalg.solve(np.eye(n) - J, b)
    np.testing.assert_allclose(x.reshape(o, o, v, v), t2, rtol=0, atol=1e-9)
    assert abs(e - lccsd.energy(t1, x.reshape(o, o, v, v), eris)) < 1e-10
