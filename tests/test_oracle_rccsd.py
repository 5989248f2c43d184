"""Pins of the NumPy restatement of update_amps (oracle/rccsd.py), without PySCF.

* One Jacobi step equals that of an independent spin-orbital CCSD
  (oracle/spinorb.py, Stanton-Gauss form) -- pins the PySCF imd.* formulas,
  the energy and the t2 equation (vvvv_t1t1="rccsd" form).
* "largest" compression with m = N makes every compressed contraction exact,
  so the FRI update equals the dense update (all branches of friccsd.py:194-279).
* The reference's t1*t1 vvvv term (quirk A.1) differs from the tau form by the
  closed-form amount t_ia t_jb S_ab - sum_cd W_abcd t_ic t_jd.
* MP2 guess energy equals the closed-form MP2 correlation energy.
"""
import numpy as np
import pytest

from oracle import liboracle, rccsd, spinorb


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
def test_rccsd_step_equals_spin_orbital(offdiag):
    eris, eri = _system(offdiag=offdiag)
    o = eris.nocc
    t1, t2 = _amps(eris)
    s = rccsd.default_settings(m_keep=t2.size, compression="largest", vvvv_t1t1="rccsd")
    t1n, t2n = rccsd.update_amps(t1, t2, eris, s, iter([]))
    asym, f = spinorb.spin_integrals(eri, eris.fock, o)
    T1, T2 = spinorb.spin_amplitudes(t1, t2)
    T1n, T2n = spinorb.update(T1, T2, asym, f, o)
    np.testing.assert_allclose(T1n[0::2, 0::2], t1n, rtol=0, atol=1e-13)
    np.testing.assert_allclose(T2n[0::2, 1::2, 0::2, 1::2], t2n, rtol=0, atol=1e-13)
    assert abs(rccsd.energy(t1, t2, eris) - spinorb.energy(T1, T2, asym, f, o)) < 1e-13


def test_spin_orbital_pin_catches_planted_sign_error():
    """Flip one sign in the restatement (a copy of Lvv) and the pin must fail."""
    eris, eri = _system()
    o = eris.nocc
    t1, t2 = _amps(eris)
    s = rccsd.default_settings(m_keep=t2.size, compression="largest", vvvv_t1t1="rccsd")
    orig = rccsd.Lvv
    try:
        rccsd.Lvv = lambda t1_, t2_, e_: orig(t1_, t2_, e_) - 2 * 2 * np.einsum("kdac,kd->ac", e_.ovvv, t1_)
        t1n, t2n = rccsd.update_amps(t1, t2, eris, s, iter([]))
    finally:
        rccsd.Lvv = orig
    asym, f = spinorb.spin_integrals(eri, eris.fock, o)
    T1, T2 = spinorb.spin_amplitudes(t1, t2)
    _, T2n = spinorb.update(T1, T2, asym, f, o)
    assert np.abs(T2n[0::2, 1::2, 0::2, 1::2] - t2n).max() > 1e-6


@pytest.mark.parametrize("ccs", [["O^2V^4"], ["O^3V^3"], ["O^4V^2"], ["O^2V^4", "O^3V^3", "O^4V^2"]])
def test_largest_m_equals_n_is_dense(ccs):
    eris, _ = _system(o=3, v=6, naux=16, seed=2)
    t1, t2 = _amps(eris, seed=5)
    dense = rccsd.default_settings(m_keep=t2.size, compression="largest", compressed_contractions=[])
    sparse = rccsd.default_settings(m_keep=t2.size, compression="largest", compressed_contractions=ccs)
    a1, a2 = rccsd.update_amps(t1, t2, eris, dense, iter([]))
    b1, b2 = rccsd.update_amps(t1, t2, eris, sparse, iter([]))
    np.testing.assert_allclose(b1, a1, rtol=0, atol=1e-13)
    np.testing.assert_allclose(b2, a2, rtol=0, atol=1e-13)


def test_reference_t1t1_vvvv_term_closed_form():
    eris, _ = _system()
    t1, t2 = _amps(eris)
    ref = rccsd.default_settings(m_keep=t2.size, compression="largest", vvvv_t1t1="reference")
    tau = rccsd.default_settings(m_keep=t2.size, compression="largest", vvvv_t1t1="rccsd")
    _, a2 = rccsd.update_amps(t1, t2, eris, ref, iter([]))
    _, b2 = rccsd.update_amps(t1, t2, eris, tau, iter([]))
    W = rccsd.cc_Wvvvv(t1, t2, eris)
    o, v = t1.shape
    mo_e = eris.mo_energy
    eia = mo_e[:o, None] - mo_e[None, o:]
    d = np.einsum("abcd,ia,jb->ijab", W, t1, t1) - np.einsum("abcd,ic,jd->ijab", W, t1, t1)
    np.testing.assert_allclose(a2 - b2, d / (eia[:, None, :, None] + eia[None, :, None, :]), atol=1e-13)


def test_mp2_guess_energy():
    eris, _ = _system(o=4, v=7)
    emp2, t1, t2 = rccsd.init_amps(eris)
    o = eris.nocc
    e = eris.mo_energy
    ov = eris.ovov
    ref = 0.0
    for i in range(o):
        for j in range(o):
            for a in range(ov.shape[1]):
                for b in range(ov.shape[1]):
                    d = e[i] + e[j] - e[o + a] - e[o + b]
                    ref += ov[i, a, j, b] * (2 * ov[i, a, j, b] - ov[i, b, j, a]) / d
    assert abs(emp2 - ref) < 1e-13
    assert abs(rccsd.energy(t1, t2, eris) - emp2) < 1e-13


def test_eri_fma_chain_is_8fold_symmetric():
    L, _ = rccsd.synthetic_L(3, 5, 11, 4)
    e = liboracle.eri_block(L, (0, 8), (0, 8), (0, 8), (0, 8))
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.array_equal(e, e.transpose(perm))
    ref = np.einsum("Qpq,Qrs->pqrs", L, L)
    np.testing.assert_allclose(e, ref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("ranges", [((0, 13), (0, 13), (0, 13), (0, 13)), ((0, 4), (4, 13), (4, 13), (4, 13)),
                                    ((2, 7), (4, 11), (0, 3), (5, 12))])
def test_eri_packed_builder_equals_plain_loop(ranges):
    """The tiled builder (packed rows, 12-column vector tiles, ragged edges)
    gives every integral bit for bit as the plain per-element chain."""
    L, _ = rccsd.synthetic_L(4, 9, 37, 6)
    a = liboracle.eri_block(L, *ranges)
    b = liboracle.eri_block(L, *ranges, plain=True)
    assert a.view(np.uint64).tolist() == b.view(np.uint64).tolist()


def test_fri_kernel_runs_and_is_near_ccsd():
    """FRI trajectories fluctuate around the CCSD energy.  The synthetic recipe
    is strongly correlated, so plain Jacobi steps need a level shift (PySCF's
    cc.level_shift, fricc/friccsd.py:83) to converge; the shift changes the
    iteration, not its fixed point."""
    eris, _ = _system(o=3, v=8, naux=60, seed=1)
    dense = rccsd.default_settings(m_keep=1, compression="largest", compressed_contractions=[], level_shift=2.0)
    conv, ecc, *_ = rccsd.kernel(eris, dense, iter([]), max_cycle=100, tol=1e-10, tolnormt=1e-8)
    assert conv
    fri = rccsd.default_settings(m_keep=int(0.5 * 9 * 64), level_shift=2.0)
    _, efri, _, _, energies = rccsd.kernel(eris, fri, iter(np.random.default_rng(0).random(400)), max_cycle=120,
                                           tol=0, tolnormt=0)
    assert len(energies) == 120
    assert abs(np.mean(energies[40:]) - ecc) < 0.02 * abs(ecc)


def _unbiased_run(eris, t1, t2, burn_in, n_iter, us, planted=None):
    """The oracle's loop body with the unbiased branch on: returns the t2
    inputs of every iteration and the running average the branch kept."""
    s = rccsd.default_settings(m_keep=t2.size // 3, unbiased=True, burn_in=burn_in)
    state, xs, avgs = {}, [], []
    uit = iter(us)
    for istep in range(n_iter):
        state["iter"] = istep
        xs.append(t2.copy())
        t1, t2 = rccsd.update_amps(t1, t2, eris, s, uit, state=state)
        avgs.append(state.get("t2_avg"))
    return xs, avgs


def test_unbiased_average_is_mean_of_inputs_after_burn_in():
    """fricc/friccsd.py:117-136: after the burn-in the running average
    Y(it) = t2/it + (it-1)/it * Y(it-1), started at Y(0) = t2, equals the
    arithmetic mean of the t2 inputs of iterations burn_in+1 .. burn_in+it
    (the input at it == 0 drops out with weight 0 at it == 1); before the
    burn-in no average exists."""
    eris, _ = _system()
    t1, t2 = _amps(eris)
    burn_in, n_iter = 2, 7
    xs, avgs = _unbiased_run(eris, t1, t2, burn_in, n_iter, np.random.default_rng(5).random(2 * n_iter))
    assert avgs[0] is None and avgs[1] is None
    np.testing.assert_array_equal(avgs[burn_in], xs[burn_in])       # it == 0: t2 itself
    np.testing.assert_array_equal(avgs[burn_in + 1], xs[burn_in + 1])  # it == 1: 0 * Y + t2/1
    for istep in range(burn_in + 2, n_iter):
        it = istep - burn_in
        mean = np.mean(xs[burn_in + 1:istep + 1], axis=0)
        np.testing.assert_allclose(avgs[istep], mean, rtol=1e-13, atol=1e-16)
        # a plausible slip (weights it/(it+1) instead of (it-1)/it) is far off
        wrong = xs[istep] / (it + 1) + it / (it + 1) * avgs[istep - 1]
        assert np.abs(wrong - mean).max() > 1e-6 * np.abs(mean).max()


def test_unbiased_first_two_averaged_steps_equal_plain_steps():
    """At it == 0 and it == 1 the average equals the current t2 exactly, so
    the update with the unbiased branch on equals the plain update with the
    same uniforms."""
    eris, _ = _system()
    t1, t2 = _amps(eris)
    m = t2.size // 3
    us = [0.11, 0.62]
    for it in (0, 1):
        plain = rccsd.update_amps(t1, t2, eris, rccsd.default_settings(m_keep=m), iter(us))
        st = {"iter": 3 + it, "t2_avg": np.full_like(t2, 7.0)}
        ub = rccsd.update_amps(t1, t2, eris, rccsd.default_settings(m_keep=m, unbiased=True, burn_in=3), iter(us),
                               state=st)
        np.testing.assert_array_equal(plain[0], ub[0])
        np.testing.assert_array_equal(plain[1], ub[1])
