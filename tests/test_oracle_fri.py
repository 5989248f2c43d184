"""Pins of the CPU oracle's FRI compression (oracle/fricc_oracle.cpp).

The reference has no test that pins its sampling (fricc/tests/test_fri_utils.py
holds only __main__ scripts), so each function is pinned here by what the
mathematics fixes:
  * one_norm / interval ends: exact rational arithmetic (math.fsum, Fraction)
  * argsort top-m: numpy's stable argsort of -|x|
  * get_d_largest: a brute-force restatement of src/fri_utils.cpp:411-421 in
    exact arithmetic
  * sample_systematic: the closed form S_k = upper_bound(P, k+u) - 1 and the
    "not found -> 0" rule (src/fri_utils.cpp:132)
  * fri_compression: invariants (values x_j/p_j, distinct samples), m = N keeps
    x exactly, and unbiasedness E[Phi(x)] = x with error ~ n^-1/2
    (src/julia/parallel_pivotal.png; fricc/tests/test_fri_utils.py:85-200).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import liboracle as lo


def _vec(n, seed, spread=8.0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) * np.exp(-rng.uniform(0, spread, n))


def test_norm_canonical_is_correctly_rounded_exact_sum():
    for seed in range(5):
        x = _vec(20000, seed)
        assert lo.norm_canonical(x) == math.fsum(np.abs(x))


def test_prefix_canonical_is_correctly_rounded_exact_prefix():
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 1, 3000)
    p[::7] = 0.0
    P = lo.prefix_canonical(p)
    acc = Fraction(0)
    assert P[0] == 0.0
    for j in range(p.size):
        acc += Fraction(p[j])
        assert P[j + 1] == float(acc)


def test_largest_matches_stable_argsort():
    x = _vec(5000, 2)
    x[100:200] = x[50]  # ties broken by ascending index
    for m in (1, 17, 150, 5000):
        idx, val, info = lo.compress(x, m, lo.LARGEST)
        ref = np.argsort(-np.abs(x), kind="stable")[:m]
        assert np.array_equal(idx, ref)
        assert np.array_equal(val, x[ref])


def _d_brute(x, m):
    """src/fri_utils.cpp:400-437 in exact arithmetic with the canonical order."""
    y = [Fraction(abs(float(v))) for v in x]
    order = sorted(range(len(x)), key=lambda i: (-abs(x[i]), i))
    R = sum(y, Fraction(0))
    D = []
    for k in range(m):
        i = order[k]
        Rk = float(R)
        yi = abs(float(x[i]))
        if float(m - k) * yi >= Rk - yi and Rk > 1e-14:
            D.append(i)
            R -= y[i]
        else:
            break
    return D, float(R)


@pytest.mark.parametrize("n,m,seed", [(20, 5, 0), (50, 30, 1), (200, 60, 2), (1000, 300, 3), (12, 12, 4)])
def test_d_set_brute_force(n, m, seed):
    x = _vec(n, seed, spread=6.0)
    idx, val, info = lo.compress(x, m, lo.FRI_SYSTEMATIC, 0.5)
    D, R = _d_brute(x, m)
    assert info["n_D"] == len(D)
    assert list(idx[: len(D)]) == D
    assert info["remaining_norm"] == R


def _closed_form_S(x, m, u, info, idx):
    K = info["n_D"]
    R = info["remaining_norm"]
    ns = m - K
    p = np.abs(x) / R * float(ns)
    p[idx[:K]] = 0.0
    P = lo.prefix_canonical(p)
    uk = np.arange(ns, dtype=np.float64) * 1.0 + u
    S = np.searchsorted(P, uk, side="right") - 1
    S[(uk >= P[-1]) | (S < 0)] = 0  # outside every interval -> index 0
    return S, p


@pytest.mark.parametrize("u", [0.0, 1e-9, 0.3, 0.5, 0.999999, 1.0])
@pytest.mark.parametrize("n,m", [(1000, 100), (5000, 2000), (40000, 3000)])
def test_systematic_closed_form(n, m, u):
    x = _vec(n, n + m)
    idx, val, info = lo.compress(x, m, lo.FRI_SYSTEMATIC, u)
    assert info["status"] == 0
    S, p = _closed_form_S(x, m, u, info, idx)
    K = info["n_D"]
    assert np.array_equal(idx[K:], S)
    found = S != 0
    # every sample that landed carries x_j / p_j (src/fri_utils.cpp:502-504)
    assert np.array_equal(val[K:][found], x[S[found]] / p[S[found]])
    # with p <= 1 an index is picked at most once
    assert np.unique(S[found]).size == found.sum()
    assert np.all(p <= 1.0)
    if u == 1.0:
        assert info["n_notfound"] >= 1


def test_m_equals_n_keeps_everything():
    x = _vec(3000, 9)
    idx, val, info = lo.compress(x, x.size, lo.FRI_SYSTEMATIC, 0.4)
    dense = np.zeros_like(x)
    np.add.at(dense, idx, val)
    assert np.array_equal(dense, x)


def test_unbiased_and_error_decays_as_inverse_sqrt():
    rng = np.random.default_rng(5)
    x = rng.random(400) + rng.random(400) * -1.0
    m = 60
    errs = []
    acc = np.zeros_like(x)
    for it in range(1, 4097):
        idx, val, info = lo.compress(x, m, lo.FRI_SYSTEMATIC, rng.random())
        phi = np.zeros_like(x)
        np.add.at(phi, idx, val)
        acc += phi
        if it in (64, 256, 1024, 4096):
            errs.append(np.abs(acc / it - x).sum() / np.abs(x).sum())
    errs = np.array(errs)
    slope = np.polyfit(np.log([64, 256, 1024, 4096]), np.log(errs), 1)[0]
    assert errs[-1] < 0.05
    assert -0.75 < slope < -0.3


def test_faithful_mode_diverges_rarely():
    """The reference's sequential sums (faithful mode) and the canonical exact
    sums pick the same indices except at rare rounding boundaries."""
    rng = np.random.default_rng(11)
    x = rng.standard_normal(200_000) * np.exp(-rng.uniform(0, 12, 200_000))
    diffs = 0
    for u in rng.random(20):
        a, _, ia = lo.compress(x, 20_000, lo.FRI_SYSTEMATIC, u, lo.CANONICAL)
        b, _, ib = lo.compress(x, 20_000, lo.FRI_SYSTEMATIC, u, lo.FAITHFUL)
        diffs += int((a != b).sum())
        assert ia["n_D"] == ib["n_D"]
    assert diffs <= 20


def test_bad_arguments_reported_not_exit():
    x = _vec(100, 1)
    assert lo.compress(x, 0, lo.FRI_SYSTEMATIC, 0.5)[2]["status"] == 2
    assert lo.compress(x, 101, lo.FRI_SYSTEMATIC, 0.5)[2]["status"] == 2
    y = x.copy()
    y[3] = np.nan
    assert lo.compress(y, 10, lo.FRI_SYSTEMATIC, 0.5)[2]["status"] == 3


def test_planted_mistake_is_caught():
    """Two plausible mistakes in the sampling step -- probabilities normalised by
    the 1-norm before D is removed, or D not excluded -- must each disagree
    with the oracle."""
    x = _vec(20000, 4)
    m = 3000
    u = 0.37
    idx, val, info = lo.compress(x, m, lo.FRI_SYSTEMATIC, u)
    K, ns = info["n_D"], m - info["n_D"]
    assert K > 0
    uk = np.arange(ns) * 1.0 + u
    for p in (np.abs(x) / info["norm0"] * float(ns), np.abs(x) / info["remaining_norm"] * float(ns)):
        if p is not None and np.all(p[idx[:K]] == 0):
            continue
        P = lo.prefix_canonical(np.minimum(p, 1.0))
        wrong = np.searchsorted(P, uk, side="right") - 1
        wrong[uk >= P[-1]] = 0
        assert not np.array_equal(idx[K:], wrong)


def test_m_equal_to_nonzeros_returns_d_alone():
    """m = nnz(x): the D loop takes every nonzero, n' = 0, nothing is sampled
    and no probability is used (the reference's p > 1 check passes on the NaN
    p of R = 0 and it returns D, src/fri_utils.cpp:383-395,494-504).  One more
    than the nonzeros leaves n' > 0 with R = 0, which is reported (quirk A.6)."""
    rng = np.random.default_rng(3)
    x = np.zeros(5000)
    nz = rng.choice(5000, 800, replace=False)
    x[nz] = rng.choice([1.0, -1.0, 0.5, 2.0 ** -30], 800)
    for mode in (0, 1):
        idx, val, info = lo.compress(x, 800, lo.FRI_SYSTEMATIC, 0.3, mode=mode)
        assert info["status"] == 0 and info["n_D"] == 800 and info["n_sampled"] == 0
        assert np.array_equal(np.sort(idx), np.sort(nz))
        assert np.array_equal(val, x[idx])
        assert lo.compress(x, 801, lo.FRI_SYSTEMATIC, 0.3, mode=mode)[2]["status"] == 1
