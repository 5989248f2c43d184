"""Pins of the CPU oracle's pivotal sampling (oracle/fricc_oracle.cpp,
orc_pivotal; reference src/fri_utils.cpp:63-127).

The reference holds no pivotal test vector (fricc/tests/test_fri_utils.py is a
__main__ script), so the oracle is pinned by what the mathematics fixes:
  * pivotal sampling is a fixed-size design with inclusion probabilities p
    (Deville and Tille, Biometrika 85, 1998): enumerating every decision path
    of small chains with exact rational weights gives P(k in S) = p_k exactly,
    each S of n' distinct indices;
  * on dyadic probabilities every FP64 running sum is exact, so the canonical
    (exact fixed-point mass) and faithful (the reference's FP64 running sum)
    modes must agree bit for bit on any uniforms;
  * the reference's exits (wrong sample count, :100-123) and the p_0 = 1 start
    it carries as a = 1 into step 1.
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import liboracle as lo


def _dyadic_p(n, seed, denom=16):
    """Probabilities k/denom in [0, 1] summing to an integer."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, denom + 1, n)
    r = int(k.sum()) % denom
    j = 0
    while r:  # trim to an integer total
        d = min(r, int(k[j]))
        k[j] -= d
        r -= d
        j += 1
    return k / denom


def _thresholds(p):
    """Exact acceptance probability of every step (the chain's mass is a
    deterministic function of p): p_j / a at a step that completes no unit,
    (1 - p_j) / (1 - a) at one that does (0 where the reference's division is
    0/0, whose comparison is false)."""
    fp = [Fraction(float(v)) for v in p]
    a = fp[0]
    q = [None]
    for j in range(1, len(p)):
        a += fp[j]
        if a < 1:
            q.append(fp[j] / a if a != 0 else Fraction(0))
        else:
            a -= 1
            q.append((1 - fp[j]) / (1 - a) if a != 1 else Fraction(0))
    return q


def _enumerate(p, ns, mode):
    """Inclusion probabilities of the oracle's chain over every decision path:
    a decision is forced by a uniform just below float(q) (taken) or equal to
    it (not taken), so the oracle's own thresholds must be float(q) to
    reproduce the path."""
    n = len(p)
    q = _thresholds(p)
    incl = [Fraction(0)] * n
    total = Fraction(0)
    for d in itertools.product((0, 1), repeat=n - 1):
        w = Fraction(1)
        U = np.zeros(n)
        for j in range(1, n):
            qj = q[j]
            w *= qj if d[j - 1] else (1 - qj)
            if w == 0:
                break
            fq = float(qj)
            U[j] = np.nextafter(fq, 0.0) if d[j - 1] else fq
        if w == 0:
            continue
        S, cnt, st = lo.pivotal(p, ns, U=U, mode=mode)
        assert st == 0 and cnt == ns, (d, st, cnt)
        assert len(set(S.tolist())) == ns
        total += w
        for k in S:
            incl[int(k)] += w
    return incl, total


@pytest.mark.parametrize("n,seed", [(5, 1), (6, 2), (7, 3), (8, 4), (9, 5), (9, 6)])
@pytest.mark.parametrize("mode", [lo.CANONICAL, lo.FAITHFUL])
def test_inclusion_probabilities_exact(n, seed, mode):
    p = _dyadic_p(n, seed)
    ns = int(round(p.sum()))
    incl, total = _enumerate(p, ns, mode)
    assert total == 1
    assert incl == [Fraction(float(v)) for v in p]


def test_inclusion_probabilities_with_certain_and_empty_entries():
    p = np.array([1.0, 0.0, 0.25, 0.75, 1.0, 0.5, 0.0, 0.5])
    for mode in (lo.CANONICAL, lo.FAITHFUL):
        incl, total = _enumerate(p, 4, mode)
        assert total == 1 and incl == [Fraction(float(v)) for v in p]


@pytest.mark.parametrize("seed", range(4))
def test_canonical_equals_faithful_on_dyadic(seed):
    p = _dyadic_p(3000, 100 + seed, denom=256)
    ns = int(round(p.sum()))
    U = np.random.default_rng(seed).random(p.size)
    a = lo.pivotal(p, ns, U=U, mode=lo.CANONICAL)
    b = lo.pivotal(p, ns, U=U, mode=lo.FAITHFUL)
    assert a[2] == 0 and b[2] == 0
    assert np.array_equal(a[0], b[0])


def test_unbiased_monte_carlo_non_dyadic():
    rng = np.random.default_rng(7)
    x = rng.random(24)
    ns = 9
    p = x / x.sum() * ns
    assert p.max() <= 1.0
    runs = 20000
    hits = np.zeros(p.size)
    for r in range(runs):
        S, cnt, st = lo.pivotal(p, ns, key=1000 + r)
        assert st == 0 and cnt == ns and len(set(S.tolist())) == ns
        hits[S] += 1
    se = np.sqrt(p * (1 - p) / runs) + 1e-12
    assert np.all(np.abs(hits / runs - p) <= 5 * se)


def test_start_at_one_chain():
    # the reference carries a = 1 from p_0 = 1 into step 1 (a -= 1 happens there)
    for p, ns in (([1.0, 1.0, 0.5, 0.5], 3), ([1.0, 0.25, 0.75], 2), ([1.0, 1.0, 1.0], 3)):
        p = np.array(p)
        for U in (np.zeros(p.size), np.full(p.size, 0.999)):
            a = lo.pivotal(p, ns, U=U, mode=lo.CANONICAL)
            b = lo.pivotal(p, ns, U=U, mode=lo.FAITHFUL)
            assert a[2] == 0 and b[2] == 0 and np.array_equal(a[0], b[0])
            assert sorted(a[0].tolist()) == sorted(set(a[0].tolist()))
    # all ones: every step emits j (0/0 acceptance), the last emits the candidate 0
    S, cnt, st = lo.pivotal(np.ones(4), 4, U=np.zeros(4))
    assert st == 0 and S.tolist() == [1, 2, 3, 0]


def test_reference_exits_are_reported():
    # wrong total mass: the last step's |a - 1| >= 1e-5 (:107-116)
    assert lo.pivotal(np.array([0.5, 0.3]), 1, U=np.zeros(2))[2] == 5
    # zero samples asked for, but step 1 completes a unit (:119-122)
    assert lo.pivotal(np.array([0.5, 0.5]), 0, U=np.zeros(2))[2] == 5
    # zero samples and no mass: the loop stops at step 1
    S, cnt, st = lo.pivotal(np.zeros(5), 0)
    assert st == 0 and cnt == 0
    # a single element never enters the loop
    assert lo.pivotal(np.array([1.0]), 1)[2] == 5
    assert lo.pivotal(np.array([0.0]), 0)[2] == 0
    # probabilities outside [0, 1]
    assert lo.pivotal(np.array([0.5, 1.5]), 2)[2] == 1
    assert lo.pivotal(np.array([0.5, np.nan]), 1)[2] == 1


def test_splitmix_stream():
    # SplitMix64 (Steele, Lea and Flood 2014): the first outputs for seed 0
    # are the published 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, ...
    u = lo.splitmix_uniforms(0, 3)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert u[1] == (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53
    assert np.all((u >= 0) & (u < 1))
    v = lo.splitmix_uniforms(lo.key_of(0.25), 100000)
    assert abs(v.mean() - 0.5) < 0.005


def _vec(n, seed, spread=8.0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) * np.exp(-rng.uniform(0, spread, n))


@pytest.mark.parametrize("n,m", [(2000, 100), (20000, 3000), (5000, 5000)])
def test_fri_pivotal_compression(n, m):
    x = _vec(n, n + m)
    i1, v1, s1 = lo.compress(x, m, lo.FRI_SYSTEMATIC, 0.3)
    i2, v2, s2 = lo.compress(x, m, lo.FRI_PIVOTAL, 0.3)
    assert s2["status"] == 0 and s2["n_D"] == s1["n_D"] and s2["remaining_norm"] == s1["remaining_norm"]
    K = s2["n_D"]
    assert np.array_equal(i2[:K], i1[:K]) and np.array_equal(v2[:K], v1[:K])
    assert len(set(i2.tolist())) == m  # pivotal never repeats an index
    ns = m - K
    if ns:
        p = np.abs(x) / s2["remaining_norm"] * ns
        S, cnt, st = lo.pivotal(np.where(np.isin(np.arange(n), i2[:K]), 0.0, p), ns, key=lo.key_of(0.3))
        assert st == 0 and np.array_equal(i2[K:], S)
        assert np.array_equal(v2[K:], x[S] / p[S])


def test_fri_pivotal_unbiased():
    x = _vec(400, 3)
    m = 60
    acc = np.zeros(x.size)
    runs = 3000
    rng = np.random.default_rng(5)
    for _ in range(runs):
        idx, val, st = lo.compress(x, m, lo.FRI_PIVOTAL, float(rng.random()))
        assert st["status"] == 0
        acc[idx] += val
    err = np.abs(acc / runs - x).sum() / np.abs(x).sum()
    assert err < 0.03, err
    assert math.isfinite(err)
