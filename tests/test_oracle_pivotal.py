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


# ---------------------------------------------------------------- parallel_sample
# parallel_sample (src/fri_utils.cpp:176-362) with T chunks for the reference's
# T OpenMP threads.  Pinned by: T = 1 is sample_pivotal itself; chunks with
# integer sums need no extra samples and no move, so the result is each
# chunk's own pivotal chain at its offset; and the inclusion probabilities,
# predicted from the reference's move (:290-316) restated in exact rationals
# (P(chunk j gets g_j + 1) = t_j by the pivotal design), match the oracle's
# frequencies over many draws.

def _chunks(n, T):
    step = n // T
    return [(j * step, (j + 1) * step + (n % T if j == T - 1 else 0)) for j in range(T)]


@pytest.mark.parametrize("mode", [lo.CANONICAL, lo.FAITHFUL])
def test_parallel_one_chunk_is_pivotal(mode):
    """One chunk: a_0 = n' |q| / |q|, which is n' whenever that FP64 quotient
    is exact (always on dyadic q); no extra sample, no move, and the chunk's
    chain is sample_pivotal on the same draws."""
    rng = np.random.default_rng(3)
    checked = 0
    for n, ns, dyadic in ((50, 7, True), (3000, 400, True), (10_000, 900, False), (777, 50, False)):
        q = _dyadic_p(n, n, denom=64) if dyadic else rng.random(n)
        if dyadic:
            ns = int(round(q.sum()))
        else:
            q = q / q.sum() * ns
        qn = math.fsum(q) if mode == lo.CANONICAL else float(np.cumsum(q)[-1])
        if ns * qn / qn != ns:
            continue
        U = rng.random(n)
        S1, st1 = lo.parallel_sample(q, ns, 1, U=U, mode=mode)
        S2, cnt, st2 = lo.pivotal(q, ns, U=U, mode=mode)
        assert st1 == 0 and st2 == 0
        assert np.array_equal(S1, S2)
        checked += 1
    assert checked >= 2


@pytest.mark.parametrize("T", [2, 5, 16])
def test_parallel_integer_chunks_are_independent_chains(T):
    rng = np.random.default_rng(T)
    n = 4000
    q = np.zeros(n)
    for lo_, hi_ in _chunks(n, T):  # dyadic values, integer sum per chunk
        k = rng.integers(0, 65, hi_ - lo_)
        r, j = int(k.sum()) % 64, 0
        while r:
            d = min(r, int(k[j]))
            k[j] -= d
            r -= d
            j += 1
        q[lo_:hi_] = k / 64.0
    ns = int(round(q.sum()))
    U = rng.random(n)
    for mode in (lo.CANONICAL, lo.FAITHFUL):
        S, st = lo.parallel_sample(q, ns, T, U=U, mode=mode)
        assert st == 0
        off = 0
        for lo_, hi_ in _chunks(n, T):
            g = int(round(q[lo_:hi_].sum()))
            Sj, cnt, stj = lo.pivotal(q[lo_:hi_], g, U=U[lo_:hi_], mode=mode)
            assert stj == 0
            assert np.array_equal(S[off:off + g], Sj + lo_)
            off += g


def _moved(q, tj, g, up):
    """The reference's move of one chunk toward a sum of g, in exact rationals."""
    qa = [Fraction(float(v)) for v in q]
    s = sum(qa)
    for i in range(len(qa)):
        y = min(Fraction(1), qa[i] / tj) if up else max(Fraction(0), (qa[i] - tj) / (1 - tj))
        s += y - qa[i]
        qa[i] = y
        if (s >= g) if up else (s <= g):
            qa[i] = y + g - s
            break
    return qa


def test_parallel_inclusion_probabilities():
    rng = np.random.default_rng(11)
    n, T, ns = 30, 3, 7
    q = rng.random(n) ** 2
    q = q / q.sum() * ns
    assert q.max() < 1
    pi = np.zeros(n)
    Fq = [Fraction(float(v)) for v in q]
    tot = sum(Fq)
    for lo_, hi_ in _chunks(n, T):
        a = ns * sum(Fq[lo_:hi_]) / tot
        g0 = math.floor(a)
        t = a - g0
        up = _moved(q[lo_:hi_], t, g0 + 1, True) if t else Fq[lo_:hi_]
        dn = _moved(q[lo_:hi_], t, g0, False) if t else Fq[lo_:hi_]
        pi[lo_:hi_] = [float(t * u + (1 - t) * d) for u, d in zip(up, dn)]
    runs = 40000
    hits = np.zeros(n)
    for r in range(runs):
        S, st = lo.parallel_sample(q, ns, T, key=77 + r)
        assert st == 0 and len(set(S.tolist())) == ns
        hits[S] += 1
    se = np.sqrt(pi * (1 - pi) / runs) + 1e-9
    assert np.all(np.abs(hits / runs - pi) <= 5 * se), np.abs(hits / runs - pi) / se
    assert abs(pi.sum() - ns) < 1e-9


@pytest.mark.parametrize("T", [2, 7, 64])
def test_parallel_invariants(T):
    rng = np.random.default_rng(100 + T)
    n, ns = 20_000, 3_000
    q = rng.random(n) * np.exp(-rng.uniform(0, 2, n))
    q = q / q.sum() * ns
    assert q.max() <= 1.0
    for mode in (lo.CANONICAL, lo.FAITHFUL):
        S, st = lo.parallel_sample(q, ns, T, key=5, mode=mode)
        assert st == 0 and len(set(S.tolist())) == ns
        off = 0
        for lo_, hi_ in _chunks(n, T):
            a = ns * q[lo_:hi_].sum() / q.sum()
            cnt = int(((S >= lo_) & (S < hi_)).sum())
            assert cnt in (math.floor(a), math.floor(a) + 1) or abs(a - round(a)) < 1e-9
            assert np.all((S[off:off + cnt] >= lo_) & (S[off:off + cnt] < hi_))  # chunk order
            off += cnt


def test_parallel_bad_arguments():
    q = np.full(10, 0.5)
    assert lo.parallel_sample(q, 5, 0)[1] == 2
    assert lo.parallel_sample(q, 5, 11)[1] == 2
    assert lo.parallel_sample(np.array([0.5, 1.5]), 2, 1)[1] == 1
    assert lo.parallel_sample(np.zeros(10), 0, 3)[1] == 0
    assert lo.parallel_sample(np.zeros(10), 2, 3)[1] == 5
