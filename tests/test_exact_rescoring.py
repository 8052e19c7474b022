"""The exact FP64 upset-mass program used to re-decide near ties (host code, no GPU).

Pinned three ways: against the oracle's FP64 posterior after the same updates
(update_in_place, lattice_base.cpp:230-244, then the upset sum of lattice_base.cpp:318-351),
against brute-force enumeration with exact rational arithmetic on tiny lattices, and
against the closed form of a single update (a two-term mixture of prior products)."""
from fractions import Fraction

import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import _native as nat
from bayesiangrouptesting_amd import workload
from oracle import pyoracle as orc


def _history(seed, n, steps):
    g = workload.SplitMix64(seed)
    return [(g.below(1 << n), g.below(2)) for _ in range(steps)]


# Up to 10 remaining pools under at most 64 updates the dense program runs (the device
# kernel's); beyond either limit the sparse one. (14, 16) has more than 10 distinct pools
# for most probes; the last case repeats 5 experiments 70 times (more than 64 updates).
@pytest.mark.parametrize("n,steps,seed", [(4, 3, 1), (8, 6, 2), (10, 10, 3), (12, 8, 4), (14, 12, 5), (14, 16, 7),
                                          (10, -70, 6)])
def test_exact_mass_matches_oracle_posterior(n, steps, seed):
    pi = np.array(workload.priors(seed, n, 0.01, 0.3))
    if steps > 0:
        hist = _history(seed + 100, n, steps)
    else:
        few = _history(seed + 100, n, 5)
        g = workload.SplitMix64(seed + 200)
        hist = [(few[g.below(5)][0], g.below(2)) for _ in range(-steps)]
    post = orc.prior(pi)
    for e, r in hist:
        orc.update_in_place(post, False, n, 1, e, r)
    g = workload.SplitMix64(seed + 7)
    probes = [0, (1 << n) - 1] + [g.below(1 << n) for _ in range(12)] + [e for e, _ in hist]
    for e in probes:
        want = orc.upset_mass(post, n, e)
        got = bgt.exact_upset_mass(n, pi, hist, e)
        assert abs(got - want) <= 1e-12 * max(want, 1e-300) + 1e-300, (e, got, want)


def _brute_exact(n, pi, hist, e):
    """M(E) with exact rationals: sum over S >= E of P(S) prod L / sum over all S."""
    fpi = [Fraction(p) for p in pi]
    neg = Fraction(99, 100)
    num = Fraction(0)
    den = Fraction(0)
    for s in range(1 << n):
        w = Fraction(1)
        for a in range(n):
            w *= (1 - fpi[a]) if (s >> a) & 1 else fpi[a]
        for ex, r in hist:
            contained = (ex & s) == ex
            w *= (neg if r else 1 - neg) if contained else ((1 - neg) if r else neg)
        den += w
        if (s & e) == e:
            num += w
    return num / den


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_exact_mass_matches_rational_brute_force(seed):
    n = 6
    pi = np.array(workload.priors(seed, n, 0.05, 0.4))
    hist = _history(seed, n, 5)
    for e in [0, 1, 5, 12, 33, 63]:
        want = float(_brute_exact(n, pi, hist, e))
        assert abs(bgt.exact_upset_mass(n, pi, hist, e) - want) <= 4e-15 * want


def test_single_update_closed_form():
    n = 20
    pi = np.array(workload.priors(3, n, 0.005, 0.05))
    q = 1.0 - pi
    e0 = 0b1011_0110_0101_1101_0011
    for r in (0, 1):
        f_in = 0.99 if r else 1.0 - 0.99
        f_out = (1.0 - 0.99) if r else 0.99
        m0 = float(np.prod([q[a] for a in range(n) if (e0 >> a) & 1]))
        z = f_out + (f_in - f_out) * m0
        for e in (0, 1 << 3, 0b111, e0, e0 | (1 << 19)):
            mf = float(np.prod([q[a] for a in range(n) if (e >> a) & 1]))
            mfe = float(np.prod([q[a] for a in range(n) if ((e | e0) >> a) & 1]))
            want = (f_out * mf + (f_in - f_out) * mfe) / z
            assert abs(bgt.exact_upset_mass(n, pi, [(e0, r)], e) - want) <= 1e-14 * want


def test_bad_arguments_are_status_codes():
    assert nat.lib().bgt_hip_exact_upset_mass(0, None, 0, None, None, 0, None, None) == nat.BAD_INPUT


def test_batch_matches_single_queries():
    n = 12
    pi = np.array(workload.priors(9, n, 0.01, 0.2))
    hist = _history(19, n, 7)
    qs = list(range(0, 1 << n, 37))
    batch = bgt.exact_upset_mass(n, pi, hist, qs)
    assert np.array_equal(batch, np.array([bgt.exact_upset_mass(n, pi, hist, e) for e in qs]))


# ---- dilution lattices: the program over counts of positive pooled atoms

def _dilution_history(seed, n, steps):
    g = workload.SplitMix64(seed)
    if steps < 0:
        # an adaptive-like history: a few experiments, repeated
        few = [g.below((1 << n) - 1) + 1 for _ in range(5)]
        return [(few[g.below(5)], g.below(2)) for _ in range(-steps)]
    return [(g.below((1 << n) - 1) + 1, g.below(2)) for _ in range(steps)]


@pytest.mark.parametrize("n,steps,seed,alpha,h", [(4, 3, 1, 0.05, 0.9), (8, 6, 2, 0.05, 0.9), (10, 10, 3, 0.05, 0.9),
                                                  (12, 8, 4, 0.99, 0.005), (14, 12, 5, 0.05, 0.9),
                                                  (12, -30, 7, 0.05, 0.9)])
def test_dilution_exact_mass_matches_oracle_posterior(n, steps, seed, alpha, h):
    pi = np.array(workload.priors(seed, n, 0.01, 0.3))
    hist = _dilution_history(seed + 100, n, steps)
    tab = bgt.DilutionTable(n, alpha, h)
    dil = orc.dilution_table(n, alpha, h)
    post = orc.prior(pi)
    for e, r in hist:
        orc.update_in_place(post, True, n, 1, e, r, dil)
    g = workload.SplitMix64(seed + 7)
    probes = [0, (1 << n) - 1] + [g.below(1 << n) for _ in range(12)] + [e for e, _ in hist]
    got = bgt.exact_upset_mass(n, pi, hist, probes, dilution=tab)
    for e, m in zip(probes, got):
        want = orc.upset_mass(post, n, e)
        assert abs(m - want) <= 1e-12 * max(want, 1e-300) + 1e-300, (e, m, want)


def _brute_exact_dilution(n, pi, hist, e, alpha, h):
    """M(E) with exact rationals under the dilution likelihood (dilution.cpp:44-67)."""
    fpi = [Fraction(p) for p in pi]
    fa, fh = Fraction(alpha), Fraction(h)

    def p1(s, m):
        return fa if m == 0 else 1 - fa * m / ((s - m) * fh + m)

    num = Fraction(0)
    den = Fraction(0)
    for st in range(1 << n):
        w = Fraction(1)
        for a in range(n):
            w *= (1 - fpi[a]) if (st >> a) & 1 else fpi[a]
        for ex, r in hist:
            s = bin(ex).count("1")
            m = s - bin(ex & st).count("1")
            w *= p1(s, m) if r else 1 - p1(s, m)
        den += w
        if (st & e) == e:
            num += w
    return num / den


@pytest.mark.parametrize("seed", [21, 22])
def test_dilution_exact_mass_matches_rational_brute_force(seed):
    n = 6
    pi = np.array(workload.priors(seed, n, 0.05, 0.4))
    hist = _dilution_history(seed, n, 5)
    tab = bgt.DilutionTable(n, 0.05, 0.9)
    for e in [0, 1, 5, 12, 33, 63]:
        want = float(_brute_exact_dilution(n, pi, hist, e, 0.05, 0.9))
        # the table's entries are FP64 roundings of the rationals: a few ulp per factor
        assert abs(bgt.exact_upset_mass(n, pi, hist, e, dilution=tab) - want) <= 1e-13 * want


def test_dilution_single_update_closed_form():
    # one pool E0 of size s: U(E) sums, over the count m of positive atoms of E0 \ E, the
    # probability of m positives among those atoms times the factor for m
    n = 9
    pi = np.array(workload.priors(5, n, 0.05, 0.3))
    q = 1.0 - pi
    e0 = 0b101101101
    tab = bgt.DilutionTable(n, 0.05, 0.9)
    s = bin(e0).count("1")
    for r in (0, 1):
        f = [(tab.values[(s - 1) * (n + 1) + m] if r else 1.0 - tab.values[(s - 1) * (n + 1) + m])
             for m in range(s + 1)]

        def u(e):
            rest = [a for a in range(n) if (e0 >> a) & 1 and not (e >> a) & 1]
            qe = float(np.prod([q[a] for a in range(n) if (e >> a) & 1]))
            dist = np.zeros(len(rest) + 1)  # distribution of the positive count over rest
            dist[0] = 1.0
            for a in rest:
                dist[1:] = dist[1:] * q[a] + dist[:-1] * pi[a]
                dist[0] *= q[a]
            return qe * float(np.dot(dist, f[:len(rest) + 1]))

        z = u(0)
        for e in (0, 1, 0b11, e0, e0 | 2, 0b111111111):
            assert abs(bgt.exact_upset_mass(n, pi, [(e0, r)], e, dilution=tab) - u(e) / z) <= 1e-14 * (u(e) / z)
