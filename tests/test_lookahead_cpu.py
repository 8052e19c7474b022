"""The oracle's lookahead scores (lattice_api.cpp:139-158) pinned on the CPU.

* against the oracle's restated greedy (its picks are the first lowest score at each step);
* against a second formulation (down-set sums and inclusion-exclusion), within the rounding
  bound of one score;
* against closed forms: after one experiment e, the candidates e and the full set both leave
  two groups empty, so each scores |P(s <= e) - 1/4| + |P(s not <= e) - 1/4| + 1/2;
* the reference's KAT (lattice_api_test.cpp lookahead case).
"""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests.lookahead_helpers import public_order, tau, tie_pick, transform_scores


def posts(n, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        if t % 2:
            p = rng.random(1 << n)
        else:
            p = orc.prior(rng.uniform(0.02, 0.3, n))
        out.append(p / p.sum())
    return out


@pytest.mark.parametrize("n", [3, 5, 8, 10])
def test_scores_reproduce_the_restated_greedy(n):
    for post in posts(n, 4, 10 + n):
        picks = orc.select_experiments(post, n, 4)
        for s in range(1, 4):
            sc = orc.lookahead_scores(post, n, picks[:s])
            assert 1 + int(np.argmin(sc[1:])) == picks[s]


@pytest.mark.parametrize("n,steps", [(4, 8), (7, 8), (9, 7), (11, 5)])
def test_scores_match_the_transform_formulation(n, steps):
    for post in posts(n, 4, 30 + n):
        sel = [orc.select_experiments(post, n, 1)[0]]
        for m in range(1, steps):
            ref = orc.lookahead_scores(post, n, sel)
            alt = transform_scores(post, n, sel)
            assert np.max(np.abs(ref[1:] - alt[1:])) <= tau(n, m)
            sel.append(tie_pick(ref, n, m))


def test_closed_forms_and_structural_tie():
    n = 6
    for post in posts(n, 6, 77):
        perm = public_order(n)
        pub = np.empty_like(post)
        pub[perm] = post
        for e in (1, 5, 22, 40):
            inside = sum(pub[s] for s in range(1 << n) if (e & s) == s)
            want = abs(inside - 0.25) + abs(1 - inside - 0.25) + 0.5
            sc = orc.lookahead_scores(post, n, [e])
            assert sc[(1 << n) - 1] == pytest.approx(want, abs=1e-14)
            assert sc[e] == pytest.approx(want, abs=1e-14)


def test_tie_rule_prefers_the_lowest_index():
    sc = np.array([np.inf, 0.5, 0.3 + 1e-15, 0.3, 0.7])
    assert tie_pick(sc, 2, 1) == 2
    sc = np.array([np.inf, 0.5, 0.3 + 1e-6, 0.3, 0.7])
    assert tie_pick(sc, 2, 1) == 3


def test_lookahead_kat(kats):
    k = kats["lattice_api"]
    post = orc.prior(k["prior"])
    picks = k["public_lookahead_op_bha_2"]
    assert orc.select_experiments(post, 2, 2) == picks
    assert tie_pick(orc.lookahead_scores(post, 2, picks[:1]), 2, 1) == picks[1]
