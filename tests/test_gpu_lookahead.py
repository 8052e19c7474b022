"""GPU parity of the k-step Op-BHA lookahead (lattice_api.cpp:116-168).

Up to four earlier picks, below 13 subjects the library scores every candidate by a pass over
all states; from 13 on it builds down-set sums once and scores each candidate from 2^(m+1) of
them. From five earlier picks on (6..16 steps) it sums, per class of states under the earlier
picks, the down-set sums of the posterior masked to that class. At every step,
given the device's own earlier picks, the pick must be the lowest public candidate whose
oracle score lies within 2 * tau of the oracle's lowest score (tau = G(G(n+1)+3) 2^-53 with
G = 2^(m+1) groups: the FP64 rounding bound of one score). Where the oracle's scores single
out one candidate, that is the oracle's own pick; candidates tied in exact arithmetic go to
the lowest index rather than to summation rounding. Beyond the sizes the oracle's O(4^n)
scan finishes in seconds, the check uses the down-set formulation on the CPU, itself pinned
to the oracle in the CPU tests.
"""
import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from oracle import pyoracle as orc
from tests.gpu_helpers import op_sequence, seeded_prior
from tests.lookahead_helpers import near_tie, tie_pick, transform_scores

pytestmark = pytest.mark.gpu

ND = bgt.LatticeType.replicated_non_dilution


def session(n, seed, updates, precision=bgt.Precision.float64, lo=0.02, hi=0.3):
    lat = bgt.Lattice(ND, n, seeded_prior(seed, n, lo, hi), precision=precision)
    for e, r in op_sequence(seed, n, 1, updates):
        lat.update(e, r)
    return lat


def check_steps(lat, n, steps, scorer):
    picks = lat.select_experiments(steps)
    assert len(picks) == steps
    assert picks[0] == lat.select_experiment(bgt.SelectorType.op_bha)
    post = lat.posterior()
    ties = 0
    for m in range(1, steps):
        sc = scorer(post, n, picks[:m])
        ties += near_tie(sc, n, m)
        assert picks[m] == tie_pick(sc, n, m), (n, m, picks)
    return picks, ties


@pytest.mark.parametrize("n,updates,steps", [(8, 0, 4), (8, 3, 5), (12, 2, 4), (13, 0, 4), (13, 3, 5),
                                             (14, 2, 4), (15, 4, 3)])
def test_lookahead_against_oracle_scores(n, updates, steps):
    lat = session(n, 300 + n + updates, updates)
    print(f"lookahead n={n} updates={updates} steps={steps}", flush=True)
    check_steps(lat, n, steps, orc.lookahead_scores)


@pytest.mark.parametrize("n,updates,steps", [(6, 1, 8), (8, 2, 8), (10, 3, 7), (11, 0, 6), (12, 2, 6), (13, 1, 7)])
def test_lookahead_many_steps_against_oracle_scores(n, updates, steps):
    # from five earlier picks on, the class-decomposed scores: in LDS up to 12 subjects,
    # through global down-set transforms from 13
    lat = session(n, 400 + n + updates, updates)
    print(f"lookahead n={n} updates={updates} steps={steps}", flush=True)
    check_steps(lat, n, steps, orc.lookahead_scores)


def test_lookahead_sixteen_steps():
    # the largest lookahead: 2^16 groups at the last step
    lat = session(4, 41, 1)
    print("lookahead n=4 steps=16", flush=True)
    check_steps(lat, 4, 16, orc.lookahead_scores)


@pytest.mark.parametrize("n", [12, 14])
def test_lookahead_fp32c_against_its_own_posterior(n):
    lat = session(n, 900 + n, 3, precision=bgt.Precision.float32_compensated)
    print(f"lookahead fp32c n={n}", flush=True)
    check_steps(lat, n, 4, orc.lookahead_scores)


@pytest.mark.parametrize("n,updates,steps,precision", [(18, 4, 5, bgt.Precision.float64),
                                                       (20, 6, 4, bgt.Precision.float64),
                                                       (20, 6, 4, bgt.Precision.float32_compensated),
                                                       (22, 5, 3, bgt.Precision.float64),
                                                       (24, 4, 2, bgt.Precision.float32_compensated),
                                                       (16, 2, 7, bgt.Precision.float64),
                                                       (16, 3, 6, bgt.Precision.float32_compensated),
                                                       (18, 2, 6, bgt.Precision.float64)])
def test_lookahead_larger_lattices(n, updates, steps, precision):
    lat = session(n, 700 + n, updates, precision=precision)
    print(f"lookahead n={n} steps={steps} {precision}", flush=True)
    check_steps(lat, n, steps, transform_scores)


def test_lookahead_structural_tie_goes_to_lowest_index():
    # A posterior whose first pick holds little mass: after it, every candidate that splits
    # the outside mass into two halves above 1/4 scores the same in exact arithmetic.
    n = 13
    lat = session(n, 55, 0, lo=0.4, hi=0.6)
    print("lookahead structural tie", flush=True)
    picks, ties = check_steps(lat, n, 3, orc.lookahead_scores)
    assert ties >= 1
    # the oracle's own first-lowest pick differs here: its summation rounding decides
    sc = orc.lookahead_scores(lat.posterior(), n, picks[:1])
    assert picks[1] != 1 + int(np.argmin(sc[1:]))


def test_lookahead_edges():
    lat = session(13, 5, 1)
    assert lat.select_experiments(0) == []
    assert lat.select_experiments(1) == [lat.select_experiment(bgt.SelectorType.op_bha)]
    with pytest.raises(bgt.Unsupported):
        lat.select_experiments(17)
    with pytest.raises(bgt.InvalidArgument):
        lat.select_experiments(-1)
    m = bgt.Lattice(ND, 4, seeded_prior(3, 8, 0.05, 0.2), variants=2)
    with pytest.raises(bgt.InvalidArgument):
        m.select_experiments(2, bgt.SelectorType.op_bha)
