"""Adaptive sessions in both storage forms, and lattices re-used from the lattice cache.

Bars:
* FP32C picks the same experiment as FP64 in every round whose FP64 pick is not itself a
  near tie, along FP64-driven session trajectories at 2^24 states (where the FP32C error
  band holds hundreds of candidates and the exact re-decision decides); every FP32C pick
  that was re-decided over a complete band has the least exact objective of the two picks.
* A lattice that comes back from the cache after use is indistinguishable from a new one:
  prior bit-identical to the oracle, no classification, selection equal to the oracle's.
"""
import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import workload
from oracle import pyoracle as orc
from tests.gpu_helpers import seeded_prior

pytestmark = pytest.mark.gpu
ND = bgt.LatticeType.replicated_non_dilution
F64 = bgt.Precision.float64
F32 = bgt.Precision.float32_compensated


def _session(n, pi, seed, prec, drive=None, max_tests=20):
    lat = bgt.Lattice(ND, n, pi, precision=prec)
    truth = bgt.reverse_bits(workload.true_state(seed, list(pi)), n)
    sampler = workload.ResponseSampler(seed, n, 1, truth)
    sel = lat.select(bgt.SelectorType.op_bha)
    picks, plan = [sel], []
    for i in range(max_tests):
        if drive is not None:
            if i >= len(drive):
                break
            e, r = drive[i]
        else:
            e = sel.experiment
            r = sampler.draw(e)
        plan.append((e, r))
        sel = lat.round(e, r, 0.01, 0.01)
        picks.append(sel)
        if drive is None and lat.is_classified():
            break
    lat.close()
    return picks, plan


@pytest.mark.parametrize("seed", [2000, 2002, 2007])
def test_fp32c_picks_equal_fp64_along_sessions(seed):
    n = 24
    pi = seeded_prior(2, n, 0.005, 0.05)
    p64, plan = _session(n, pi, seed, F64)
    p32, _ = _session(n, pi, seed, F32, drive=plan)
    assert len(p64) == len(p32)
    history = []
    for i, (a, b) in enumerate(zip(p64, p32)):
        if not a.near_tie:
            assert a.experiment == b.experiment, (seed, i)
        if b.redecided == 1 and a.experiment != b.experiment:
            ea = bgt.reverse_bits(a.experiment, n)
            eb = bgt.reverse_bits(b.experiment, n)
            ma, mb = bgt.exact_upset_mass(n, pi, history, [ea, eb])
            assert abs(mb - 0.5) <= abs(ma - 0.5) + 1e-15, (seed, i)
        if i < len(plan):
            history.append((bgt.reverse_bits(plan[i][0], n), plan[i][1]))


@pytest.mark.parametrize("prec", [F64, F32])
def test_recycled_lattice_is_like_new(prec):
    n = 17
    used = bgt.Lattice(ND, n, seeded_prior(41, n), precision=prec)
    for e, r in [(5, 1), (0x1F00, 0), (3, 1)]:
        used.update(e, r)
    used.update_classification(0.2, 0.2)
    used.select(bgt.SelectorType.op_bha)
    used.close()  # back to the cache
    pi = seeded_prior(42, n)
    lat = bgt.Lattice(ND, n, pi, precision=prec)
    ref = orc.prior(pi)
    got = lat.posterior()
    if prec == F64:
        np.testing.assert_array_equal(got, ref)
    else:
        np.testing.assert_allclose(got, ref, rtol=2 ** -23, atol=0)
    assert lat.positive_classification_mask() == 0 and lat.negative_classification_mask() == 0
    assert lat.error_bound()[0] <= (2.0 * 2 ** -24 if prec == F32 else 64 * 2 ** -53)
    assert bgt.reverse_bits(lat.select_experiment(bgt.SelectorType.op_bha), n) == orc.op_bha(ref, n)[0]
    lat.close()
