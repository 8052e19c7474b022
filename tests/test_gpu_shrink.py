"""Device projection (project_posterior, shrink.hpp:70-86) and the shrinking bookkeeping of
model::Lattice (update_metadata_with_shrinking, lattice_base.cpp:266-316) against the oracle.

Bars:
* The projection of a device posterior is bit-identical to the oracle's projection of the
  same values when it removes at most 6 atoms (2^6 terms per kept entry, summed in the
  reference's serial order). FP64 projects the device's FP64 values; FP32C projects its
  FP32 values and stores each FP64 sum rounded to FP32, which the oracle's sums rounded the
  same way give. Larger removals sum each kept entry's terms in chunks (a fixed order, not
  the serial one): within the serial sum's own rounding bound (2^r roundings) of the
  oracle's FP64 sums, FP32C within one FP32 rounding of them.
* A shrinking session (update, classification, projection whenever a subject becomes fully
  classified) follows the oracle's own update_metadata_with_shrinking step for step: equal
  classification masks, removed subjects and subject counts, and posteriors within 1e-12
  relative (the device's denominator is a tree sum, the reference's a serial one).
* Bad plans are refused with the ABI's status codes.
"""
import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import workload
from oracle import pyoracle as orc
from tests.gpu_helpers import seeded_prior

pytestmark = pytest.mark.gpu
ND = bgt.LatticeType.replicated_non_dilution
DL = bgt.LatticeType.replicated_dilution
F64 = bgt.Precision.float64
F32 = bgt.Precision.float32_compensated


def _responses(seed, k, steps):
    g = workload.SplitMix64(seed)
    return [g.below(1 << k) for _ in range(steps)]


@pytest.mark.parametrize("n,k,dil,prec,removed_subjects", [
    (10, 1, False, F64, [1, 3]), (10, 1, False, F32, [0, 9]), (6, 2, True, F64, [2]),
    (5, 3, True, F32, [0, 4]), (16, 1, False, F32, [5, 6, 7]), (12, 2, False, F64, [11]),
])
def test_projection_is_bit_identical(n, k, dil, prec, removed_subjects):
    pi = seeded_prior(700 + n * k, n * k, 0.02, 0.2)
    table = bgt.DilutionTable(n, 0.05, 0.9) if dil else None
    sl = bgt.ShrinkingLattice(DL if dil else ND, n, pi, variants=k, precision=prec, dilution=table)
    g = workload.SplitMix64(n * 31 + k)
    for r in _responses(n + k, k, 3):
        e = g.below(1 << n) | 1
        sl.update(e, r)
    x = sl.lattice.posterior()
    classified = 0
    for i in removed_subjects:
        for v in range(k):
            classified |= 1 << (v * n + i)
    plan = bgt.make_shrink_plan(n, n, k, classified, 0)
    assert plan.reduced_atom_count == len(removed_subjects) * k
    proj = sl.lattice.project(plan)
    got = proj.posterior()
    want = orc.project_posterior(x, orc.make_shrink_plan(n, n, k, classified, 0))
    if prec == F32:
        want = want.astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(got, want)
    # the source is unchanged and still usable
    np.testing.assert_array_equal(sl.lattice.posterior(), x)
    proj.close()
    sl.close()


@pytest.mark.parametrize("n,k,dil,prec,removed_subjects", [
    (14, 1, False, F64, list(range(2, 12))), (12, 1, True, F32, [0, 1, 2, 4, 5, 7, 8, 9, 11]),
    (9, 2, False, F64, [0, 1, 2, 3, 5, 6, 7]), (16, 1, False, F64, list(range(13))), (7, 3, True, F32, [1, 2, 5]),
])
def test_wide_projection_matches_the_oracle(n, k, dil, prec, removed_subjects):
    pi = seeded_prior(760 + n * k, n * k, 0.02, 0.2)
    table = bgt.DilutionTable(n, 0.05, 0.9) if dil else None
    sl = bgt.ShrinkingLattice(DL if dil else ND, n, pi, variants=k, precision=prec, dilution=table)
    g = workload.SplitMix64(n * 37 + k)
    for r in _responses(n + 2 * k, k, 3):
        sl.update(g.below(1 << n) | 1, r)
    x = sl.lattice.posterior()
    classified = 0
    for i in removed_subjects:
        for v in range(k):
            classified |= 1 << (v * n + i)
    plan = bgt.make_shrink_plan(n, n, k, classified, 0)
    r = plan.reduced_atom_count
    assert r > 6
    print(f"wide projection n={n} k={k} removes {r} atoms, {prec.name}", flush=True)
    proj = sl.lattice.project(plan)
    got = proj.posterior()
    want = orc.project_posterior(x, orc.make_shrink_plan(n, n, k, classified, 0))
    if prec == F64:
        # both are sums of 2^r positive terms: the oracle's serial order carries up to 2^r
        # roundings of the sum, the device's chunked order fewer
        np.testing.assert_array_less(np.abs(got - want), (2.0 ** r + 64) * 2.0 ** -53 * want + 1e-300)
    else:
        # one FP32 rounding of sums that differ far below FP32's resolution
        w32 = want.astype(np.float32).astype(np.float64)
        assert np.all(np.abs(got - w32) <= np.spacing(np.abs(w32).astype(np.float32)).astype(np.float64)), \
            np.max(np.abs(got - w32) / w32)
    # the mass is kept (FP32C: each stored entry one FP32 rounding)
    assert abs(got.sum() - x.sum()) <= (1e-12 if prec == F64 else 2.0 ** -23) * x.sum()
    proj.close()
    sl.close()


@pytest.mark.parametrize("n,k,dil,seed", [(12, 1, False, 1), (14, 1, False, 2), (6, 2, True, 3), (5, 3, False, 4)])
def test_shrinking_session_follows_oracle(n, k, dil, seed):
    pi = seeded_prior(900 + seed, n * k, 0.01, 0.12)
    up = lo = 0.05
    table = bgt.DilutionTable(n, 0.05, 0.9) if dil else None
    dflat = table.values if dil else None  # D[s-1][m], stride n + 1
    sl = bgt.ShrinkingLattice(DL if dil else ND, n, pi, variants=k, precision=F64, dilution=table)
    post = orc.prior(np.array(pi))
    curr, pos, neg, clas = n, 0, 0, 0
    g = workload.SplitMix64(seed * 977)
    shrinks = 0
    for step in range(12):
        e = sl.select_experiment(bgt.SelectorType.op_bha if k == 1 else bgt.SelectorType.op_bbpa)
        if e == 0 or e == 0xFFFFFFFF:
            e = 1
        r = g.below(1 << k)
        sl.update(e, r)
        orc.update_in_place(post, dil, curr, k, e, r, dflat)
        # stop before a step that would classify every subject (the device lattice keeps one)
        nxt = orc.shrink_step(n, k, up, lo, post, curr, pos, neg, clas)
        if nxt[1] == 0:
            break
        sl.update_metadata_with_shrinking(up, lo)
        shrunk = nxt[1] < curr
        post, curr, pos, neg, clas = nxt
        shrinks += shrunk
        assert (sl.positive, sl.negative) == (pos, neg), step
        assert (sl.classified_subjects, sl.current_subjects) == (clas, curr), step
        np.testing.assert_allclose(sl.lattice.posterior(), post, rtol=1e-12, atol=0)
    assert shrinks >= 1, "the session never shrank: pick a case that classifies"
    sl.close()


def test_bad_plans_are_refused():
    n = 8
    lat = bgt.Lattice(ND, n, seeded_prior(5, n))
    other = bgt.make_shrink_plan(9, 9, 1, 0b11, 0)
    with pytest.raises(bgt.InvalidArgument):
        lat.project(other)  # plan for another shape
    everything = bgt.make_shrink_plan(n, n, 1, (1 << n) - 1, 0)
    with pytest.raises(bgt.Unsupported):
        lat.project(everything)  # would remove every subject
    lat.close()
