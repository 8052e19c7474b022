"""Batched steps (bgt_hip_batch_step / bgt_hip_batch_round) against per-lattice calls.

Bar: bit-identical. Each lattice in a batch runs the same kernels with the same launch
shapes as the per-lattice calls: the fused round (update + classification + auto_select
selection, DeviceLattice::round) for a step that updates, and update / classify / select
otherwise. Selections, objectives, classification masks and the stored posteriors must match
exactly, whatever the mix."""
import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import _native as nat
from bayesiangrouptesting_amd import workload
from tests.gpu_helpers import seeded_prior

pytestmark = pytest.mark.gpu
ND = bgt.LatticeType.replicated_non_dilution
DL = bgt.LatticeType.replicated_dilution
F64 = bgt.Precision.float64
F32 = bgt.Precision.float32_compensated

# (subjects, variants, dilution, precision): every selection path (LDS, single tile, two
# passes with the old and the persistent pass 2, gather BBPA), repeated shapes included
MIX = [(10, 1, False, F64), (10, 1, False, F64), (7, 2, True, F64), (14, 1, True, F32), (15, 1, False, F64),
       (16, 1, False, F32), (8, 2, False, F64), (20, 1, True, F64), (11, 2, True, F32), (22, 1, False, F64),
       (22, 1, False, F64), (24, 1, False, F32), (6, 3, True, F64)]


def _make(i, n, k, dil, prec):
    pi = seeded_prior(500 + i, n * k, 0.005, 0.1)
    return bgt.Lattice(DL if dil else ND, n, pi, variants=k, precision=prec)


def test_batch_round_equals_per_lattice_calls():
    a = [_make(i, *m) for i, m in enumerate(MIX)]
    b = [_make(i, *m) for i, m in enumerate(MIX)]
    tables = [bgt.DilutionTable(m[0], 0.05, 0.9) if m[2] else None for m in MIX]
    g = workload.SplitMix64(7)
    pending = [None] * len(MIX)
    for step in range(4):
        resps = [None if e is None else g.below(1 << m[1]) for e, m in zip(pending, MIX)]
        sels, masks = bgt.batch_round(b, pending, resps, bgt.SelectorType.auto_select, 0.01, 0.01, tables)
        for i, (lat, e, r, t) in enumerate(zip(a, pending, resps, tables)):
            if r is not None:
                s = lat.round(e, r, 0.01, 0.01, t)
            else:
                lat.update_classification(0.01, 0.01)
                s = lat.select(bgt.SelectorType.auto_select)
            assert (s.experiment, s.objective, s.runner_up, s.near_tie, s.redecided) == \
                (sels[i].experiment, sels[i].objective, sels[i].runner_up, sels[i].near_tie, sels[i].redecided), (step, i)
            assert masks[i] == (lat.positive_classification_mask(), lat.negative_classification_mask()), (step, i)
        pending = [s.experiment for s in sels]
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.posterior(), y.posterior())


def _paired_dilution(i, n, shrink):
    # subjects 2j and 2j+1 share a prior, so an experiment holding one of a pair ties exactly
    # with its mirror image: the top two tie and FP32C re-decides by enumeration every step
    base = seeded_prior(900 + i, (n + 1) // 2, 0.02, 0.2)
    pi = np.array([base[j // 2] for j in range(n)])
    lat = bgt.Lattice(DL, n, pi, precision=F32)
    if not shrink:
        return lat
    lat.update(bgt.reverse_bits(0b111, n), 0, bgt.DilutionTable(n, 0.05, 0.9))
    plan = bgt.make_shrink_plan(n, n, 1, 0b101, 0)  # subjects 0 and 2 leave
    out = lat.project(plan)
    lat.close()
    return out


def test_batched_dilution_redecisions_equal_per_lattice_calls():
    # the batched step enumerates the first exact masses of every tied binary dilution
    # lattice in one launch; per-lattice calls enumerate their own
    shapes = [(10, False), (12, False), (13, True), (14, False), (16, False), (17, True), (18, False), (20, False)]
    a = [_paired_dilution(i, n, s) for i, (n, s) in enumerate(shapes)]
    b = [_paired_dilution(i, n, s) for i, (n, s) in enumerate(shapes)]
    subjects = [lat.subjects() for lat in a]
    tables = [bgt.DilutionTable(m, 0.05, 0.9) for m in subjects]
    g = workload.SplitMix64(11)
    pending = [None] * len(a)
    redecided = 0
    for step in range(4):
        resps = [None if e is None else g.below(2) for e in pending]
        print(f"batched dilution re-decisions: step {step}", flush=True)
        sels, masks = bgt.batch_round(b, pending, resps, bgt.SelectorType.auto_select, 0.01, 0.01, tables)
        for i, (lat, e, r, t) in enumerate(zip(a, pending, resps, tables)):
            if r is not None:
                s = lat.round(e, r, 0.01, 0.01, t)
            else:
                lat.update_classification(0.01, 0.01)
                s = lat.select(bgt.SelectorType.auto_select)
            got = sels[i]
            assert (s.experiment, s.objective, s.runner_up, s.runner_up_objective, s.near_tie, s.redecided,
                    s.candidates, s.band) == (got.experiment, got.objective, got.runner_up, got.runner_up_objective,
                                              got.near_tie, got.redecided, got.candidates, got.band), (step, i)
            assert masks[i] == (lat.positive_classification_mask(), lat.negative_classification_mask()), (step, i)
            redecided += s.redecided == 1 and s.candidates > 0
        pending = [s.experiment for s in sels]
    assert redecided >= len(a), redecided
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.posterior(), y.posterior())
        x.close()
        y.close()


def test_batch_step_without_classification_and_errors():
    n = 12
    a = [bgt.Lattice(ND, n, seeded_prior(3 + i, n, 0.01, 0.1)) for i in range(3)]
    b = [bgt.Lattice(ND, n, seeded_prior(3 + i, n, 0.01, 0.1)) for i in range(3)]
    es = [5, 17, 300]
    sels = bgt.batch_round(b, es, [1, 0, 1], bgt.SelectorType.op_bbpa, classify=False)
    for lat, e, r, s in zip(a, es, [1, 0, 1], sels):
        lat.update(e, r)
        want = lat.select(bgt.SelectorType.op_bbpa)
        assert (want.experiment, want.objective) == (s.experiment, s.objective)
    with pytest.raises(bgt.InvalidArgument):
        bgt.batch_round([b[0], b[0]], [1, 2], [0, 1])
    # a bad experiment anywhere rejects the whole batch before any lattice changes
    before = [x.posterior() for x in b]
    with pytest.raises(bgt.InvalidArgument):
        bgt.batch_round(b, [1, 3, 2], [0, 2, 0])
    for x, p in zip(b, before):
        np.testing.assert_array_equal(x.posterior(), p)


def test_footprint_matches_allocation():
    for n, k, dil, prec in MIX[:6]:
        lat = _make(0, n, k, dil, prec)
        lat.select(bgt.SelectorType.auto_select)  # allocates the intermediate
        import ctypes
        got = ctypes.c_uint64()
        nat.check(nat.lib().bgt_hip_lattice_info(lat.handle, None, None, None, ctypes.byref(got)))
        assert got.value == bgt.lattice_footprint(DL if dil else ND, n, k, prec)
