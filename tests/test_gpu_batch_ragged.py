"""Batched steps over lattices of uneven (n, k) in merged launches, against per-lattice calls.

A batch step launches each kernel once per group: updates and marginals once per storage
form, selections once per transform regime (whole transform in LDS, one tile, 16..20 atoms),
form, objective and variant count, whatever the lattices' sizes. Every lattice takes its own
sizes and workgroup counts from its descriptor, the counts its per-lattice launches use, so
its work is split the same way. Bar: bit-identical selections, objectives, runner-ups,
classification masks and stored posteriors, with BHA, k = 1 BBPA and k >= 2 BBPA, both forms.
Under auto_select a step that updates runs as the lattice's fused round (update in the
selection's first pass above one tile), so the per-lattice reference is round(); under
op_bbpa it is update / classify / select.
"""
import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import workload
from tests.gpu_helpers import seeded_prior

pytestmark = pytest.mark.gpu
ND = bgt.LatticeType.replicated_non_dilution
DL = bgt.LatticeType.replicated_dilution
F64 = bgt.Precision.float64
F32 = bgt.Precision.float32_compensated

# (subjects, variants, dilution): within each form, several sizes of every merged regime
SHAPES = [
    # whole transform in LDS (at most 14 atoms): k = 1 of several sizes, k = 2 and k = 3 gather
    (8, 1, False), (11, 1, True), (13, 1, False), (14, 1, True), (5, 2, False), (6, 2, True), (7, 2, False),
    (4, 3, True),
    # one tile (15 atoms): k = 1 both likelihoods, k = 3 gather from the posterior
    (15, 1, False), (15, 1, True), (5, 3, False),
    # 16..20 atoms (pass 1 + a tiled final pass, band rescan): k = 1 of every size, k = 4 gather
    (16, 1, True), (17, 1, False), (18, 1, True), (19, 1, False), (20, 1, False), (4, 4, False), (5, 4, True),
    # shape groups beside them: one-pass BBPA tables, the persistent pass 2
    (9, 2, False), (21, 1, True), (22, 1, False),
]


def _make(i, n, k, dil, prec):
    pi = seeded_prior(1300 + i, n * k, 0.005, 0.1)
    return bgt.Lattice(DL if dil else ND, n, pi, variants=k, precision=prec)


def _key(s):
    return (s.experiment, s.objective, s.runner_up, s.runner_up_objective, s.near_tie, s.redecided)


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("selector", [bgt.SelectorType.auto_select, bgt.SelectorType.op_bbpa])
def test_merged_launches_equal_per_lattice_calls(prec, selector):
    a = [_make(i, n, k, d, prec) for i, (n, k, d) in enumerate(SHAPES)]
    b = [_make(i, n, k, d, prec) for i, (n, k, d) in enumerate(SHAPES)]
    tables = [bgt.DilutionTable(n, 0.05, 0.9) if d else None for (n, k, d) in SHAPES]
    g = workload.SplitMix64(11 + int(selector))
    pending = [None] * len(SHAPES)
    for step in range(4):
        resps = [None if e is None else g.below(1 << k) for e, (n, k, d) in zip(pending, SHAPES)]
        print(f"ragged batch: form {prec.name} selector {selector.name} step {step}", flush=True)
        sels, masks = bgt.batch_round(b, pending, resps, selector, 0.01, 0.01, tables)
        for i, (lat, e, r, t) in enumerate(zip(a, pending, resps, tables)):
            if r is not None and selector == bgt.SelectorType.auto_select:
                s = lat.round(e, r, 0.01, 0.01, t)
            else:
                if r is not None:
                    lat.update(e, r, t)
                lat.update_classification(0.01, 0.01)
                s = lat.select(selector)
            assert _key(s) == _key(sels[i]), (step, SHAPES[i])
            assert masks[i] == (lat.positive_classification_mask(), lat.negative_classification_mask()), \
                (step, SHAPES[i])
        pending = [s.experiment for s in sels]
    for x, y, shape in zip(a, b, SHAPES):
        np.testing.assert_array_equal(x.posterior(), y.posterior(), err_msg=str(shape))
        np.testing.assert_array_equal(x.atom_masses(), y.atom_masses(), err_msg=str(shape))
    for x in a + b:
        x.close()


def test_merged_launches_alone_in_a_group():
    # one lattice per regime group, and a group of one shape repeated: the merged launch of
    # a group of equal shapes is the shape-group launch
    shapes = [(12, 1, False), (12, 1, False), (12, 1, False), (18, 1, True)]
    a = [_make(40 + i, n, k, d, F64) for i, (n, k, d) in enumerate(shapes)]
    b = [_make(40 + i, n, k, d, F64) for i, (n, k, d) in enumerate(shapes)]
    tables = [bgt.DilutionTable(n, 0.05, 0.9) if d else None for (n, k, d) in shapes]
    sels, _ = bgt.batch_round(b, [None] * 4, [None] * 4, bgt.SelectorType.auto_select, 0.01, 0.01, tables)
    es = [s.experiment for s in sels]
    sels2, masks2 = bgt.batch_round(b, es, [1, 0, 1, 0], bgt.SelectorType.auto_select, 0.01, 0.01, tables)
    for lat, e, r, t, s in zip(a, es, [1, 0, 1, 0], tables, sels2):
        lat.select(bgt.SelectorType.auto_select)
        assert _key(lat.round(e, r, 0.01, 0.01, t)) == _key(s)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.posterior(), y.posterior())
