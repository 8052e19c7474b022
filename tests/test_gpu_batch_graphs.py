"""Graph replay of batched steps (bgt_hip_batch_graphs) against launched steps.

Two copies of one mix of lattices step side by side, one with graph replay off and one with
it on, through the same (experiment, response) sequence. A step is captured into a hipGraph
once per launch structure and replayed by later steps of that structure, with the step's own
descriptors uploaded by the graph's first copy. Bar: bit-identical selections, objectives,
masks and stored posteriors at every step, across several replays of one graph with launched
work and synchronisations between them, and after a change of structure and back.
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

# every selection path of the batch step: LDS, single tile, two passes (k_zeta_final with
# the band rescan, and the persistent pass 2), the one-pass and the gather BBPA; repeated shapes
MIX = [(10, 1, False, F64), (10, 1, False, F64), (7, 2, True, F64), (14, 1, True, F32), (15, 1, False, F64),
       (16, 1, False, F32), (8, 2, False, F64), (20, 1, True, F64), (11, 2, True, F32), (22, 1, False, F64),
       (6, 3, True, F64)]


def _make(i, n, k, dil, prec):
    return bgt.Lattice(DL if dil else ND, n, seeded_prior(900 + i, n * k, 0.005, 0.1), variants=k, precision=prec)


def _key(s):
    return (s.experiment, s.objective, s.runner_up, s.near_tie, s.redecided)


@pytest.fixture
def graphs():
    prev = bgt.set_batch_graphs(False)
    yield
    bgt.set_batch_graphs(prev)


def _steps(lats, tables, steps, use_graphs, seed, mix):
    """Run `steps` batched rounds; returns every step's selections and masks."""
    g = workload.SplitMix64(seed)
    pending = [None] * len(lats)
    out = []
    for step in range(steps):
        resps = [None if e is None else g.below(1 << m[1]) for e, m in zip(pending, mix)]
        bgt.set_batch_graphs(use_graphs)
        sels, masks = bgt.batch_round(lats, pending, resps, bgt.SelectorType.auto_select, 0.01, 0.01, tables)
        bgt.set_batch_graphs(False)
        out.append(([_key(s) for s in sels], list(masks)))
        pending = [s.experiment for s in sels]
    return out


def test_graph_replay_equals_launched_steps(graphs):
    tables = [bgt.DilutionTable(m[0], 0.05, 0.9) if m[2] else None for m in MIX]
    a = [_make(i, *m) for i, m in enumerate(MIX)]
    b = [_make(i, *m) for i, m in enumerate(MIX)]
    other = _make(99, 12, 1, False, F64)
    c0, r0 = bgt.batch_graph_counts()
    want = _steps(a, tables, 6, False, 11, MIX)
    # graph steps one at a time, with launched work and host reads between the replays
    g = workload.SplitMix64(11)
    pending = [None] * len(b)
    for step in range(6):
        resps = [None if e is None else g.below(1 << m[1]) for e, m in zip(pending, MIX)]
        bgt.set_batch_graphs(True)
        sels, masks = bgt.batch_round(b, pending, resps, bgt.SelectorType.auto_select, 0.01, 0.01, tables)
        bgt.set_batch_graphs(False)
        assert [_key(s) for s in sels] == want[step][0], step
        assert list(masks) == want[step][1], step
        other.select(bgt.SelectorType.op_bha)
        other.update(1 + step, step & 1)
        _ = other.atom_masses()
        pending = [s.experiment for s in sels]
    c1, r1 = bgt.batch_graph_counts()
    # the first step (no updates) and the second (updates) have their own structures; the
    # other four replay the second's graph
    assert c1 - c0 == 2
    assert r1 - r0 == 4
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.posterior(), y.posterior())


def test_graph_replay_across_structure_changes(graphs):
    # the mix steps, then a smaller mix (one lattice out: a new structure, captured), then the
    # full mix again (its graph replayed from the cache)
    mix = MIX[:6]
    tables = [bgt.DilutionTable(m[0], 0.05, 0.9) if m[2] else None for m in mix]
    a = [_make(i, *m) for i, m in enumerate(mix)]
    b = [_make(i, *m) for i, m in enumerate(mix)]
    g = workload.SplitMix64(5)
    pending = [None] * len(mix)
    c0, r0 = bgt.batch_graph_counts()
    for step, keep in enumerate([6, 6, 6, 5, 5, 6, 6]):
        resps = [None if e is None else g.below(1 << m[1]) for e, m in zip(pending, mix)]
        la, lb = a[:keep], b[:keep]
        pe, rs, ts = pending[:keep], resps[:keep], tables[:keep]
        sa, ma = bgt.batch_round(la, pe, rs, bgt.SelectorType.auto_select, 0.01, 0.01, ts)
        bgt.set_batch_graphs(True)
        sb, mb = bgt.batch_round(lb, pe, rs, bgt.SelectorType.auto_select, 0.01, 0.01, ts)
        bgt.set_batch_graphs(False)
        assert [_key(s) for s in sa] == [_key(s) for s in sb], step
        assert list(ma) == list(mb), step
        for i in range(keep):
            pending[i] = sa[i].experiment
    c1, r1 = bgt.batch_graph_counts()
    # structures: select-only (6), update (6), update (5), update (6) again
    assert c1 - c0 == 3
    assert r1 - r0 == 4
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.posterior(), y.posterior())
