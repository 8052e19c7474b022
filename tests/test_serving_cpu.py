"""Host logic of shrinking served sessions, no GPU.

A shrinking session keeps, after each projection, the original index of every current subject
and maps the projected lattice's experiments and classification masks back through it; a
projected dilution lattice reads the table built for its current subject count. Pinned by:
  * the projected posterior is the marginal of the original over the removed subjects
    (project_posterior, shrink.hpp:70-86, the oracle's plain sums): every upset mass of a set
    of current atoms equals the original's upset mass of the set mapped back, after one
    projection and after a projection of a projection, k = 1, 2 and 3;
  * a wrong map (kept subjects reversed, or the variant stride of the current lattice kept)
    fails the same check;
  * DilutionTable(s') agrees with the oracle's table for s >= s' subjects on every (p, m) with
    p <= s' (dilution.cpp:44-67 makes D[p-1][m] a function of p and m alone).
"""
import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import serving, workload
from oracle import pyoracle as orc


def _posterior(g, n, k, steps):
    pi = np.array(workload.priors(int(g.next_u64() & 0xFFFFFFFF), n * k, 0.05, 0.4))
    post = orc.prior(pi)
    for _ in range(steps):
        e = int(g.next_u64() & ((1 << n) - 1)) or 1
        orc.update_in_place(post, False, n, k, e, g.below(1 << k))
    return post


def _classify(g, n, k, kept, removed, count):
    """Classified atoms (original numbering): the removed subjects' and `count` more subjects'."""
    subjects = [s for s in range(n) if (removed >> s) & 1]
    live = list(kept)
    for _ in range(count):
        subjects.append(live.pop(g.below(len(live))))
    return sum(1 << (v * n + s) for s in subjects for v in range(k))


def _project(g, post, n, k, kept, removed, count):
    cur = len(kept)
    classified = _classify(g, n, k, kept, removed, count)
    plan = bgt.make_shrink_plan(n, cur, k, classified, removed)
    proj = orc.project_posterior(post, orc.make_shrink_plan(n, cur, k, classified, removed))
    kept = serving._kept_after(kept, plan)
    assert len(kept) == plan.new_current_subjects and len(proj) == 1 << (len(kept) * k)
    return proj, kept, plan.new_classified_subjects


def _check_marginals(g, orig_post, n, k, proj, kept, mapper=serving._to_original):
    cur = len(kept)
    atoms = cur * k
    sets = [0, (1 << atoms) - 1] + [int(g.next_u64() & g.next_u64() & ((1 << atoms) - 1)) for _ in range(40)]
    worst = 0.0
    for x in sets:
        got = orc.upset_mass(proj, atoms, x)
        want = orc.upset_mass(orig_post, n * k, mapper(x, kept, cur, n, k))
        worst = max(worst, abs(got - want) / want)
    return worst


@pytest.mark.parametrize("n,k,first,second", [(8, 1, 2, 3), (10, 1, 1, 4), (6, 2, 1, 2), (5, 2, 2, 1), (4, 3, 1, 1)])
def test_projected_sets_map_back_to_the_original(n, k, first, second):
    g = workload.SplitMix64(n * 1000 + k * 10 + first)
    post = _posterior(g, n, k, 6)
    proj, kept, removed = _project(g, post, n, k, list(range(n)), 0, first)
    assert _check_marginals(g, post, n, k, proj, kept) <= 1e-12
    # a projection of the projection, kept still holding original indices
    proj2, kept2, removed2 = _project(g, proj, n, k, kept, removed, second)
    assert bin(removed2).count("1") == first + second and kept2 == [s for s in kept if not (removed2 >> s) & 1]
    assert _check_marginals(g, post, n, k, proj2, kept2) <= 1e-12


@pytest.mark.parametrize("n,k", [(8, 1), (6, 2)])
def test_a_wrong_map_fails_the_marginal_check(n, k):
    g = workload.SplitMix64(77 + n + k)
    post = _posterior(g, n, k, 6)
    proj, kept, _ = _project(g, post, n, k, list(range(n)), 0, 2)

    def reversed_kept(x, kept, cur, orig, variants):
        return serving._to_original(x, kept[::-1], cur, orig, variants)

    def current_stride(x, kept, cur, orig, variants):
        out = 0
        for v in range(variants):
            for j, s in enumerate(kept):
                if (x >> (v * cur + j)) & 1:
                    out |= 1 << (v * cur + s)
        return out

    assert _check_marginals(g, post, n, k, proj, kept, reversed_kept) > 1e-6
    if k > 1:
        assert _check_marginals(g, post, n, k, proj, kept, current_stride) > 1e-6


def test_projected_dilution_tables_are_the_leading_rows():
    for s in range(1, 13):
        full = orc.dilution_table(s, 0.05, 0.9).reshape(s, s + 1)
        for sp in range(1, s + 1):
            t = serving._table(sp)
            assert t.subjects() == sp
            for p in range(1, sp + 1):
                for m in range(p + 1):
                    assert t.at(p, m) == full[p - 1, m], (s, sp, p, m)
