"""Pin the CPU restatement (oracle/) to the reference's own known-answer tests.

Every check here comes from a reference test (cited in tests/golden/reference_kats.json)
or from what the mathematics fixes independently of the restated loop.
"""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests.conftest import expr

TOL = 1e-12


def public_posterior(post, n):
    # lattice_api.cpp:89-93: public state s reads internal reverse_bits(s, n) for k=1
    return np.array([post[orc.reverse_bits(s, n)] for s in range(1 << n)])


def test_model_math_prior_and_update(kats):
    k = kats["model_math"]
    post = orc.prior(k["prior"])
    np.testing.assert_allclose(post, k["internal_prior_posterior"], atol=TOL, rtol=0)
    u = k["update_out_of_place"]
    nxt, den = orc.update_out(post, False, 2, 1, u["experiment"], u["response"])
    terms = [a * b for a, b in u["denominator_terms"]]
    assert abs(den - sum(terms)) < TOL
    np.testing.assert_allclose(nxt, np.array(terms) / sum(terms), atol=TOL, rtol=0)


def test_model_math_dilution_table(kats):
    k = kats["model_math"]["dilution_2_099_0005"]
    tab = orc.dilution_table(2, k["alpha"], k["h"]).reshape(2, 3)
    for row, col, val in k["entries"]:
        assert abs(tab[row, col] - expr(val)) < TOL


def test_lattice_api_public_prior_masses_and_selection(kats):
    k = kats["lattice_api"]
    post = orc.prior(k["prior"])
    np.testing.assert_allclose(public_posterior(post, 2), k["public_prior_posterior"], atol=TOL, rtol=0)
    for e_pub, val in k["public_upset_mass"].items():
        e = orc.reverse_bits(int(e_pub), 2)
        assert abs(orc.upset_mass(post, 2, e) - val) < TOL
    sel = {"auto_select": 0, "op_bha": 1, "op_bbpa": 2}
    for name, want in k["public_select"].items():
        assert orc.reverse_bits(orc.select(post, 2, 1, sel[name]), 2) == want
    assert orc.select_experiments(post, 2, 2) == k["public_lookahead_op_bha_2"]


def test_lattice_api_update_and_classification(kats):
    k = kats["lattice_api"]
    post = orc.prior(k["prior"])
    u = k["public_update"]
    orc.update_in_place(post, False, 2, 1, orc.reverse_bits(u["experiment"], 2), u["response"])
    den = expr(u["denominator"])
    want = [a * b / den for a, b in u["numerators"]]
    np.testing.assert_allclose(public_posterior(post, 2), want, atol=TOL, rtol=0)
    masses = orc.atom_masses(post, 2)
    pos, neg = orc.classify(masses, 0.01, 0.01)
    c = k["classification_0_01"]
    assert orc.reverse_bits(pos, 2) == c["positive_mask"]
    assert orc.reverse_bits(neg, 2) == c["negative_mask"]
    assert (bin(pos | neg).count("1") == 2) == c["is_classified"]


def test_lattice_api_multinomial_bbpa_equals_brute_force(kats):
    m = kats["lattice_api"]["multinomial"]
    post = orc.prior(m["prior"])
    n, k = m["subjects"], m["variants"]
    e_bbpa, s_bbpa = orc.bbpa(post, n, k)
    e_bf, s_bf = orc.brute_force(post, n, k)
    assert e_bbpa == e_bf
    assert abs(s_bbpa - s_bf) < TOL


def test_dilution_response_probabilities(kats):
    k = kats["dilution"]
    dil = orc.dilution_table(2, k["alpha"], k["h"])
    for e, r, s, val in k["public_response_probability"]:
        got = orc.response_prob(True, 2, 1, orc.reverse_bits(e, 2), r, orc.reverse_bits(s, 2), dil)
        assert abs(got - expr(val)) < TOL


def test_shrink_plan_fields(kats):
    for c in kats["shrink_plan"]["cases"]:
        p = orc.make_shrink_plan(c["orig"], c["curr"], c["k"], c["classified"], c["already"])
        for field in ("current_classified_atoms", "removable_current_atoms", "base_atom_count",
                      "reduced_atom_count", "new_current_subjects"):
            if field in c:
                assert getattr(p, field) == c[field], field
        for a, want in c.get("original_to_current_atom_position", {}).items():
            assert p.original_to_current_atom_position[int(a)] == want


@pytest.mark.parametrize("orig,curr,k,cls,already", [
    (2, 2, 1, 1, 0), (2, 2, 2, 5, 0), (3, 2, 2, 36, 2)])
def test_shrink_projection_brute_force_and_conservation(orig, curr, k, cls, already):
    # shrink_projection_regression_test.cpp:53-82
    src = np.arange(1, (1 << (curr * k)) + 1, dtype=np.float64)
    src /= src.sum()
    p = orc.make_shrink_plan(orig, curr, k, cls, already)
    got = orc.project_posterior(src, p)
    base = [p.base_atom_masks[i] for i in range(p.base_atom_count)]
    red = [p.reduced_atom_masks[i] for i in range(p.reduced_atom_count)]
    want = np.zeros(1 << p.base_atom_count)
    for b in range(len(want)):
        bs = sum(m for i, m in enumerate(base) if b >> i & 1)
        for r in range(1 << len(red)):
            rs = sum(m for i, m in enumerate(red) if r >> i & 1)
            want[b] += src[bs | rs]
    np.testing.assert_allclose(got, want, atol=TOL, rtol=0)
    assert abs(got.sum() - src.sum()) < TOL


def test_shrink_property_random():
    # shrink_property_regression_test.cpp:116-161 (own generator; same properties)
    rng = np.random.default_rng(20260625)
    for subjects, variants in [(3, 1), (2, 2), (3, 2)]:
        for it in range(32):
            already = 2 if (it % 5 == 0 and subjects > 2) else 0
            curr = subjects - bin(already).count("1")
            src = rng.uniform(0.01, 1.0, 1 << (curr * variants))
            src /= src.sum()
            cls = 0
            for s in range(subjects):
                if already >> s & 1:
                    continue
                ch = rng.integers(0, 4)
                if ch == 0:
                    continue
                if ch == 1 and variants > 1:
                    cls |= 1 << s
                    continue
                for v in range(variants):
                    cls |= 1 << (v * subjects + s)
            p = orc.make_shrink_plan(subjects, curr, variants, cls, already)
            got = orc.project_posterior(src, p)
            assert abs(got.sum() - src.sum()) < 1e-11
            for i in range(p.base_atom_count):
                m = p.base_atom_masks[i]
                src_mass = src[[s for s in range(len(src)) if s & m]].sum()
                proj_mass = got[[s for s in range(len(got)) if s >> i & 1]].sum()
                assert abs(src_mass - proj_mass) < 1e-11


def _assert_same_stats(a, b, tol=TOL):
    assert a.total_leaves == b.total_leaves
    for f in ("correct_probability", "incorrect_probability", "false_positive_probability",
              "false_negative_probability", "unclassified_probability", "expected_stages", "expected_tests"):
        assert abs(getattr(a, f) - getattr(b, f)) <= tol, f


def test_tree_stats_binary_single_subject_kat(kats):
    k = kats["tree_stats_binary_single_subject"]
    st = orc.run_simulation(False, 1, 1, k["prior"], k["search_depth"], k["threshold_up"], k["threshold_lo"],
                            k["branch_threshold"])
    e = k["expected"]
    assert st.total_leaves == e["total_leaves"]
    for f, v in e.items():
        assert abs(getattr(st, f) - v) <= TOL, f


def test_tree_multinomial_leaf_count(kats):
    k = kats["tree_multinomial_leaves"]
    st = orc.run_simulation(False, k["subjects"], k["variants"], k["prior"], k["search_depth"])
    assert st.total_leaves == k["total_leaves"]


def test_tree_shrinking_matches_no_shrink_restatement(kats):
    for c in kats["tree_no_shrink_cases"]["cases"]:
        n, k = c["subjects"], c["variants"]
        pi = c["prior"] if "prior" in c else [c["prior_scalar"]] * (n * k)
        a = orc.run_simulation(c["dilution"], n, k, pi, c["depth"], c["up"], c["lo"], c["branch"])
        b = orc.run_simulation_noshrink(c["dilution"], n, k, pi, c["depth"], c["up"], c["lo"], c["branch"])
        _assert_same_stats(a, b)


def test_cli_scalar_prior_leaf_count(kats):
    c = kats["cli_scalar_prior_e2e"]
    st = orc.run_simulation(False, c["subjects"], 1, [c["prior_scalar"]] * c["subjects"], c["search_depth"],
                            c["threshold_up"], c["threshold_lo"], c["branch_threshold"], selector=2)
    assert st.total_leaves == c["total_leaves"]
