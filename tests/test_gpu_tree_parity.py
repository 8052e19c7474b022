"""run_simulation -> TreeStats on the GPU against the CPU restatement's shrinking tree
(simulation_runner.cpp:92-132) and the reference's own known answers. Bar: 1e-12
absolute (tests/cpp/support/assertions.hpp:14-31), identical leaf counts. Every case runs
twice: with the device tree's shrinking children and with full-size children restricted to
the subjects still in play (the default)."""
import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import workload
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("correct_probability", "incorrect_probability", "false_positive_probability",
          "false_negative_probability", "unclassified_probability", "expected_stages", "expected_tests")
SEL = {"auto_select": 0, "op_bha": 1, "op_bbpa": 2, "brute_force": 3}


@pytest.fixture(autouse=True, params=[True, False], ids=["shrink", "full"])
def tree_mode(request):
    prev = bgt.set_tree_shrink(request.param)
    yield request.param
    bgt.set_tree_shrink(prev)


def gpu_stats(dilution, n, k, pi, depth, up=0.01, lo=0.01, branch=0.0, selector=0, alpha=0.99, h=0.005,
              precision=bgt.Precision.float64):
    cfg = bgt.SimulationConfig(
        lattice_type=bgt.LatticeType.replicated_dilution if dilution else bgt.LatticeType.replicated_non_dilution,
        subjects=n, variants=k, prior=list(pi))
    cfg.options.search_depth = depth
    cfg.options.threshold_up = up
    cfg.options.threshold_lo = lo
    cfg.options.branch_threshold = branch
    cfg.options.selector = bgt.SelectorType(selector)
    cfg.options.precision = precision
    cfg.dilution.alpha = alpha
    cfg.dilution.h = h
    return bgt.run_simulation(cfg).stats


def assert_same(a, b, tol=1e-12):
    assert a.total_leaves == b.total_leaves
    for f in FIELDS:
        assert abs(getattr(a, f) - getattr(b, f)) <= tol, (f, getattr(a, f), getattr(b, f))


def test_binary_single_subject_kat(kats):
    k = kats["tree_stats_binary_single_subject"]
    st = gpu_stats(False, 1, 1, k["prior"], k["search_depth"], k["threshold_up"], k["threshold_lo"],
                   k["branch_threshold"])
    e = k["expected"]
    assert st.total_leaves == e["total_leaves"]
    for f, v in e.items():
        assert abs(getattr(st, f) - v) <= 1e-12, f


def test_multinomial_leaf_count(kats):
    k = kats["tree_multinomial_leaves"]
    assert gpu_stats(False, 1, 2, k["prior"], 1).total_leaves == k["total_leaves"]


def test_no_shrink_reference_cases(kats):
    for c in kats["tree_no_shrink_cases"]["cases"]:
        n, k = c["subjects"], c["variants"]
        pi = c["prior"] if "prior" in c else [c["prior_scalar"]] * (n * k)
        g = gpu_stats(c["dilution"], n, k, pi, c["depth"], c["up"], c["lo"], c["branch"])
        ref = orc.run_simulation(c["dilution"], n, k, pi, c["depth"], c["up"], c["lo"], c["branch"])
        assert_same(g, ref)


def test_cli_scalar_prior_leaves(kats):
    c = kats["cli_scalar_prior_e2e"]
    g = gpu_stats(False, 3, 1, [0.02] * 3, 5, 0.01, 0.01, 0.001, selector=2)
    assert g.total_leaves == c["total_leaves"]


def test_cuda_parity_configs(kats):
    for c in kats["cuda_parity_configs"]["cases"]:
        sel = SEL[c["selector"]]
        g = gpu_stats(c["dilution"], c["subjects"], c["variants"], c["prior"], c["depth"], branch=c["branch"],
                      selector=sel)
        ref = orc.run_simulation(c["dilution"], c["subjects"], c["variants"], c["prior"], c["depth"],
                                 branch=c["branch"], selector=sel)
        assert_same(g, ref)


@pytest.mark.parametrize("n,k,dilution,depth,selector,branch", [
    (12, 1, False, 3, 1, 0.0),     # the plain CPU reference workload (n=12, depth 3)
    (8, 1, True, 4, 2, 0.001),
    (6, 2, False, 3, 2, 0.0),
    (4, 3, True, 2, 2, 0.0),
    (10, 1, True, 5, 3, 0.001),
])
def test_random_configs_against_shrinking_oracle(n, k, dilution, depth, selector, branch):
    pi = workload.priors(300 + n * k + depth, n * k, 0.01, 0.15)
    g = gpu_stats(dilution, n, k, pi, depth, branch=branch, selector=selector, alpha=0.05, h=0.9)
    ref = orc.run_simulation(dilution, n, k, pi, depth, branch=branch, selector=selector, alpha=0.05, h=0.9)
    assert_same(g, ref)


def test_batched_tree_workload_n16_depth6():
    # binary dilution n=16, DilutionTable(16, 0.05, 0.9), priors U[0.01, 0.10], Op-BHA, depth 6
    c = workload.C4
    pi = workload.priors(c.seed, c.subjects, c.prior_lo, c.prior_hi)
    g = gpu_stats(True, 16, 1, pi, 6, selector=1, alpha=0.05, h=0.9)
    ref = orc.run_simulation(True, 16, 1, pi, 6, selector=1, alpha=0.05, h=0.9)
    assert_same(g, ref)
    f = gpu_stats(True, 16, 1, pi, 6, selector=1, alpha=0.05, h=0.9, precision=bgt.Precision.float32_compensated)
    assert f.total_leaves == ref.total_leaves
    for fld in FIELDS:
        assert abs(getattr(f, fld) - getattr(ref, fld)) <= 1e-6 * max(1.0, abs(getattr(ref, fld))), fld
