"""C ABI checks that need no GPU: the library loads, exports every declared entry point,
and reports the absence of a device through status codes rather than crashing."""
import ctypes

import numpy as np
import pytest

from bayesiangrouptesting_amd import _native as nat
import bayesiangrouptesting_amd as bgt


def test_library_exports_every_declared_symbol():
    lib = nat.lib()
    declared = nat.declared_symbols()
    assert len(declared) >= 24
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []


def test_status_codes_match_reference_convention():
    assert (nat.OK, nat.BAD_INPUT, nat.UNSUPPORTED, nat.NO_BACKEND, nat.INTERNAL) == (0, 1, 2, 3, 4)


def test_distributed_entry_point_reports_single_gpu():
    st = nat.TreeStatsC()
    prior = np.array([0.1], dtype=np.float64)
    rc = nat.lib().bgt_hip_provider_run_distributed(0, 1, 1, prior.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                    1, 0.01, 0.01, 0.0, 0, 0, None, ctypes.byref(st))
    assert rc == nat.NO_BACKEND
    assert b"one GPU" in nat.lib().bgt_hip_last_error()


def test_null_arguments_are_rejected_without_device_work():
    assert nat.lib().bgt_hip_lattice_update(None, 1, 1, None) == nat.BAD_INPUT
    assert nat.lib().bgt_hip_lattice_destroy(None) == nat.OK
    sel = nat.SelectionC()
    assert nat.lib().bgt_hip_lattice_select(None, 0, ctypes.byref(sel)) == nat.BAD_INPUT


def test_python_mirror_validation_matches_reference_messages():
    with pytest.raises(ValueError, match="subjects and variants must be positive"):
        bgt.Lattice(bgt.LatticeType.replicated_non_dilution, 0, [])
    with pytest.raises(ValueError, match="prior length must equal subjects"):
        bgt.Lattice(bgt.LatticeType.replicated_non_dilution, 2, [0.1])
    cfg = bgt.SimulationConfig(subjects=2, variants=2, prior=[0.1] * 4)
    cfg.options.selector = bgt.SelectorType.op_bha
    with pytest.raises(ValueError, match="Op-BHA tree selection is only supported for binary lattices"):
        bgt.run_simulation(cfg)


def test_dilution_table_mirror_matches_oracle():
    from oracle import pyoracle as orc
    t = bgt.DilutionTable(7, 0.05, 0.9)
    np.testing.assert_array_equal(t.values, orc.dilution_table(7, 0.05, 0.9))
    assert t.at(3, 0) == 0.05
    with pytest.raises(IndexError):
        t.at(8, 0)


def test_no_device_is_reported_as_unsupported_when_absent():
    if nat.available():
        pytest.skip("a device is present")
    with pytest.raises(bgt.Unsupported):
        bgt.Lattice(bgt.LatticeType.replicated_non_dilution, 2, [0.2, 0.4])


def test_public_bit_flip_matches_the_oracle():
    # lattice_api.cpp:17-26, the low `width` bits reversed; the mirror's byte-table form
    # against the oracle's bit loop, widths 0..32
    from oracle import pyoracle as orc
    rng = np.random.default_rng(5)
    for width in range(0, 33):
        for v in rng.integers(0, 1 << 32, 64, dtype=np.uint64):
            assert bgt.reverse_bits(int(v), width) == (orc.reverse_bits(int(v), width) if width else 0)


def test_dilution_response_probability_kat(kats):
    # replicated_lattice.cpp:17-28 through the host mirror, public coordinates flipped as
    # bgt::Lattice does for a binary lattice (lattice_api.cpp:101-106)
    from tests.conftest import expr
    k = kats["dilution"]
    t = bgt.DilutionTable(2, k["alpha"], k["h"])
    for e, r, s, want in k["public_response_probability"]:
        got = bgt.response_probability(bgt.LatticeType.replicated_dilution, 2, 1, bgt.reverse_bits(e, 2), r,
                                       bgt.reverse_bits(s, 2), t)
        assert abs(got - expr(want)) <= 1e-12


def test_response_probability_defines_the_empty_pool():
    # no table row exists for |E| = 0 (the reference would read row -1): P(r = 1 | E = 0) = 0,
    # the device's definition; non-dilution lattices treat E = 0 as inside every state
    t = bgt.DilutionTable(3, 0.05, 0.9)
    dl, nd = bgt.LatticeType.replicated_dilution, bgt.LatticeType.replicated_non_dilution
    assert bgt.response_probability(dl, 3, 1, 0, 1, 5, t) == 0.0
    assert bgt.response_probability(dl, 3, 2, 0, 0, 5, t) == 1.0
    assert bgt.response_probability(nd, 3, 1, 0, 1, 0, None) == 0.99
    with pytest.raises(bgt.InvalidArgument):
        bgt.response_probability(dl, 3, 1, 1, 1, 0, None)


def test_response_probability_matches_the_oracle():
    from oracle import pyoracle as orc
    n, k = 3, 2
    t = bgt.DilutionTable(n, 0.05, 0.9)
    for dilution, lt in ((True, bgt.LatticeType.replicated_dilution), (False, bgt.LatticeType.replicated_non_dilution)):
        for e in range(1, 1 << n):
            for r in range(1 << k):
                for s in range(1 << (n * k)):
                    want = orc.response_prob(dilution, n, k, e, r, s, orc.dilution_table(n, 0.05, 0.9))
                    assert bgt.response_probability(lt, n, k, e, r, s, t if dilution else None) == want


def test_benchmark_select_argument_codes():
    # cuda_provider.cu:2449-2486: a missing output is 1, a shape or selector it cannot run is 2
    prior = np.full(4, 0.1)
    p = prior.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    c = ctypes.c_uint32()
    d = [ctypes.c_double() for _ in range(5)]
    refs = [ctypes.byref(x) for x in d]
    L = nat.lib()
    assert L.bgt_hip_provider_benchmark_select(4, 1, p, 1, 1, 0, None, *refs) == nat.BAD_INPUT
    assert L.bgt_hip_provider_benchmark_select(2, 2, p, 1, 1, 0, ctypes.byref(c), *refs) == nat.UNSUPPORTED
    assert L.bgt_hip_provider_benchmark_select(4, 1, p, 1, 0, 0, ctypes.byref(c), *refs) == nat.UNSUPPORTED
    assert L.bgt_hip_provider_benchmark_select(31, 1, p, 1, 1, 0, ctypes.byref(c), *refs) == nat.UNSUPPORTED
