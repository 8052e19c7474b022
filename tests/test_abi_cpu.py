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
