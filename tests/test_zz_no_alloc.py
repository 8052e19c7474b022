"""bgt_hip_select_noalloc (selection with no device allocation, scratch in a static device
array) against bgt_hip_lattice_select on the same posterior, through the C ABI.

Bar: the same kernels run on the same values, so experiment, runner-up and both objectives
are bit-identical to the lattice's selection whenever the lattice did not re-decide a near
tie exactly (dilution lattices never do; the others are compared only then); the atom
marginals within 1e-12 of the lattice's (FP32C: 1e-6). Calls follow each other on one GPU's scratch array,
each pair with a smaller lattice second, so a stale larger lattice's contents would show.
The test allocates its own input and output buffers with the HIP runtime the library uses.
"""
import ctypes

import numpy as np
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import _native as nat
from oracle import pyoracle as orc
from tests.gpu_helpers import op_sequence, seeded_prior

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """The HIP runtime the provider library is linked against (one copy per process)."""

    def __init__(self):
        nat.lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    path = line.split()[-1]
                    break
        assert path is not None, "the provider library did not load the HIP runtime"
        h = ctypes.CDLL(path)
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        self.h = h

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, H2D) == 0
        return p

    def download(self, p, out):
        assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self, p):
        assert self.h.hipFree(p) == 0


@pytest.fixture(scope="module")
def hip():
    return Hip()


def lattice_selection(lat, selector):
    s = nat.SelectionC()
    nat.check(nat.lib().bgt_hip_lattice_select(lat.handle, selector, ctypes.byref(s)))
    return s


def noalloc_selection(hip, post_dev, n, precision, selector, scale=1.0):
    res = hip.alloc(ctypes.sizeof(nat.SelectionC))
    mdev = hip.alloc(8 * n)
    storage = nat.FP64 if precision == bgt.Precision.float64 else nat.FP32C
    rc = nat.lib().bgt_hip_select_noalloc(post_dev, n, storage, selector, scale, res, mdev, None)
    assert rc == nat.OK, nat.lib().bgt_hip_last_error()
    assert hip.h.hipDeviceSynchronize() == 0
    s = nat.SelectionC()
    assert hip.h.hipMemcpy(ctypes.addressof(s), res, ctypes.sizeof(s), D2H) == 0
    masses = hip.download(mdev, np.empty(n))
    hip.free(res)
    hip.free(mdev)
    return s, masses


def case(hip, n, precision, selector, dilution, updates):
    """One lattice: its selection and the no-alloc selection on its posterior."""
    pi = seeded_prior(900 + n + 7 * updates, n, 0.02, 0.3)
    lt = bgt.LatticeType.replicated_dilution if dilution else bgt.LatticeType.replicated_non_dilution
    lat = bgt.Lattice(lt, n, pi, precision=precision)
    tab = bgt.DilutionTable(n, 0.05, 0.9) if dilution else None
    for e, r in op_sequence(n + updates, n, 1, updates):
        lat.update(e, r, tab)
    if precision == bgt.Precision.float64:
        # the normalised posterior: each entry y * (1/D), as the lattice's kernels read it
        host = lat.posterior()
    else:
        # a prior lattice (pending scale 1): the stored FP32 entries are the FP64 product rounded
        assert updates == 0
        host = orc.prior(pi).astype(np.float32)
    want = lattice_selection(lat, selector)
    m_want = lat.atom_masses()
    post_dev = hip.upload(host)
    got, masses = noalloc_selection(hip, post_dev, n, precision, selector)
    hip.free(post_dev)
    lat.close()
    return want, got, masses, m_want


# (subjects of the first call, subjects of the second): every path (one workgroup in LDS up
# to 14, one tile at 15, pass 1 + final pass above) and back down
PAIRS = [(20, 16), (18, 15), (16, 14), (15, 9), (12, 3)]


@pytest.mark.parametrize("precision", [bgt.Precision.float64, bgt.Precision.float32_compensated])
@pytest.mark.parametrize("selector", [1, 2])
@pytest.mark.parametrize("pair", PAIRS)
def test_noalloc_matches_lattice_select(hip, precision, selector, pair):
    compared = 0
    for n in pair:
        fp64 = precision == bgt.Precision.float64
        # dilution lattices have no exact re-decision, so their selections always compare
        for dilution, updates in ((True, 3), (False, 0)) if fp64 else ((True, 0), (False, 0)):
            want, got, masses, m_want = case(hip, n, precision, selector, dilution, updates)
            assert got.redecided == 0 and got.candidates == 0
            # FP32C above 2^15 states spills the transform's partial sums as FP32 (one more
            # rounding than the lattice's direct marginal sums): the FP32C bar
            tol = 1e-12 if fp64 else 1e-6
            assert np.max(np.abs(masses - m_want)) <= tol, (n, dilution)
            if want.redecided:
                continue  # the lattice re-decided a near tie exactly; the no-alloc call does not
            assert (got.experiment, got.runner_up) == (want.experiment, want.runner_up), (n, dilution)
            assert (got.objective, got.runner_up_objective) == (want.objective, want.runner_up_objective)
            assert got.near_tie == (1 if got.runner_up_objective - got.objective <= got.gap_bound else 0)
            compared += 1
    assert compared >= 1


def test_noalloc_scale_is_the_pending_normalisation(hip):
    # entries stored as y with the normalisation pending: y * scale is the posterior
    n = 17
    pi = seeded_prior(77, n, 0.02, 0.3)
    x = orc.prior(pi)
    a = hip.upload(x)
    b = hip.upload(x * 4.0)
    sa, ma = noalloc_selection(hip, a, n, bgt.Precision.float64, 1)
    sb, mb = noalloc_selection(hip, b, n, bgt.Precision.float64, 1, scale=0.25)
    hip.free(a)
    hip.free(b)
    assert (sa.experiment, sa.objective) == (sb.experiment, sb.objective)
    np.testing.assert_array_equal(ma, mb)
    e_ref, upset = orc.op_bha(x, n)
    assert sa.experiment == e_ref
    assert abs(sa.objective - abs(upset[e_ref] - 0.5)) <= 1e-13


def test_noalloc_rejects_what_does_not_fit(hip):
    res = hip.alloc(ctypes.sizeof(nat.SelectionC))
    dummy = hip.alloc(64)
    L = nat.lib()
    assert L.bgt_hip_select_noalloc(dummy, 21, 0, 1, 1.0, res, None, None) == nat.UNSUPPORTED
    assert b"20 subjects" in L.bgt_hip_last_error()
    assert L.bgt_hip_select_noalloc(None, 10, 0, 1, 1.0, res, None, None) == nat.BAD_INPUT
    assert L.bgt_hip_select_noalloc(dummy, 10, 0, 3, 1.0, res, None, None) == nat.BAD_INPUT
    assert L.bgt_hip_select_noalloc(dummy, 10, 0, 1, 0.0, res, None, None) == nat.BAD_INPUT
    assert L.bgt_hip_select_noalloc(dummy, 10, 5, 1, 1.0, res, None, None) == nat.BAD_INPUT
    hip.free(dummy)
    hip.free(res)
