"""ctypes binding of the C restatement (oracle/bgt_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Coordinates are the reference's internal ones (atom a = v*n + i); the public
k=1 bit flip (src/core/lattice_api.cpp:17-26) is applied by callers that mirror
the public API.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict[str, ctypes.CDLL] = {}

_dptr = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class OrcTreeStats(ctypes.Structure):
    _fields_ = [
        ("total_leaves", ctypes.c_int),
        ("correct_probability", ctypes.c_double),
        ("incorrect_probability", ctypes.c_double),
        ("false_positive_probability", ctypes.c_double),
        ("false_negative_probability", ctypes.c_double),
        ("unclassified_probability", ctypes.c_double),
        ("expected_stages", ctypes.c_double),
        ("expected_tests", ctypes.c_double),
    ]


class OrcShrinkPlan(ctypes.Structure):
    _fields_ = [
        ("original_subjects", ctypes.c_int),
        ("current_subjects", ctypes.c_int),
        ("variants", ctypes.c_int),
        ("source_atoms", ctypes.c_int),
        ("base_atom_count", ctypes.c_int),
        ("reduced_atom_count", ctypes.c_int),
        ("new_current_subjects", ctypes.c_int),
        ("current_classified_atoms", ctypes.c_uint32),
        ("removable_current_atoms", ctypes.c_uint32),
        ("new_classified_subjects", ctypes.c_uint32),
        ("original_to_current_atom_position", ctypes.c_int * 32),
        ("base_atom_masks", ctypes.c_uint32 * 32),
        ("reduced_atom_masks", ctypes.c_uint32 * 32),
    ]


def _build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(variant: str = "strict") -> ctypes.CDLL:
    """Load liboracle.so ("strict", the checker) or liboracle_fast.so ("fast", the timed baseline)."""
    if variant in _LIBS:
        return _LIBS[variant]
    name = "liboracle.so" if variant == "strict" else "liboracle_fast.so"
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        _build()
    L = ctypes.CDLL(path)
    L.orc_dilution_table.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dptr]
    L.orc_prior_prob.argtypes = [ctypes.c_uint32, ctypes.c_int, _dptr]
    L.orc_prior_prob.restype = ctypes.c_double
    L.orc_prior.argtypes = [ctypes.c_int, _dptr, _dptr]
    L.orc_response_prob.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_uint32, _dptr, ctypes.c_int]
    L.orc_response_prob.restype = ctypes.c_double
    L.orc_update_in_place.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dptr, ctypes.c_uint32,
                                      ctypes.c_uint32, _dptr, ctypes.c_int]
    L.orc_update_in_place.restype = ctypes.c_double
    L.orc_update_out.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dptr, _dptr, ctypes.c_uint32,
                                 ctypes.c_uint32, _dptr, ctypes.c_int]
    L.orc_update_out.restype = ctypes.c_double
    L.orc_zeta.argtypes = [_dptr, ctypes.c_int]
    L.orc_op_bha.argtypes = [_dptr, ctypes.c_int, _dptr]
    L.orc_op_bha.restype = ctypes.c_uint32
    L.orc_op_bha_omp.argtypes = [_dptr, ctypes.c_int, _dptr]
    L.orc_op_bha_omp.restype = ctypes.c_uint32
    L.orc_threads.argtypes = []
    L.orc_threads.restype = ctypes.c_int
    L.orc_bbpa.argtypes = [_dptr, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dptr]
    L.orc_bbpa.restype = ctypes.c_uint32
    L.orc_bbpa_scores.argtypes = [_dptr, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dptr]
    L.orc_bbpa_scores.restype = ctypes.c_uint32
    L.orc_brute_force.argtypes = [_dptr, ctypes.c_int, ctypes.c_int, _dptr]
    L.orc_brute_force.restype = ctypes.c_uint32
    L.orc_select.argtypes = [_dptr, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_select.restype = ctypes.c_uint32
    L.orc_atom_masses.argtypes = [_dptr, ctypes.c_int, _dptr]
    L.orc_classify.argtypes = [_dptr, ctypes.c_int, ctypes.c_double, ctypes.c_double, _u32p, _u32p]
    L.orc_upset_mass.argtypes = [_dptr, ctypes.c_int, ctypes.c_uint32]
    L.orc_upset_mass.restype = ctypes.c_double
    L.orc_select_experiments.argtypes = [_dptr, ctypes.c_int, ctypes.c_int, _u32p]
    L.orc_select_experiments.restype = ctypes.c_int
    L.orc_lookahead_scores.argtypes = [_dptr, ctypes.c_int, _u32p, ctypes.c_int, _dptr]
    L.orc_lookahead_scores.restype = ctypes.c_int
    L.orc_reverse_bits.argtypes = [ctypes.c_uint32, ctypes.c_int]
    L.orc_reverse_bits.restype = ctypes.c_uint32
    L.orc_make_shrink_plan.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.POINTER(OrcShrinkPlan)]
    L.orc_project_posterior.argtypes = [_dptr, _dptr, ctypes.POINTER(OrcShrinkPlan)]
    L.orc_shrink_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dptr, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                  ctypes.POINTER(ctypes.c_uint32), _dptr]
    L.orc_shrink_step.restype = ctypes.c_int
    L.orc_true_ex.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    L.orc_true_ex.restype = ctypes.c_uint32
    L.orc_run_simulation.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dptr, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     _dptr, ctypes.c_int, ctypes.POINTER(OrcTreeStats)]
    L.orc_run_simulation.restype = ctypes.c_int
    L.orc_run_simulation_noshrink.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dptr, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, ctypes.POINTER(OrcTreeStats)]
    L.orc_run_simulation_noshrink.restype = ctypes.c_int
    L.orc_now.restype = ctypes.c_double
    _LIBS[variant] = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dptr)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def dilution_table(n: int, alpha: float, h: float) -> np.ndarray:
    out = np.zeros(n * (n + 1), dtype=np.float64)
    lib().orc_dilution_table(n, alpha, h, _p(out))
    return out


def prior(pi, atoms: int | None = None, variant: str = "strict") -> np.ndarray:
    pi = _f64(pi)
    a = len(pi) if atoms is None else atoms
    out = np.empty(1 << a, dtype=np.float64)
    lib(variant).orc_prior(a, _p(pi), _p(out))
    return out


def prior_prob(state: int, pi) -> float:
    pi = _f64(pi)
    return lib().orc_prior_prob(state, len(pi), _p(pi))


def response_prob(dilution: bool, n: int, k: int, e: int, r: int, s: int, dil=None) -> float:
    d, stride = _dil_args(dil)
    return lib().orc_response_prob(int(dilution), n, k, e, r, s, _p(d), stride)


def _dil_args(dil):
    if dil is None:
        return np.zeros(1), 0
    d = _f64(dil)
    n = int(round((-1 + np.sqrt(1 + 4 * len(d))) / 2))
    return d, n + 1


def update_in_place(post: np.ndarray, dilution: bool, n: int, k: int, e: int, r: int, dil=None,
                    variant: str = "strict") -> float:
    d, stride = _dil_args(dil)
    return lib(variant).orc_update_in_place(int(dilution), n, k, _p(post), e, r, _p(d), stride)


def update_out(post: np.ndarray, dilution: bool, n: int, k: int, e: int, r: int, dil=None):
    d, stride = _dil_args(dil)
    out = np.empty_like(post)
    den = lib().orc_update_out(int(dilution), n, k, _p(post), _p(out), e, r, _p(d), stride)
    return out, den


def zeta(a: np.ndarray, n: int) -> np.ndarray:
    out = _f64(a).copy()
    lib().orc_zeta(_p(out), n)
    return out


def op_bha(post: np.ndarray, n: int, variant: str = "strict"):
    upset = np.empty(1 << n, dtype=np.float64)
    e = lib(variant).orc_op_bha(_p(post), n, _p(upset))
    return int(e), upset


def bbpa(post: np.ndarray, n: int, k: int, target: float | None = None, variant: str = "strict"):
    t = 1.0 / (1 << k) if target is None else target
    best = ctypes.c_double()
    e = lib(variant).orc_bbpa(_p(post), n, k, t, ctypes.byref(best))
    return int(e), best.value


def bbpa_scores(post: np.ndarray, n: int, k: int, target: float | None = None) -> tuple[int, np.ndarray]:
    """The Op-BBPA pick and every experiment's objective (index e holds experiment e's)."""
    t = 1.0 / (1 << k) if target is None else target
    scores = np.empty(1 << n, dtype=np.float64)
    e = lib().orc_bbpa_scores(_p(post), n, k, t, _p(scores))
    return int(e), scores


def brute_force(post: np.ndarray, n: int, k: int):
    best = ctypes.c_double()
    e = lib().orc_brute_force(_p(post), n, k, ctypes.byref(best))
    return int(e), best.value


def select(post: np.ndarray, n: int, k: int, selector: int) -> int:
    return int(lib().orc_select(_p(post), n, k, selector))


def atom_masses(post: np.ndarray, atoms: int, variant: str = "strict") -> np.ndarray:
    out = np.empty(atoms, dtype=np.float64)
    lib(variant).orc_atom_masses(_p(post), atoms, _p(out))
    return out


def classify(masses: np.ndarray, up: float, lo: float, pos: int = 0, neg: int = 0):
    p = ctypes.c_uint32(pos)
    q = ctypes.c_uint32(neg)
    m = _f64(masses)
    lib().orc_classify(_p(m), len(m), up, lo, ctypes.byref(p), ctypes.byref(q))
    return p.value, q.value


def upset_mass(post: np.ndarray, atoms: int, e: int) -> float:
    return lib().orc_upset_mass(_p(post), atoms, e)


def select_experiments(post: np.ndarray, n: int, steps: int) -> list[int]:
    out = (ctypes.c_uint32 * max(steps, 1))()
    lib().orc_select_experiments(_p(post), n, steps, out)
    return [int(out[i]) for i in range(steps)]


def lookahead_scores(post: np.ndarray, n: int, selected) -> np.ndarray:
    """Every public candidate's lookahead score after `selected` (index 0 is +inf)."""
    sel = np.ascontiguousarray(selected, dtype=np.uint32)
    out = np.full(1 << n, np.inf)
    p = _f64(post)
    if lib().orc_lookahead_scores(_p(p), n, sel.ctypes.data_as(_u32p), len(sel), _p(out)) != 0:
        raise ValueError("lookahead_scores: bad selection")
    return out


def reverse_bits(v: int, width: int) -> int:
    return int(lib().orc_reverse_bits(v, width))


def make_shrink_plan(orig: int, curr: int, k: int, classified_atoms: int, already: int) -> OrcShrinkPlan:
    plan = OrcShrinkPlan()
    lib().orc_make_shrink_plan(orig, curr, k, classified_atoms, already, ctypes.byref(plan))
    return plan


def project_posterior(src: np.ndarray, plan: OrcShrinkPlan) -> np.ndarray:
    out = np.empty(1 << plan.base_atom_count, dtype=np.float64)
    lib().orc_project_posterior(_p(_f64(src)), _p(out), ctypes.byref(plan))
    return out


def shrink_step(orig: int, k: int, up: float, lo: float, post: np.ndarray, curr: int, pos: int, neg: int,
                clas_subjs: int):
    """One update_metadata_with_shrinking (lattice_base.cpp:266-316): returns (posterior,
    current subjects, positive mask, negative mask, classified subjects) after it."""
    src = _f64(post)
    out = np.empty(src.size, dtype=np.float64)
    p, q, c = ctypes.c_uint32(pos), ctypes.c_uint32(neg), ctypes.c_uint32(clas_subjs)
    cc = lib().orc_shrink_step(orig, k, up, lo, _p(src), curr, ctypes.byref(p), ctypes.byref(q), ctypes.byref(c),
                               _p(out))
    return out[: 1 << (cc * k)].copy(), cc, p.value, q.value, c.value


def true_ex(halving: int, clas_subjs: int, orig: int) -> int:
    return int(lib().orc_true_ex(halving, clas_subjs, orig))


@dataclass
class TreeStats:
    total_leaves: int
    correct_probability: float
    incorrect_probability: float
    false_positive_probability: float
    false_negative_probability: float
    unclassified_probability: float
    expected_stages: float
    expected_tests: float

    @classmethod
    def from_c(cls, s) -> "TreeStats":
        return cls(*(getattr(s, f) for f, _ in OrcTreeStats._fields_))


def run_simulation(dilution: bool, n: int, k: int, pi, depth: int, up: float = 0.01, lo: float = 0.01,
                   branch: float = 0.0, selector: int = 0, alpha: float = 0.99, h: float = 0.005,
                   variant: str = "strict") -> TreeStats:
    pi = _f64(pi)
    dil = dilution_table(n, alpha, h) if dilution else np.zeros(1)
    st = OrcTreeStats()
    rc = lib(variant).orc_run_simulation(int(dilution), n, k, _p(pi), depth, up, lo, branch, selector,
                                         _p(dil), n + 1, ctypes.byref(st))
    if rc != 0:
        raise ValueError("invalid simulation configuration")
    return TreeStats.from_c(st)


def run_simulation_noshrink(dilution: bool, n: int, k: int, pi, depth: int, up: float = 0.01, lo: float = 0.01,
                            branch: float = 0.0, alpha: float = 0.99, h: float = 0.005) -> TreeStats:
    pi = _f64(pi)
    st = OrcTreeStats()
    rc = lib().orc_run_simulation_noshrink(int(dilution), n, k, _p(pi), depth, up, lo, branch, alpha, h,
                                           ctypes.byref(st))
    if rc != 0:
        raise ValueError("invalid simulation configuration")
    return TreeStats.from_c(st)


def now() -> float:
    return lib().orc_now()
