"""ctypes binding of libbgt_hip.so (the C ABI declared in include/bgt_hip.h).

The product path has no CPU fallback: if the library is missing, importing this
module's ``lib()`` raises. Build it with ``python -m bayesiangrouptesting_amd.build``.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbgt_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "bgt_hip.h")

OK, BAD_INPUT, UNSUPPORTED, NO_BACKEND, INTERNAL = 0, 1, 2, 3, 4
NO_UPDATE = 0xFFFFFFFF
FP64, FP32C = 0, 1
SELECT_AUTO, SELECT_OP_BHA, SELECT_OP_BBPA, SELECT_BRUTE_FORCE = 0, 1, 2, 3

_dptr = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class TreeStatsC(ctypes.Structure):
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


class SelectionC(ctypes.Structure):
    _fields_ = [
        ("experiment", ctypes.c_uint32),
        ("runner_up", ctypes.c_uint32),
        ("objective", ctypes.c_double),
        ("runner_up_objective", ctypes.c_double),
        ("gap_bound", ctypes.c_double),
        ("near_tie", ctypes.c_int),
        ("redecided", ctypes.c_int),
        ("candidates", ctypes.c_int),
        ("band", ctypes.c_int),
        ("collected", ctypes.c_int),
    ]


class ShrinkPlanC(ctypes.Structure):
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
    ]


class Error(RuntimeError):
    """A provider call failed; ``status`` carries the C ABI status code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class InvalidArgument(Error, ValueError):
    pass


class Unsupported(Error):
    pass


_LIB = None
_H = ctypes.c_void_p


def _sig(L):
    c_int, c_double, c_u32, c_u64 = ctypes.c_int, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64
    table = {
        "bgt_hip_provider_available": ([], c_int),
        "bgt_hip_provider_run": ([c_int, c_int, c_int, _dptr, c_int, c_double, c_double, c_double, c_int, _dptr,
                                  ctypes.POINTER(TreeStatsC)], c_int),
        "bgt_hip_provider_run_precision": ([c_int, c_int, c_int, _dptr, c_int, c_double, c_double, c_double, c_int,
                                            _dptr, c_int, ctypes.POINTER(TreeStatsC)], c_int),
        "bgt_hip_provider_run_distributed": ([c_int, c_int, c_int, _dptr, c_int, c_double, c_double, c_double, c_int,
                                              c_int, _dptr, ctypes.POINTER(TreeStatsC)], c_int),
        "bgt_hip_provider_benchmark_select": ([c_int, c_int, _dptr, c_int, c_int, c_int, _u32p, _dptr, _dptr, _dptr,
                                               _dptr, _dptr], c_int),
        "bgt_hip_tune": ([c_int, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_int), _dptr, _dptr,
                          ctypes.c_char_p, c_int], c_int),
        "bgt_hip_tune_force": ([c_int], c_int),
        "bgt_hip_tune_force_pass1": ([c_int], c_int),
        "bgt_hip_tune_forms": ([c_int, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                _dptr, ctypes.c_char_p, c_int], c_int),
        "bgt_hip_select_noalloc": ([ctypes.c_void_p, c_int, c_int, c_int, c_double, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p], c_int),
        "bgt_hip_lattice_create": ([c_int, c_int, c_int, _dptr, c_int, ctypes.POINTER(_H)], c_int),
        "bgt_hip_lattice_destroy": ([_H], c_int),
        "bgt_hip_lattice_update": ([_H, c_u32, c_u32, _dptr], c_int),
        "bgt_hip_lattice_select": ([_H, c_int, ctypes.POINTER(SelectionC)], c_int),
        "bgt_hip_lattice_posterior": ([_H, c_u32, _dptr], c_int),
        "bgt_hip_lattice_read_posterior": ([_H, _dptr, c_u64], c_int),
        "bgt_hip_lattice_upset_mass": ([_H, c_u32, _dptr], c_int),
        "bgt_hip_lattice_exact_masses": ([_H, c_int, _u32p, _dptr, _dptr], c_int),
        "bgt_hip_lattice_atom_masses": ([_H, _dptr], c_int),
        "bgt_hip_lattice_classify": ([_H, c_double, c_double, _u32p, _u32p], c_int),
        "bgt_hip_lattice_masks": ([_H, _u32p, _u32p], c_int),
        "bgt_hip_lattice_set_candidates": ([_H, c_u32], c_int),
        "bgt_hip_shrink_plan_make": ([c_int, c_int, c_int, c_u32, c_u32, ctypes.POINTER(ShrinkPlanC)], c_int),
        "bgt_hip_lattice_project": ([_H, ctypes.POINTER(ShrinkPlanC), ctypes.POINTER(_H)], c_int),
        "bgt_hip_lattice_select_experiments": ([_H, c_int, _u32p], c_int),
        "bgt_hip_lattice_round": ([_H, c_u32, c_u32, _dptr, c_double, c_double, ctypes.POINTER(SelectionC), _u32p,
                                   _u32p, _dptr], c_int),
        "bgt_hip_lattice_error_bound": ([_H, _dptr, _dptr], c_int),
        "bgt_hip_lattice_kernel_times": ([_H, c_int, _dptr, ctypes.POINTER(c_int), c_int, ctypes.c_char_p, c_int],
                                         c_int),
        "bgt_hip_lattice_info": ([_H, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                                  ctypes.POINTER(c_u64)], c_int),
        "bgt_hip_lattice_footprint": ([c_int, c_int, c_int, c_int, ctypes.POINTER(c_u64)], c_int),
        "bgt_hip_batch_step": ([ctypes.POINTER(_H), _u32p, _u32p, ctypes.POINTER(_dptr), c_int, c_int,
                                ctypes.POINTER(SelectionC)], c_int),
        "bgt_hip_batch_round": ([ctypes.POINTER(_H), _u32p, _u32p, ctypes.POINTER(_dptr), c_int, c_int, c_double,
                                 c_double, ctypes.POINTER(SelectionC), _u32p, _u32p], c_int),
        "bgt_hip_batch_graphs": ([c_int, ctypes.POINTER(c_int)], c_int),
        "bgt_hip_tree_shrink": ([c_int, ctypes.POINTER(c_int)], c_int),
        "bgt_hip_batch_graph_counts": ([ctypes.POINTER(c_u64), ctypes.POINTER(c_u64)], c_int),
        "bgt_hip_exact_upset_mass": ([c_int, _dptr, c_int, _u32p, _u32p, c_int, _u32p, _dptr], c_int),
        "bgt_hip_exact_upset_mass_dilution": ([c_int, _dptr, _dptr, c_int, _u32p, _u32p, c_int, _u32p, _dptr], c_int),
        "bgt_hip_last_error": ([], ctypes.c_char_p),
        "bgt_hip_device_synchronize": ([], c_int),
        "bgt_hip_device_name": ([ctypes.c_char_p, c_int], c_int),
    }
    for name, (args, res) in table.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


def declared_symbols() -> list[str]:
    """Every entry point include/bgt_hip.h declares."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(bgt_hip_\w+)\s*\(", text, re.M)))


def lib():
    """Load the provider library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: the HIP provider has not been built "
                               "(python -m bayesiangrouptesting_amd.build)")
        _LIB = _sig(ctypes.CDLL(LIB_PATH))
    return _LIB


def check(status: int) -> None:
    if status == OK:
        return
    msg = lib().bgt_hip_last_error().decode(errors="replace")
    if status == BAD_INPUT:
        raise InvalidArgument(status, msg)
    if status in (UNSUPPORTED, NO_BACKEND):
        raise Unsupported(status, msg)
    raise Error(status, msg)


def available() -> bool:
    return bool(lib().bgt_hip_provider_available())
