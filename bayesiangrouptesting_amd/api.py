"""Host-side mirror of the reference's public API for the hot path, over the HIP provider.

Names, argument meanings and error behaviour follow include/bgt/lattice.hpp,
include/bgt/types.hpp, include/bgt/simulation.hpp and include/bgt/runtime.hpp. Every
numerical operation runs on the GPU through the C ABI (include/bgt_hip.h); this module
only validates arguments, applies the public bit flip of binary lattices
(src/core/lattice_api.cpp:17-26) and converts types.
"""
from __future__ import annotations

import ctypes
import enum
import time
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native as nat
from . import reporting
from ._native import Error, InvalidArgument, Unsupported  # noqa: F401


class LatticeType(enum.IntEnum):
    # include/bgt/types.hpp:81-87
    distributed_non_dilution = 0
    distributed_dilution = 1
    replicated_non_dilution = 2
    replicated_dilution = 3


class SelectorType(enum.IntEnum):
    # include/bgt/types.hpp:89-94, plus brute_force (advertised in README.md:87-89)
    auto_select = 0
    op_bha = 1
    op_bbpa = 2
    brute_force = 3


class Provider(enum.IntEnum):
    # include/bgt/types.hpp:96-101, plus the HIP provider
    auto_select = 0
    cpu = 1
    cuda = 2
    hip = 3


class Precision(enum.IntEnum):
    # include/bgt/compile_options.hpp:19-23 plus the compensated FP32 storage form
    float32 = 0
    float64 = 1
    float32_compensated = 2


class DilutionMode(enum.IntEnum):
    non_dilution = 0
    dilution = 1


def _storage(precision: Precision) -> int:
    if precision == Precision.float64:
        return nat.FP64
    if precision in (Precision.float32_compensated, Precision.float32):
        return nat.FP32C
    raise InvalidArgument(nat.BAD_INPUT, "unknown precision")


# bit reversal of every byte value
_REV8 = bytes(int(format(i, "08b")[::-1], 2) for i in range(256))


def reverse_bits(value: int, width: int) -> int:
    """lattice_api.cpp:17-26: the low `width` bits of `value` (width <= 32) in reverse order.
    By bytes through a table: a round calls it several times, on its host critical path."""
    if width <= 0:
        return 0
    v = value & 0xFFFFFFFF
    r = (_REV8[v & 0xFF] << 24) | (_REV8[(v >> 8) & 0xFF] << 16) | (_REV8[(v >> 16) & 0xFF] << 8) | _REV8[v >> 24]
    return r >> (32 - width)


def _uses_dilution(t: LatticeType) -> bool:
    return t in (LatticeType.replicated_dilution, LatticeType.distributed_dilution)


class DilutionTable:
    """include/bgt/lattice.hpp:17-32; entries as dilution.cpp:44-67 (row s-1: pool size s)."""

    def __init__(self, subjects: int = 0, alpha: float = 0.99, h: float = 0.005):
        if subjects <= 0:
            raise InvalidArgument(nat.BAD_INPUT, "subjects must be positive.")
        self._subjects = subjects
        vals = np.zeros(subjects * (subjects + 1), dtype=np.float64)
        for rk in range(1, subjects + 1):
            base = (rk - 1) * (subjects + 1)
            vals[base] = alpha
            for r in range(1, rk + 1):
                k = rk - r
                vals[base + r] = 1 - alpha * r / (k * h + r)
        self.values = vals

    def subjects(self) -> int:
        return self._subjects

    def at(self, group_size: int, positives: int) -> float:
        if group_size <= 0 or group_size > self._subjects or positives < 0 or positives > group_size:
            raise IndexError("dilution table index is out of range.")
        return float(self.values[(group_size - 1) * (self._subjects + 1) + positives])

    def pointer(self):
        return self.values.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Selection:
    experiment: int
    runner_up: int
    objective: float
    runner_up_objective: float
    gap_bound: float
    near_tie: bool
    redecided: int  # 0 no, 1 exact over the whole error band, 2 exact over the top two only
    candidates: int = 0
    band: int = 0
    collected: int = 0

    @classmethod
    def from_c(cls, s: nat.SelectionC, flip=None) -> "Selection":
        f = flip or (lambda v: v)
        return cls(f(s.experiment), f(s.runner_up) if s.runner_up != 0xFFFFFFFF else s.runner_up,
                   s.objective, s.runner_up_objective, s.gap_bound, bool(s.near_tie), int(s.redecided),
                   int(s.candidates), int(s.band), int(s.collected))


def response_probability(lattice_type: "LatticeType", subjects: int, variants: int, experiment: int,
                         response: int, true_state: int, dilution: Optional[DilutionTable] = None) -> float:
    """L(E, r, S) in internal coordinates (replicated_lattice.cpp:8-28): the product over
    variants of P(r_v | E, S_v). Non-dilution: P(r_v = 1) = 0.99 when the pool E lies inside
    S_v's negatives, 0.01 otherwise. Dilution: P(r_v = 1) = D[|E| - 1][missing], missing =
    |E| - |E & S_v|; the empty pool has no table row (the reference would index row -1), and
    P(r_v = 1 | E = 0) = 0 here as on the device."""
    ret = 1.0
    mask = (1 << subjects) - 1
    if not _uses_dilution(lattice_type):
        for v in range(variants):
            inside = (experiment & (true_state >> (v * subjects)) & mask) == experiment
            rb = (response >> v) & 1
            ret *= (0.99 if rb else 1.0 - 0.99) if inside else (1.0 - 0.99 if rb else 0.99)
        return ret
    if dilution is None:
        raise InvalidArgument(nat.BAD_INPUT, "dilution table is required for dilution lattices")
    size = bin(experiment).count("1")
    for v in range(variants):
        if size == 0:
            p1 = 0.0
        else:
            missing = size - bin(experiment & (true_state >> (v * subjects)) & mask).count("1")
            p1 = dilution.values[(size - 1) * (dilution.subjects() + 1) + missing]
        ret *= p1 if (response >> v) & 1 else 1.0 - p1
    return ret


class Lattice:
    """bgt::Lattice (include/bgt/lattice.hpp:34-67) backed by a device-resident posterior."""

    def __init__(self, type: LatticeType, subjects: int, prior: Sequence[float], variants: int = 1,
                 precision: Precision = Precision.float64):
        # lattice_api.cpp:66-77
        if subjects <= 0 or variants <= 0:
            raise InvalidArgument(nat.BAD_INPUT, "subjects and variants must be positive.")
        prior = np.ascontiguousarray(np.asarray(prior, dtype=np.float64))
        if prior.size != subjects * variants:
            raise InvalidArgument(nat.BAD_INPUT, "prior length must equal subjects * variants.")
        if type in (LatticeType.distributed_non_dilution, LatticeType.distributed_dilution):
            raise Unsupported(nat.UNSUPPORTED, "distributed lattices are multi-rank; this provider is single GPU")
        self._type = LatticeType(type)
        self._n = subjects
        self._k = variants
        self._prior = prior
        self._h = nat._H()
        nat.check(nat.lib().bgt_hip_lattice_create(int(_uses_dilution(self._type)), subjects, variants,
                                                   prior.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                   _storage(precision), ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                nat.lib().bgt_hip_lattice_destroy(h)
            except Exception:
                pass
            self._h = None

    def close(self) -> None:
        self.__del__()

    # -- shape
    def subjects(self) -> int:
        return self._n

    def variants(self) -> int:
        return self._k

    def type(self) -> LatticeType:
        return self._type

    def _flip(self, v: int) -> int:
        return reverse_bits(v, self._n) if self._k == 1 else v

    @property
    def handle(self):
        return self._h

    # -- queries
    def posterior_probability(self, state: int) -> float:
        out = ctypes.c_double()
        nat.check(nat.lib().bgt_hip_lattice_posterior(self._h, self._flip(state), ctypes.byref(out)))
        return out.value

    def posterior(self) -> np.ndarray:
        """All 2^(n*k) posterior entries in INTERNAL state order."""
        out = np.empty(1 << (self._n * self._k), dtype=np.float64)
        nat.check(nat.lib().bgt_hip_lattice_read_posterior(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
        return out

    def upset_probability_mass(self, experiment: int) -> float:
        out = ctypes.c_double()
        nat.check(nat.lib().bgt_hip_lattice_upset_mass(self._h, self._flip(experiment), ctypes.byref(out)))
        return out.value

    def exact_masses(self, experiments) -> tuple[np.ndarray, float]:
        """Exact FP64 upset masses of public-coordinate experiments, recomputed from the prior
        and the updates applied so far (the values near ties are re-decided with), and their
        relative rounding bound."""
        qs = np.ascontiguousarray(np.array([self._flip(int(e)) for e in np.atleast_1d(experiments)], dtype=np.uint32))
        out = np.empty(qs.size, dtype=np.float64)
        bound = ctypes.c_double()
        nat.check(nat.lib().bgt_hip_lattice_exact_masses(
            self._h, qs.size, qs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(bound)))
        return out, bound.value

    def atom_masses(self) -> np.ndarray:
        """Atom marginals in INTERNAL atom order."""
        out = np.empty(self._n * self._k, dtype=np.float64)
        nat.check(nat.lib().bgt_hip_lattice_atom_masses(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def response_probability(self, experiment: int, response: int, true_state: int,
                             dilution: Optional[DilutionTable] = None) -> float:
        # lattice_api.cpp:101-106 -> replicated_lattice.cpp:8-28 (host-side scalar query)
        return response_probability(self._type, self._n, self._k, self._flip(experiment), response,
                                    self._flip(true_state), dilution)

    # -- selection
    def select(self, selector: SelectorType = SelectorType.auto_select) -> Selection:
        """Full selection record (experiment in PUBLIC coordinates)."""
        s = nat.SelectionC()
        nat.check(nat.lib().bgt_hip_lattice_select(self._h, int(selector), ctypes.byref(s)))
        return Selection.from_c(s, self._flip)

    def select_experiment(self, selector: SelectorType = SelectorType.auto_select) -> int:
        # lattice_api.cpp:108-114
        return self.select(selector).experiment

    def select_experiments(self, k: int, selector: SelectorType = SelectorType.op_bha) -> list[int]:
        # lattice_api.cpp:116-168
        if k < 0:
            raise InvalidArgument(nat.BAD_INPUT, "k must be non-negative.")
        if k == 0:
            return []
        resolved = SelectorType.op_bha if (selector == SelectorType.auto_select and self._k == 1) else selector
        if selector == SelectorType.auto_select and self._k != 1:
            resolved = SelectorType.op_bbpa
        if resolved == SelectorType.op_bbpa:
            if k != 1:
                raise InvalidArgument(nat.BAD_INPUT, "Op-BBPA currently selects one experiment at a time.")
            return [self.select_experiment(selector)]
        if resolved != SelectorType.op_bha or self._k != 1:
            raise InvalidArgument(nat.BAD_INPUT,
                                  "Op-BHA multi-experiment selection is only supported for binary lattices.")
        out = (ctypes.c_uint32 * k)()
        nat.check(nat.lib().bgt_hip_lattice_select_experiments(self._h, k, out))
        return [int(v) for v in out]

    # -- update / classification
    def update(self, experiment: int, response: int, dilution: Optional[DilutionTable] = None) -> None:
        # lattice_api.cpp:170-174
        ptr = dilution.pointer() if dilution is not None else None
        if _uses_dilution(self._type) and dilution is None:
            raise InvalidArgument(nat.BAD_INPUT, "dilution table is required for dilution lattices")
        nat.check(nat.lib().bgt_hip_lattice_update(self._h, self._flip(experiment), response, ptr))

    def round(self, experiment: int, response: int, threshold_up: float, threshold_lo: float,
              dilution: Optional[DilutionTable] = None, masses: Optional[np.ndarray] = None) -> Selection:
        """Fused update + classification + next selection (Op-BHA for binary lattices, Op-BBPA
        for multinomial ones)."""
        s = nat.SelectionC()
        ptr = dilution.pointer() if dilution is not None else None
        mp = masses.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if masses is not None else None
        nat.check(nat.lib().bgt_hip_lattice_round(self._h, self._flip(experiment), response, ptr, threshold_up,
                                                  threshold_lo, ctypes.byref(s), None, None, mp))
        return Selection.from_c(s, self._flip)

    def restrict_candidates(self, subjects_mask: int) -> None:
        """Consider only experiments that are subsets of subjects_mask (public coordinates)."""
        nat.check(nat.lib().bgt_hip_lattice_set_candidates(self._h, self._flip(subjects_mask)))

    def update_classification(self, threshold_up: float, threshold_lo: float) -> None:
        nat.check(nat.lib().bgt_hip_lattice_classify(self._h, threshold_up, threshold_lo, None, None))

    def _masks(self):
        p = ctypes.c_uint32()
        q = ctypes.c_uint32()
        nat.check(nat.lib().bgt_hip_lattice_masks(self._h, ctypes.byref(p), ctypes.byref(q)))
        return p.value, q.value

    def positive_classification_mask(self) -> int:
        return self._flip(self._masks()[0])

    def negative_classification_mask(self) -> int:
        return self._flip(self._masks()[1])

    def is_classified(self) -> bool:
        p, q = self._masks()
        return bin(p | q).count("1") == self._n * self._k

    def project(self, plan: "ShrinkPlan") -> "Lattice":
        """A new lattice without the plan's removable atoms, holding this one's posterior
        summed over them (project_posterior, shrink.hpp:70-86). This lattice is unchanged."""
        h = nat._H()
        c = plan.to_c()
        nat.check(nat.lib().bgt_hip_lattice_project(self._h, ctypes.byref(c), ctypes.byref(h)))
        out = Lattice.__new__(Lattice)
        out._type = self._type
        out._n = plan.new_current_subjects
        out._k = self._k
        out._prior = None
        out._h = h
        return out

    def error_bound(self) -> tuple[float, float]:
        a = ctypes.c_double()
        b = ctypes.c_double()
        nat.check(nat.lib().bgt_hip_lattice_error_bound(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def lattice_footprint(type: LatticeType, subjects: int, variants: int = 1,
                      precision: Precision = Precision.float64) -> int:
    """Device bytes a lattice of this shape holds once it has selected (admission control)."""
    out = ctypes.c_uint64()
    nat.check(nat.lib().bgt_hip_lattice_footprint(int(_uses_dilution(LatticeType(type))), subjects, variants,
                                                  _storage(precision), ctypes.byref(out)))
    return out.value


@dataclass
class ShrinkPlan:
    """model::ShrinkPlan (shrink.hpp:27-46). Atoms are variant-major (atom v*n + i)."""
    original_subjects: int
    current_subjects: int
    variants: int
    source_atoms: int
    base_atom_count: int
    reduced_atom_count: int
    new_current_subjects: int
    current_classified_atoms: int
    removable_current_atoms: int
    new_classified_subjects: int
    original_to_current_atom_position: list

    def has_shrink(self) -> bool:
        return self.removable_current_atoms != 0

    @property
    def base_atom_masks(self) -> list[int]:
        return [1 << a for a in range(self.source_atoms) if not (self.removable_current_atoms >> a) & 1]

    @property
    def reduced_atom_masks(self) -> list[int]:
        return [1 << a for a in range(self.source_atoms) if (self.removable_current_atoms >> a) & 1]

    @classmethod
    def from_c(cls, c: nat.ShrinkPlanC) -> "ShrinkPlan":
        atoms = c.original_subjects * c.variants
        return cls(c.original_subjects, c.current_subjects, c.variants, c.source_atoms, c.base_atom_count,
                   c.reduced_atom_count, c.new_current_subjects, c.current_classified_atoms,
                   c.removable_current_atoms, c.new_classified_subjects,
                   list(c.original_to_current_atom_position[:atoms]))

    def to_c(self) -> nat.ShrinkPlanC:
        c = nat.ShrinkPlanC()
        for f, _ in nat.ShrinkPlanC._fields_:
            if f != "original_to_current_atom_position":
                setattr(c, f, getattr(self, f))
        for i in range(32):
            c.original_to_current_atom_position[i] = (self.original_to_current_atom_position[i]
                                                      if i < len(self.original_to_current_atom_position) else -1)
        return c


def make_shrink_plan(original_subjects: int, current_subjects: int, variants: int, classified_atoms: int,
                     already_classified_subjects: int) -> ShrinkPlan:
    """make_shrink_plan (shrink_plan.cpp:124-177); classified_atoms in the original numbering."""
    c = nat.ShrinkPlanC()
    nat.check(nat.lib().bgt_hip_shrink_plan_make(original_subjects, current_subjects, variants, classified_atoms,
                                                 already_classified_subjects, ctypes.byref(c)))
    return ShrinkPlan.from_c(c)


class ShrinkingLattice:
    """The shrinking bookkeeping of model::Lattice (lattice_base.cpp:266-316) over a device
    lattice: classification masks and removed subjects in the ORIGINAL numbering, the device
    lattice holding only the subjects still in play (INTERNAL coordinates, as model::Lattice).
    Each update_metadata_with_shrinking classifies from the device's atom masses and, when
    some subject became fully classified, replaces the lattice by its projection."""

    def __init__(self, type: LatticeType, subjects: int, prior: Sequence[float], variants: int = 1,
                 precision: Precision = Precision.float64, dilution: Optional[DilutionTable] = None):
        self.lattice = Lattice(type, subjects, prior, variants=variants, precision=precision)
        self.original_subjects = subjects
        self.current_subjects = subjects
        self.variants = variants
        self.classified_subjects = 0
        self.positive = 0
        self.negative = 0
        self._table = dilution
        self._tables: dict[int, np.ndarray] = {}

    def _flat_table(self) -> Optional[np.ndarray]:
        # the device takes D[s-1][m] for the current subject count (pool sizes 1..n)
        if self._table is None:
            return None
        n, full = self.current_subjects, self._table.subjects()
        if n not in self._tables:
            rows = self._table.values.reshape(full, full + 1)
            self._tables[n] = np.ascontiguousarray(rows[:n, : n + 1])  # D[s-1][m], s, m <= n
        return self._tables[n]

    def update(self, experiment: int, response: int) -> None:
        """update_probs_in_place in the current numbering (internal coordinates)."""
        t = self._flat_table()
        ptr = t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if t is not None else None
        nat.check(nat.lib().bgt_hip_lattice_update(self.lattice.handle, experiment, response, ptr))

    def select_experiment(self, selector: SelectorType = SelectorType.auto_select) -> int:
        """The selection in the current numbering (internal coordinates)."""
        s = nat.SelectionC()
        nat.check(nat.lib().bgt_hip_lattice_select(self.lattice.handle, int(selector), ctypes.byref(s)))
        return s.experiment

    def update_metadata_with_shrinking(self, threshold_up: float, threshold_lo: float) -> None:
        """lattice_base.cpp:266-305, then apply_shrink_plan (307-316)."""
        k, orig = self.variants, self.original_subjects
        classified = self.positive | self.negative
        old = make_shrink_plan(orig, self.current_subjects, k, classified, self.classified_subjects)
        masses = self.lattice.atom_masses()
        new_pos = new_neg = 0
        for i in range(orig * k):
            if (classified >> i) & 1:
                continue
            cur = old.original_to_current_atom_position[i]
            if cur < 0:
                continue
            m = masses[cur]
            if m < threshold_lo:
                new_pos |= 1 << i
            elif m > 1 - threshold_up:
                new_neg |= 1 << i
        self.positive |= new_pos
        self.negative |= new_neg
        plan = make_shrink_plan(orig, self.current_subjects, k, self.positive | self.negative,
                                self.classified_subjects)
        if plan.has_shrink() and plan.new_current_subjects > 0:
            projected = self.lattice.project(plan)
            self.lattice.close()
            self.lattice = projected
            self.classified_subjects = plan.new_classified_subjects
            self.current_subjects = plan.new_current_subjects

    def close(self) -> None:
        self.lattice.close()


def batch_round(lattices: Sequence["Lattice"], experiments: Sequence[Optional[int]], responses: Sequence[Optional[int]],
                selector: SelectorType = SelectorType.auto_select, threshold_up: float = 0.01,
                threshold_lo: float = 0.01, dilutions: Optional[Sequence[Optional[DilutionTable]]] = None,
                classify: bool = True):
    """One batched step over many lattices (public coordinates): update the lattices whose
    response is not None, take marginals and classify (classify=True), then select. The
    results equal per-lattice update / update_classification / select_experiment calls.
    Returns the selections and, with classify, the (positive, negative) public masks."""
    count = len(lattices)
    hs = (nat._H * count)(*[l.handle for l in lattices])
    es = np.zeros(count, dtype=np.uint32)
    rs = np.full(count, nat.NO_UPDATE, dtype=np.uint32)
    for i, (lat, e, r) in enumerate(zip(lattices, experiments, responses)):
        if r is not None:
            es[i] = lat._flip(int(e))
            rs[i] = int(r)
    dptrs = None
    keep = []
    if dilutions is not None:
        dptrs = (ctypes.POINTER(ctypes.c_double) * count)()
        for i, t in enumerate(dilutions):
            if t is not None:
                dptrs[i] = t.pointer()
                keep.append(t)
    sels = (nat.SelectionC * count)()
    pos = np.zeros(count, dtype=np.uint32)
    neg = np.zeros(count, dtype=np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    if classify:
        nat.check(nat.lib().bgt_hip_batch_round(hs, es.ctypes.data_as(u32p), rs.ctypes.data_as(u32p), dptrs, count,
                                                int(selector), threshold_up, threshold_lo, sels,
                                                pos.ctypes.data_as(u32p), neg.ctypes.data_as(u32p)))
    else:
        nat.check(nat.lib().bgt_hip_batch_step(hs, es.ctypes.data_as(u32p), rs.ctypes.data_as(u32p), dptrs, count,
                                               int(selector), sels))
    out = [Selection.from_c(sels[i], lattices[i]._flip) for i in range(count)]
    if not classify:
        return out
    masks = [(lattices[i]._flip(int(pos[i])), lattices[i]._flip(int(neg[i]))) for i in range(count)]
    return out, masks


def set_batch_graphs(enable: bool) -> bool:
    """Replay batched steps from captured hipGraphs (one per launch structure, see
    bgt_hip_batch_graphs); returns the former setting."""
    prev = ctypes.c_int(0)
    nat.check(nat.lib().bgt_hip_batch_graphs(1 if enable else 0, ctypes.byref(prev)))
    return bool(prev.value)


def set_tree_shrink(enable: bool) -> bool:
    """Shrink the children of run_simulation's tree, as the reference's LocalTree does, or
    keep them at full size with a candidate restriction (the default, equal in exact
    arithmetic; bgt_hip_tree_shrink); returns the former setting."""
    prev = ctypes.c_int(0)
    nat.check(nat.lib().bgt_hip_tree_shrink(1 if enable else 0, ctypes.byref(prev)))
    return bool(prev.value)


def batch_graph_counts() -> tuple[int, int]:
    """(steps captured into a new graph, steps that replayed one) since the process started."""
    c, r = ctypes.c_uint64(0), ctypes.c_uint64(0)
    nat.check(nat.lib().bgt_hip_batch_graph_counts(ctypes.byref(c), ctypes.byref(r)))
    return int(c.value), int(r.value)


def exact_upset_mass(subjects: int, prior: Sequence[float], history: Sequence[tuple[int, int]],
                     experiments, dilution: Optional[DilutionTable] = None) -> "float | np.ndarray":
    """Exact FP64 upset mass(es) of a binary lattice after `history` updates (internal
    coordinates), from the prior and the history alone (no device involved): non-dilution,
    or dilution with the table the updates were applied with. `experiments` is one
    experiment or a sequence of them."""
    pi = np.ascontiguousarray(np.asarray(prior, dtype=np.float64))
    es = np.ascontiguousarray(np.array([e for e, _ in history], dtype=np.uint32))
    rs = np.ascontiguousarray(np.array([r for _, r in history], dtype=np.uint32))
    scalar = np.isscalar(experiments)
    qs = np.ascontiguousarray(np.atleast_1d(np.asarray(experiments, dtype=np.uint32)))
    out = np.empty(qs.size, dtype=np.float64)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    dptr = ctypes.POINTER(ctypes.c_double)
    if dilution is None:
        nat.check(nat.lib().bgt_hip_exact_upset_mass(subjects, pi.ctypes.data_as(dptr), len(history),
                                                     es.ctypes.data_as(u32p), rs.ctypes.data_as(u32p), qs.size,
                                                     qs.ctypes.data_as(u32p), out.ctypes.data_as(dptr)))
    else:
        nat.check(nat.lib().bgt_hip_exact_upset_mass_dilution(
            subjects, pi.ctypes.data_as(dptr), dilution.pointer(), len(history), es.ctypes.data_as(u32p),
            rs.ctypes.data_as(u32p), qs.size, qs.ctypes.data_as(u32p), out.ctypes.data_as(dptr)))
    return float(out[0]) if scalar else out


# ------------------------------------------------------------------------------ simulation

@dataclass
class TreeStats:
    # include/bgt/types.hpp:109-119
    total_leaves: int = 0
    correct_probability: float = 0.0
    incorrect_probability: float = 0.0
    false_positive_probability: float = 0.0
    false_negative_probability: float = 0.0
    unclassified_probability: float = 0.0
    expected_stages: float = 0.0
    expected_tests: float = 0.0

    @classmethod
    def from_c(cls, s: nat.TreeStatsC) -> "TreeStats":
        return cls(*(getattr(s, f) for f, _ in nat.TreeStatsC._fields_))


@dataclass
class SimulationOptions:
    # include/bgt/types.hpp:121-130
    search_depth: int = 1
    threshold_up: float = 0.01
    threshold_lo: float = 0.01
    branch_threshold: float = 0.0
    selector: SelectorType = SelectorType.auto_select
    provider: Provider = Provider.auto_select
    precision: Precision = Precision.float64


@dataclass
class DilutionParameters:
    # include/bgt/simulation.hpp:44-48
    alpha: float = 0.99
    h: float = 0.005


class SimulationMode(enum.IntEnum):
    # include/bgt/simulation.hpp:13-21
    local_tree = 0
    parallel_global_tree = 1
    parallel_dynamic_tree = 2
    parallel_hybrid_tree = 3
    parallel_partial_tree = 4
    parallel_fusion_tree = 5


class TrueStatePolicy(enum.IntEnum):
    # include/bgt/simulation.hpp:23-28
    all = 0
    trimmed = 1
    symmetric = 2


class SimulationRunStatus(enum.IntEnum):
    # include/bgt/simulation.hpp:30-35
    completed = reporting.STATUS_COMPLETED
    interrupted = reporting.STATUS_INTERRUPTED
    running = reporting.STATUS_RUNNING


class TerminationReason(enum.IntEnum):
    # include/bgt/simulation.hpp:37-42
    none = reporting.REASON_NONE
    user_signal = reporting.REASON_USER_SIGNAL
    cancellation_requested = reporting.REASON_CANCELLATION_REQUESTED


@dataclass
class ReportConfig:
    # include/bgt/simulation.hpp:50-55
    write_csv: bool = False
    output_directory: str = "."
    run_name: str = ""


@dataclass
class SimulationTimings:
    # include/bgt/simulation.hpp:57-63
    setup_seconds: float = 0.0
    tree_build_seconds: float = 0.0
    evaluation_seconds: float = 0.0
    total_seconds: float = 0.0


@dataclass
class ProgressConfig:
    # include/bgt/simulation.hpp:65-72
    enabled: bool = False
    interval_seconds: float = 5.0
    output_path: str = ""
    write_jsonl: bool = True
    print_stderr: bool = False


@dataclass
class SimulationConfig:
    # include/bgt/simulation.hpp:87-103. mode, true_state_policy, global_tree_depth,
    # workload_granularity, trim_percent and evaluation_prior configure the reference's
    # multi-rank modes; they are validated here as there, and the local tree ignores them.
    lattice_type: LatticeType = LatticeType.replicated_non_dilution
    subjects: int = 0
    variants: int = 1
    prior: list = field(default_factory=list)
    evaluation_prior: Optional[list] = None
    options: SimulationOptions = field(default_factory=SimulationOptions)
    mode: SimulationMode = SimulationMode.local_tree
    true_state_policy: TrueStatePolicy = TrueStatePolicy.all
    global_tree_depth: int = 0
    workload_granularity: int = 1
    trim_percent: float = 1.0
    dilution: DilutionParameters = field(default_factory=DilutionParameters)
    report: ReportConfig = field(default_factory=ReportConfig)
    progress: ProgressConfig = field(default_factory=ProgressConfig)


@dataclass
class KernelSpec:
    # include/bgt/types.hpp:132-181; the arch string names the GPU target
    provider: Provider = Provider.auto_select
    subjects: int = 0
    variants: int = 1
    state_bits: int = 32
    dilution: DilutionMode = DilutionMode.non_dilution
    selector: SelectorType = SelectorType.auto_select
    precision: Precision = Precision.float64
    accumulator_precision: Precision = Precision.float64
    gpu_arch: str = "gfx950"

    def key(self) -> tuple:
        return (int(self.provider), self.subjects, self.variants, self.state_bits, int(self.dilution),
                int(self.selector), int(self.precision), int(self.accumulator_precision), self.gpu_arch)


@dataclass
class SimulationResult:
    # include/bgt/simulation.hpp:105-121
    stats: TreeStats = field(default_factory=TreeStats)
    kernel: KernelSpec = field(default_factory=KernelSpec)
    mode: SimulationMode = SimulationMode.local_tree
    provider: Provider = Provider.auto_select
    rank: int = 0
    world_size: int = 1
    total_states: int = 0
    evaluated_states: int = 0
    report_path: str = ""
    timings: SimulationTimings = field(default_factory=SimulationTimings)
    status: SimulationRunStatus = SimulationRunStatus.completed
    termination_reason: TerminationReason = TerminationReason.none
    evaluated_prior_mass: float = 0.0
    progress_path: str = ""


@dataclass
class ProgressSnapshot:
    # include/bgt/simulation.hpp:74-85
    status: SimulationRunStatus = SimulationRunStatus.completed
    reason: TerminationReason = TerminationReason.none
    total_states: int = 0
    evaluated_states: int = 0
    state_coverage_fraction: float = 0.0
    evaluated_prior_mass: float = 0.0
    stats: TreeStats = field(default_factory=TreeStats)
    timings: SimulationTimings = field(default_factory=SimulationTimings)
    message: str = ""


def clear_simulation_cancellation() -> None:
    # progress.cpp:87-90
    reporting.clear_cancellation()


def request_simulation_cancellation(reason: TerminationReason = TerminationReason.cancellation_requested) -> None:
    # progress.cpp:92-98: the first reason requested stays until cleared
    reporting.request_cancellation(int(reason))


def simulation_cancellation_requested() -> bool:
    return reporting.cancellation_reason() != reporting.REASON_NONE


def simulation_cancellation_reason() -> TerminationReason:
    return TerminationReason(reporting.cancellation_reason())


def _resolve_selector(selector: SelectorType, variants: int) -> SelectorType:
    # simulation_helpers.cpp:40-45
    if selector == SelectorType.auto_select:
        return SelectorType.op_bha if variants == 1 else SelectorType.op_bbpa
    return selector


def _validate(config: SimulationConfig) -> None:
    # simulation_runner.cpp:30-68
    if config.subjects <= 0 or config.variants <= 0 or config.options.search_depth < 0:
        raise InvalidArgument(nat.BAD_INPUT, "invalid simulation dimensions.")
    if config.workload_granularity <= 0:
        raise InvalidArgument(nat.BAD_INPUT, "workload_granularity must be positive.")
    if config.progress.interval_seconds <= 0.0:
        raise InvalidArgument(nat.BAD_INPUT, "progress interval must be positive.")
    if len(config.prior) != config.subjects * config.variants:
        raise InvalidArgument(nat.BAD_INPUT, "prior length must equal subjects * variants.")
    if config.evaluation_prior is not None and len(config.evaluation_prior) != config.subjects * config.variants:
        raise InvalidArgument(nat.BAD_INPUT, "evaluation_prior length must equal subjects * variants.")
    # simulation_helpers.cpp:34-38 admits auto/op_bha/op_bbpa; brute_force is this provider's
    # addition (the device tree runs it), anything else is refused as there
    if config.options.selector not in (SelectorType.auto_select, SelectorType.op_bha, SelectorType.op_bbpa,
                                       SelectorType.brute_force):
        raise InvalidArgument(nat.BAD_INPUT, "unsupported selector.")
    if _resolve_selector(config.options.selector, config.variants) == SelectorType.op_bha and config.variants != 1:
        raise InvalidArgument(nat.BAD_INPUT, "Op-BHA tree selection is only supported for binary lattices.")
    if _uses_dilution(config.lattice_type) and (config.dilution.alpha <= 0.0 or config.dilution.h <= 0.0):
        raise InvalidArgument(nat.BAD_INPUT, "dilution parameters must be positive.")
    distributed = config.lattice_type in (LatticeType.distributed_dilution, LatticeType.distributed_non_dilution)
    if config.mode == SimulationMode.local_tree and distributed:
        raise InvalidArgument(nat.BAD_INPUT, "local_tree requires a replicated lattice.")
    if config.mode in (SimulationMode.parallel_dynamic_tree, SimulationMode.parallel_hybrid_tree) and distributed:
        raise InvalidArgument(nat.BAD_INPUT,
                              "dynamic and hybrid parallel tree modes currently require replicated lattices.")
    if config.mode == SimulationMode.parallel_partial_tree:
        if not distributed:
            raise InvalidArgument(nat.BAD_INPUT, "partial parallel tree mode requires a distributed lattice.")
        if config.true_state_policy == TrueStatePolicy.all:
            raise InvalidArgument(nat.BAD_INPUT,
                                  "partial parallel tree mode requires trimmed or symmetric true_state_policy.")
    if config.mode != SimulationMode.local_tree:
        raise Unsupported(nat.UNSUPPORTED, f"{reporting.mode_name(config.mode)} runs over several ranks; this "
                          "provider runs the local tree on one GPU")


class Runtime:
    """include/bgt/runtime.hpp:58-78: provider registry keyed by KernelSpec."""

    def __init__(self):
        self._registry: set = set()

    def available_providers(self) -> list[Provider]:
        out = [Provider.cpu]
        if nat.available():
            out.append(Provider.hip)
        return out

    def make_kernel_spec(self, config: SimulationConfig) -> KernelSpec:
        # simulation_runner.cpp:166-179
        return KernelSpec(provider=config.options.provider, subjects=config.subjects, variants=config.variants,
                          dilution=DilutionMode.dilution if _uses_dilution(config.lattice_type)
                          else DilutionMode.non_dilution,
                          selector=_resolve_selector(config.options.selector, config.variants),
                          precision=config.options.precision,
                          accumulator_precision=Precision.float64)

    def registry_size(self) -> int:
        return len(self._registry)

    def run_simulation(self, config: SimulationConfig) -> SimulationResult:
        # simulation_runner.cpp:181-237: the GPU provider's path (progress forced at the
        # start, one device evaluation of every true state, the final snapshot, the report)
        total_start = time.perf_counter()
        _validate(config)
        dil = None
        if _uses_dilution(config.lattice_type):
            dil = DilutionTable(config.subjects, config.dilution.alpha, config.dilution.h)
        spec = self.make_kernel_spec(config)
        self._registry.add(spec.key())
        result = SimulationResult(kernel=spec, mode=config.mode, rank=0, world_size=1,
                                  total_states=1 << (config.subjects * config.variants), evaluated_states=0)

        def snapshot(status, reason, message):
            # progress.cpp:203-221
            t = SimulationTimings(**vars(result.timings))
            t.total_seconds = time.perf_counter() - total_start
            done = progress.evaluated_states
            return ProgressSnapshot(SimulationRunStatus(status), TerminationReason(reason), result.total_states, done,
                                    0.0 if result.total_states == 0 else done / result.total_states,
                                    progress.evaluated_prior_mass,
                                    progress.stats if progress.stats is not None else TreeStats(), t, message)

        progress = reporting.ProgressTracker(config, result.rank, result.total_states, total_start, snapshot)
        eval_start = time.perf_counter()
        if config.options.provider == Provider.cpu:
            raise Unsupported(nat.UNSUPPORTED, "cpu provider is not part of this build")
        if progress.should_stop():
            result.status = SimulationRunStatus.interrupted
            result.termination_reason = simulation_cancellation_reason()
            result.progress_path = progress.progress_path
            result.timings.total_seconds = time.perf_counter() - total_start
            progress.finish(result.status, result.termination_reason, "cancelled before evaluation started")
            result.report_path = reporting.write_report_if_requested(config, result)
            return result
        progress.maybe_emit("starting HIP evaluation", True)
        prior = np.ascontiguousarray(np.asarray(config.prior, dtype=np.float64))
        st = nat.TreeStatsC()
        o = config.options
        nat.check(nat.lib().bgt_hip_provider_run_precision(
            int(_uses_dilution(config.lattice_type)), config.subjects, config.variants,
            prior.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), o.search_depth, o.threshold_up, o.threshold_lo,
            o.branch_threshold, int(o.selector), dil.pointer() if dil is not None else None,
            _storage(o.precision), ctypes.byref(st)))
        result.stats = TreeStats.from_c(st)
        result.provider = Provider.hip
        spec.provider = Provider.hip
        result.evaluated_states = result.total_states
        result.evaluated_prior_mass = 1.0
        progress.replace_state(result.evaluated_states, result.evaluated_prior_mass, TreeStats(**vars(result.stats)))
        result.timings.evaluation_seconds = time.perf_counter() - eval_start
        result.timings.total_seconds = time.perf_counter() - total_start
        result.progress_path = progress.progress_path
        progress.finish(result.status, result.termination_reason, "completed")
        result.report_path = reporting.write_report_if_requested(config, result)
        return result


def run_simulation(config: SimulationConfig) -> SimulationResult:
    return Runtime().run_simulation(config)


def available_providers() -> list[Provider]:
    return Runtime().available_providers()


def hip_available() -> bool:
    return nat.available()
