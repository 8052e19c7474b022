"""A serving loop for concurrent adaptive-testing sessions with continuous batching.

Each session is one lattice resident in HBM, running the adaptive loop: select an
experiment, receive its response, update, classify, until every atom is classified or a
test budget is used up. Sessions arrive over time (a Poisson process in the benchmark
workload). The engine admits arrived sessions while their device footprint fits the HBM
budget, and every engine step advances every active session by one round with a single
batched call (bgt_hip_batch_round: sessions of uneven (n, k) share each kernel's launch,
one descriptor upload, one result download). An admitted session's lattice is
built (allocation and prior kernel on its own stream) by a host worker thread while the
current step runs on the GPU, and the session joins the first step after it is ready;
finished sessions free their memory at once, so admitted work keeps flowing rather than
waiting for a batch to drain.

With shrinking (ServingEngine(shrink=True)), a session whose classification covers every
variant of some subject replaces its lattice by the projection onto the remaining subjects
(update_metadata_with_shrinking, lattice_base.cpp:266-316, as the reference's trees do for
every child, global_tree_intra.cpp:8-40), frees the difference of the footprints, and
selects again on the smaller lattice in the next step before its next test. Its experiments
and masks are reported in its original numbering.

Responses come from a seeded simulated lab: each session draws its true state from its
prior and samples responses with the reference's likelihood (the client side of the
loop). A session's selections are identical whether it is served alone or in any mix,
because each lattice runs the same kernels with the same launch shapes either way.
"""
from __future__ import annotations

import concurrent.futures
import dataclasses
import functools
import math
import time
from typing import Optional, Sequence

import numpy as np

from . import api
from . import workload


@dataclasses.dataclass
class SessionSpec:
    sid: int
    subjects: int
    variants: int
    dilution: bool
    prior: list
    seed: int
    arrival: float = 0.0  # seconds after the start of the run


@dataclasses.dataclass
class SessionResult:
    sid: int
    subjects: int
    variants: int
    dilution: bool
    tests: int
    classified: bool
    positive: int
    negative: int
    experiments: list
    arrival: float
    admitted: float
    finished: float
    near_ties: list = dataclasses.field(default_factory=list)  # per experiment: reported near tie
    masks: list = dataclasses.field(default_factory=list)  # per experiment: (positive, negative) before it
    shrinks: int = 0  # projections of the session's lattice (shrinking engines)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def serving_mix(count: int, rate: float, seed: int = 5, lo: float = 0.005, hi: float = 0.10) -> list[SessionSpec]:
    """The serving workload: n uniform in 10..24, k in {1, 2} with n*k <= 26, dilution and
    non-dilution 50/50, priors U[lo, hi], Poisson arrivals at `rate` sessions/s."""
    g = workload.SplitMix64(seed)
    out = []
    t = 0.0
    for sid in range(count):
        while True:
            n = 10 + g.below(15)
            k = 1 + g.below(2)
            if n * k <= 26:
                break
        dil = g.below(2) == 1
        s_seed = g.next_u64()
        prior = workload.priors(s_seed & 0xFFFFFFFF, n * k, lo, hi)
        t += -math.log(1.0 - g.uniform()) / rate if rate > 0 else 0.0
        out.append(SessionSpec(sid, n, k, dil, prior, s_seed & 0xFFFFFFFF, t))
    return out


@functools.lru_cache(maxsize=None)
def _table(subjects: int) -> api.DilutionTable:
    # DilutionTable(s, a, h) holds D[p-1][m] for pool sizes p <= s from (p, m) alone, so the
    # table of a projected lattice is the leading rows of its original's
    return api.DilutionTable(subjects, 0.05, 0.9)


def _to_original(x: int, kept: Sequence[int], current: int, original: int, variants: int) -> int:
    """A set of a projected lattice's atoms (internal numbering, variant-major) in the
    numbering of the lattice it came from; kept[j] is current subject j's original index."""
    out = 0
    for v in range(variants):
        for j, sbj in enumerate(kept):
            if (x >> (v * current + j)) & 1:
                out |= 1 << (v * original + sbj)
    return out


def _kept_after(kept: Sequence[int], plan: api.ShrinkPlan) -> list[int]:
    """The original index of each subject a projection keeps: current subject j leaves when
    its atoms are removable (they are whole subjects, so bit j of variant 0 tells)."""
    return [sbj for j, sbj in enumerate(kept) if not (plan.removable_current_atoms >> j) & 1]


class _Live:
    """An admitted session: its lattice, simulated lab and progress."""

    def __init__(self, spec: SessionSpec, precision: api.Precision, admitted: float, shrink: bool = False):
        self.spec = spec
        t = api.LatticeType.replicated_dilution if spec.dilution else api.LatticeType.replicated_non_dilution
        self.lattice = api.Lattice(t, spec.subjects, spec.prior, variants=spec.variants, precision=precision)
        self.table = _table(spec.subjects) if spec.dilution else None
        truth = workload.true_state(spec.seed, spec.prior)
        if spec.variants == 1:
            truth = api.reverse_bits(truth, spec.subjects)  # public coordinates
        self.sampler = workload.ResponseSampler(spec.seed, spec.subjects, spec.variants, truth,
                                                list(self.table.values) if self.table is not None else None)
        self.admitted = admitted
        self.pending: Optional[int] = None  # experiment selected and not yet applied (lattice's coordinates)
        self.pending_original = 0  # the same in the session's original public coordinates
        self.shrink = shrink
        self.kept = list(range(spec.subjects))  # original (internal) index of each current subject
        self.removed = 0  # subjects summed out (original internal numbering)
        self.pos_internal = self.neg_internal = 0  # classification, original internal atoms
        self._planned = None  # the classification the last shrink plan was made for
        self.shape = (spec.dilution, spec.subjects, spec.variants)  # the current lattice's
        self.tests = 0
        self.experiments: list[int] = []
        self.near_ties: list[bool] = []
        self.masks: list[tuple[int, int]] = []
        self.pos = self.neg = 0
        self.shrinks = 0

    def classified(self) -> bool:
        return bin(self.pos | self.neg).count("1") == self.spec.subjects * self.spec.variants

    def _public(self, x: int) -> int:
        # original internal numbering -> original public coordinates (bit-reversed for k = 1)
        return api.reverse_bits(x, self.spec.subjects) if self.spec.variants == 1 else x

    def absorb(self, p: int, q: int) -> None:
        """A step's classification masks (public coordinates of the current lattice)."""
        if not self.shrink:
            self.pos, self.neg = p, q
            return
        lat, n0, k = self.lattice, self.spec.subjects, self.spec.variants
        nc = len(self.kept)
        self.pos_internal |= _to_original(lat._flip(p), self.kept, nc, n0, k)
        self.neg_internal |= _to_original(lat._flip(q), self.kept, nc, n0, k)
        self.pos, self.neg = self._public(self.pos_internal), self._public(self.neg_internal)

    def original(self, e: int) -> int:
        """A selected experiment (current lattice's public coordinates) in the original ones."""
        if not self.shrink or e == 0xFFFFFFFF:
            return e
        nc = len(self.kept)
        return self._public(_to_original(self.lattice._flip(e), self.kept, nc, self.spec.subjects, 1))

    def shrink_plan(self) -> Optional[api.ShrinkPlan]:
        """make_shrink_plan over the classification so far, when it removes some subject."""
        if not self.shrink:
            return None
        classified = self.pos_internal | self.neg_internal
        if classified == self._planned:
            return None  # nothing classified since the last plan, which removed what it could
        self._planned = classified
        plan = api.make_shrink_plan(self.spec.subjects, len(self.kept), self.spec.variants, classified, self.removed)
        return plan if plan.has_shrink() and plan.new_current_subjects > 0 else None

    def apply(self, plan: api.ShrinkPlan) -> None:
        """The lattice replaced by its projection (project_posterior, shrink.hpp:70-86)."""
        projected = self.lattice.project(plan)
        self.lattice.close()
        self.lattice = projected
        self.kept = _kept_after(self.kept, plan)
        assert len(self.kept) == plan.new_current_subjects
        self.removed = plan.new_classified_subjects
        self.shape = (self.spec.dilution, len(self.kept), self.spec.variants)
        if self.spec.dilution:
            self.table = _table(len(self.kept))
        self.pending = None
        self.shrinks += 1


class ServingEngine:
    def __init__(self, precision: api.Precision = api.Precision.float64, hbm_budget: int = 64 << 30,
                 max_tests: int = 20, threshold_up: float = 0.01, threshold_lo: float = 0.01,
                 max_active: int = 1024, shrink: bool = False):
        self.precision = precision
        self.shrink = shrink
        self.hbm_budget = hbm_budget
        self.max_tests = max_tests
        self.up = threshold_up
        self.lo = threshold_lo
        self.max_active = max_active

    def footprint(self, spec: SessionSpec) -> int:
        return self._footprint((spec.dilution, spec.subjects, spec.variants))

    def _footprint(self, shape: tuple) -> int:
        dil, n, k = shape
        t = api.LatticeType.replicated_dilution if dil else api.LatticeType.replicated_non_dilution
        return api.lattice_footprint(t, n, k, self.precision)

    def run(self, specs: Sequence[SessionSpec], realtime: bool = True) -> dict:
        """Serve `specs` (arrival times relative to the start). With realtime, sessions
        become visible at their arrival times on the wall clock; otherwise all are
        queued at time zero (and latency counts from there). Returns per-session results
        and aggregate statistics.

        Admission overlaps the device work: an admitted session's lattice (allocation and
        prior kernel, on the lattice's own stream) is built by a host worker thread while
        the current batched step runs, and the session joins the first step after it is
        ready. The provider's calls release the interpreter lock, and distinct lattices may
        be used from distinct threads."""
        queue = sorted(specs, key=lambda s: (s.arrival, s.sid))
        qi = 0
        live: list[_Live] = []
        building: list = []  # (future, spec) of admitted sessions whose lattice is being built
        used = 0
        peak_used = 0
        done: list[SessionResult] = []
        step_ms: list[float] = []
        batch_sizes: list[int] = []
        admit_wait_s = 0.0  # time the serving thread waited for admissions with nothing to step
        fp = {}

        def footprint(shape) -> int:
            if shape not in fp:
                fp[shape] = self._footprint(shape)
            return fp[shape]

        t0 = time.perf_counter()
        builder = concurrent.futures.ThreadPoolExecutor(max_workers=1)

        def now() -> float:
            return time.perf_counter() - t0

        try:
            while qi < len(queue) or live or building:
                # admission: arrived sessions, in order, while their footprint fits
                tnow = now()
                while qi < len(queue) and len(live) + len(building) < self.max_active:
                    spec = queue[qi]
                    if realtime and spec.arrival > tnow:
                        break
                    need = footprint((spec.dilution, spec.subjects, spec.variants))
                    if used + need > self.hbm_budget:
                        if not live and not building:
                            raise RuntimeError(f"session {spec.sid} alone exceeds the HBM budget")
                        break
                    building.append((builder.submit(_Live, spec, self.precision, tnow, self.shrink), spec))
                    used += need
                    peak_used = max(peak_used, used)
                    qi += 1
                # sessions whose lattices are ready join this step
                if building:
                    if not live:
                        ta = time.perf_counter()
                        concurrent.futures.wait([f for f, _ in building],
                                                return_when=concurrent.futures.FIRST_COMPLETED)
                        admit_wait_s += time.perf_counter() - ta
                    still = []
                    for fut, spec in building:
                        if fut.done():
                            live.append(fut.result())
                        else:
                            still.append((fut, spec))
                    building = still
                if not live:
                    if qi < len(queue):
                        # idle until the next arrival
                        wait = queue[qi].arrival - now()
                        if wait > 0:
                            time.sleep(min(wait, 0.01))
                    continue
                # one batched round for every active session
                exps, resps = [], []
                for s in live:
                    if s.pending is None:
                        exps.append(None)
                        resps.append(None)
                    else:
                        exps.append(s.pending)
                        resps.append(s.sampler.draw(s.pending_original))
                ts = time.perf_counter()
                sels, masks = api.batch_round([s.lattice for s in live], exps, resps, api.SelectorType.auto_select,
                                              self.up, self.lo, [s.table for s in live])
                step_ms.append((time.perf_counter() - ts) * 1e3)
                batch_sizes.append(len(live))
                tdone = now()
                keep = []
                for s, sel, (p, q), r in zip(live, sels, masks, resps):
                    if r is not None:
                        s.tests += 1
                    s.absorb(p, q)
                    finished = s.classified() or s.tests >= self.max_tests
                    if not finished:
                        plan = s.shrink_plan()
                        if plan is not None:
                            # projected: this step's selection was on the larger lattice, so the
                            # next step selects again before the next test
                            before = s.shape
                            s.apply(plan)
                            used += footprint(s.shape) - footprint(before)
                            keep.append(s)
                            continue
                        s.pending = sel.experiment
                        s.pending_original = s.original(sel.experiment)
                        s.experiments.append(s.pending_original)
                        s.near_ties.append(bool(sel.near_tie))
                        s.masks.append((s.pos, s.neg))
                        keep.append(s)
                        continue
                    spec = s.spec
                    done.append(SessionResult(spec.sid, spec.subjects, spec.variants, spec.dilution, s.tests,
                                              s.classified(), s.pos, s.neg, s.experiments,
                                              spec.arrival if realtime else 0.0, s.admitted,
                                              tdone, s.near_ties, s.masks, s.shrinks))
                    used -= footprint(s.shape)
                    s.lattice.close()
                live = keep
        finally:
            for fut, _ in building:
                fut.cancel()
            builder.shutdown(wait=True)
        wall = now()
        lat = np.array([r.latency for r in done]) * 1e3
        return {
            "sessions": len(done),
            "wall_s": wall,
            "sessions_per_s": len(done) / wall if wall > 0 else float("inf"),
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max())},
            "step_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99)),
                        "steps": len(step_ms), "total_s": float(np.sum(step_ms)) / 1e3},
            "admission_wait_s": admit_wait_s,
            "batch": {"mean": float(np.mean(batch_sizes)), "max": int(np.max(batch_sizes))},
            "tests_mean": float(np.mean([r.tests for r in done])),
            "classified": sum(r.classified for r in done),
            "peak_hbm_bytes": peak_used,
            "results": done,
        }


def serve_alone(spec: SessionSpec, precision: api.Precision = api.Precision.float64, max_tests: int = 20,
                up: float = 0.01, lo: float = 0.01, shrink: bool = False) -> SessionResult:
    """The same session through per-lattice calls, one at a time, in the order of the
    engine's steps: the fused round (update, classification and the next selection) after a
    test, classification and selection otherwise (the first step, and after a projection)."""
    s = _Live(spec, precision, 0.0, shrink)
    while True:
        if s.pending is not None:
            r = s.sampler.draw(s.pending_original)
            sel = s.lattice.round(s.pending, r, up, lo, s.table)
            s.tests += 1
        else:
            s.lattice.update_classification(up, lo)
            sel = None
        s.absorb(s.lattice.positive_classification_mask(), s.lattice.negative_classification_mask())
        if s.classified() or s.tests >= max_tests:
            break
        plan = s.shrink_plan()
        if plan is not None:
            s.apply(plan)
            continue
        if sel is None:
            sel = s.lattice.select(api.SelectorType.auto_select)
        s.pending = sel.experiment
        s.pending_original = s.original(sel.experiment)
        s.experiments.append(s.pending_original)
    out = SessionResult(spec.sid, spec.subjects, spec.variants, spec.dilution, s.tests, s.classified(), s.pos, s.neg,
                        s.experiments, 0.0, 0.0, 0.0, shrinks=s.shrinks)
    s.lattice.close()
    return out
