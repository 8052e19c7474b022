"""Benchmark: Op-BHA select + update + atom marginals at n = 28 on one MI355X.

Headline workload (one step = one round of the adaptive loop on the 2^28-state binary
non-dilution lattice, priors U[0.005, 0.05] from seed 2): apply the response to the current
experiment, take the atom marginals of the updated posterior and classify, and select the
next experiment by Op-BHA -- the fused two-pass round of the HIP provider
(bgt_hip_lattice_round). Rounds come from seeded adaptive sessions run back to back: session
i draws its true state from the prior with seed 2000 + i and stops, as the reference's
adaptive loop does, once every subject is classified or after 20 tests. Each session starts
from a lattice holding the prior and the prior's Op-BHA selection (the same for every
session), prepared before the timed region; the timed region is exactly `steps` rounds. The
headline value is rounds/s of the compensated FP32 form. The plain FP64 form runs first and
drives the trajectory (its picks and the responses sampled for them); the FP32C form replays
the same (experiment, response) sequence, so the two forms' selections are compared round by
round.

Roofline: the dominant kernel of the FP32C round is pass 2 (the transform's bits 15..27 fused
with the Op-BHA objective): it reads the transform intermediate once, S*b_z bytes per launch
(b = 4 for FP32C, 8 for FP64). Pass 1 (the update fused with bits 0..14) reads the posterior
and writes it and the intermediate, S*(2 b_p + b_z). Both are timed with HIP events on the
lattice's stream over the timed rounds, one bracket per round (pass 1 and pass 2 in turn). A
round moves S*(2 b_p + 2 b_z) bytes; its share of peak is reported beside the kernel's. Before
timing, the autotuner measures the transform forms for this KernelSpec (untimed); the line
names the form it chose.

Also reported, each for both storage forms: the batched tree simulation (n = 16 dilution,
depth 6) to TreeStats, with full-size and with shrinking children, Op-BBPA rounds of the
n = 8, k = 3 dilution lattice, and the serving loop over a mix of adaptive sessions with
Poisson arrivals (p50/p99 latency, sessions/s), with and without shrinking sessions.

CPU baseline: the C restatement of the reference CPU provider (oracle/, built with the
reference's default flags) runs one full n = 28 round on one host thread.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Op-BHA select+update/s @n=28 (HBM GB/s %peak); served p50/p99 ms; FP32c vs FP64"
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_round_n28.json")
MAX_TESTS = 20
THRESHOLD = 0.01


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_times(lat) -> dict:
    from bayesiangrouptesting_amd import _native as nat
    ms = (ctypes.c_double * 16)()
    cnt = (ctypes.c_int * 16)()
    names = ctypes.create_string_buffer(512)
    nat.check(nat.lib().bgt_hip_lattice_kernel_times(lat.handle, -1, ms, cnt, 16, names, 512))
    out = {}
    for i, name in enumerate(names.value.decode().split(";")):
        if name:
            out[name] = (ms[i], cnt[i])
    return out


def enable_clock(lat, mode) -> None:
    """0 off, 1 every kernel class and host phase, 2 light (events around one transform pass per round)."""
    from bayesiangrouptesting_amd import _native as nat
    nat.check(nat.lib().bgt_hip_lattice_kernel_times(lat.handle, int(mode), None, None, 0, None, 0))


def device_sync() -> None:
    from bayesiangrouptesting_amd import _native as nat
    nat.check(nat.lib().bgt_hip_device_synchronize())


class Timer:
    """Wall time of device work in segments, each bracketed by a device synchronisation."""

    def __init__(self):
        self.total = 0.0
        self._t0 = None

    @property
    def running(self) -> bool:
        return self._t0 is not None

    def start(self) -> None:
        device_sync()
        self._t0 = time.perf_counter()

    def stop(self) -> None:
        device_sync()
        self.total += time.perf_counter() - self._t0
        self._t0 = None


def run_sessions(precision, n, pi, steps, warmup, seed, drive=None, clock=2):
    """`warmup` + `steps` rounds of seeded adaptive sessions. Without `drive` each session
    samples responses for its own picks and ends when classified or after MAX_TESTS tests;
    with it, the given (session, experiment, response) rounds are replayed. Returns the
    timed seconds, per-kernel times, the plan that was run, the selection after every round,
    per-round host latencies (ms) of the timed rounds, the classification of every round and
    the last session's lattice."""
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import workload

    total = warmup + steps
    timer = Timer()
    done, clocked = [], []
    plan, picks, lat_ms, classified = [], [], [], []
    lat = None
    sess = -1
    e = 0
    sampler = None
    tests = 0
    for i in range(total):
        if lat is None:
            # a new session: its lattice is prepared outside the timed region
            was = timer.running
            if was:
                timer.stop()
            sess += 1
            lat = bgt.Lattice(bgt.LatticeType.replicated_non_dilution, n, pi, precision=precision)
            first = lat.select(bgt.SelectorType.op_bha)
            e = first.experiment
            s_seed = seed * 1000 + sess
            sampler = workload.ResponseSampler(s_seed, n, 1, bgt.reverse_bits(workload.true_state(s_seed, pi), n))
            tests = 0
            if i >= warmup:
                enable_clock(lat, clock)
                clocked.append(lat)
            if was:
                timer.start()
        if i == warmup:
            if lat not in clocked:
                enable_clock(lat, clock)
                clocked.append(lat)
            timer.start()
        if drive is not None:
            s_drive, e, r = drive[i]
            assert s_drive == sess, "replay out of step with the session plan"
        else:
            r = sampler.draw(e)
        plan.append((sess, e, r))
        t1 = time.perf_counter()
        sel = lat.round(e, r, THRESHOLD, THRESHOLD)
        if i >= warmup:
            lat_ms.append((time.perf_counter() - t1) * 1e3)
        picks.append(sel)
        tests += 1
        classified.append(lat.is_classified())
        if drive is not None:
            end = i + 1 < total and drive[i + 1][0] != sess
        else:
            end = classified[-1] or tests >= MAX_TESTS
        if end and i + 1 < total:
            done.append(lat)
            lat = None
        else:
            e = sel.experiment
    timer.stop()
    # per-kernel averages over the timed rounds of every clocked lattice
    agg: dict = {}
    for l in clocked:
        for name, (ms, cnt) in kernel_times(l).items():
            t, c = agg.get(name, (0.0, 0))
            agg[name] = (t + ms * cnt, c + cnt)
        enable_clock(l, 0)
    kt = {k: (t / c if c else 0.0, c) for k, (t, c) in agg.items()}
    for l in done:
        l.close()
    return {"seconds": timer.total, "kernels": kt, "plan": plan, "picks": picks, "lat_ms": lat_ms,
            "classified": classified, "sessions": sess + 1, "lattice": lat}


PASS2_FORMS = ["two-pass split on blocked z", "three-pass", "two-pass plain on row z"]
PASS1_FORMS = ["persistent with prefetch", "one workgroup per tile"]


def tune_forms(n) -> dict:
    """The autotuner's choice of the FP32C transform's launch forms at this size (measured once,
    untimed): the high-bit form and the pass-1 grid, with the round time of every pair."""
    from bayesiangrouptesting_amd import _native as nat
    p2, p1 = ctypes.c_int(-1), ctypes.c_int(-1)
    ms = (ctypes.c_double * 6)()
    sid = ctypes.create_string_buffer(256)
    nat.check(nat.lib().bgt_hip_tune_forms(0, n, 1, nat.FP32C, 1, 1, ctypes.byref(p2), ctypes.byref(p1), ms,
                                           sid, 256))
    return {"spec": sid.value.decode(), "form": PASS2_FORMS[p2.value], "pass1": PASS1_FORMS[p1.value],
            "round_ms": {f"{PASS2_FORMS[a]} / {PASS1_FORMS[b]}": ms[a * 2 + b] for a in range(3) for b in range(2)}}


def cpu_baseline(n, pi, steps, all_cores: bool = False):
    """Full rounds of the C restatement at n = 28 (select, update, marginals), built with the
    reference's default flags, along the given (internal experiment, response) steps; the
    per-round mean. all_cores: the select with the reference's own OpenMP loops (prior,
    copy, zeta sweeps, argmin; bbpa.cpp:517-571,920-940), the update and the marginals
    serial as in the reference."""
    from oracle import pyoracle as orc
    L = orc.lib("fast")
    # prior built with numpy (construction is not part of a round)
    post = np.ones(1, dtype=np.float64)
    for a in range(n):
        post = np.concatenate([post * pi[a], post * (1.0 - pi[a])])
    upset = np.empty(1 << n, dtype=np.float64)
    masses = np.empty(n, dtype=np.float64)
    dptr = ctypes.POINTER(ctypes.c_double)
    select = L.orc_op_bha_omp if all_cores else L.orc_op_bha
    t = {"select_s": 0.0, "update_s": 0.0, "classify_s": 0.0}
    for e_int, r in steps:
        t0 = time.perf_counter()
        select(post.ctypes.data_as(dptr), n, upset.ctypes.data_as(dptr))
        t1 = time.perf_counter()
        L.orc_update_in_place(0, n, 1, post.ctypes.data_as(dptr), e_int, r, None, 0)
        t2 = time.perf_counter()
        L.orc_atom_masses(post.ctypes.data_as(dptr), n, masses.ctypes.data_as(dptr))
        t3 = time.perf_counter()
        t["select_s"] += t1 - t0
        t["update_s"] += t2 - t1
        t["classify_s"] += t3 - t2
    del post, upset
    out = {k: v / len(steps) for k, v in t.items()}
    out["round_s"] = out["select_s"] + out["update_s"] + out["classify_s"]
    out["threads"] = L.orc_threads() if all_cores else 1
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_traffic():
    """HBM bytes per pass-1 launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


# ----------------------------------------------------------------- other workloads

def tree_workload(repeats: int = 3) -> dict:
    """Batched tree simulation: binary dilution n = 16, DilutionTable(16, 0.05, 0.9), priors
    U[0.01, 0.10], Op-BHA to depth 6 -> TreeStats; best of `repeats` runs per form."""
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import workload
    from oracle import pyoracle as orc

    c = workload.C4
    pi = workload.priors(c.seed, c.subjects, c.prior_lo, c.prior_hi)
    out = {"workload": "binary dilution n=16 k=1, DilutionTable(16,0.05,0.9), priors U[0.01,0.10] seed 4, "
                       "Op-BHA, depth 6, one batched launch per kernel per tree level"}
    stats = {}
    for prec in (bgt.Precision.float64, bgt.Precision.float32_compensated):
        cfg = bgt.SimulationConfig(lattice_type=bgt.LatticeType.replicated_dilution, subjects=c.subjects,
                                   variants=1, prior=pi)
        cfg.options.search_depth = 6
        cfg.options.precision = prec
        cfg.dilution = bgt.DilutionParameters(c.alpha, c.h)
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            st = bgt.run_simulation(cfg).stats
            best = min(best, time.perf_counter() - t0)
        stats[prec] = st
        out["fp64" if prec == bgt.Precision.float64 else "fp32c"] = {"ms": best * 1e3, "leaves": st.total_leaves}
    # the same trees with shrinking children (projections, as the reference's LocalTree)
    prev = bgt.set_tree_shrink(True)
    try:
        for prec in (bgt.Precision.float64, bgt.Precision.float32_compensated):
            cfg = bgt.SimulationConfig(lattice_type=bgt.LatticeType.replicated_dilution, subjects=c.subjects,
                                       variants=1, prior=pi)
            cfg.options.search_depth = 6
            cfg.options.precision = prec
            cfg.dilution = bgt.DilutionParameters(c.alpha, c.h)
            best = float("inf")
            for _ in range(repeats):
                t0 = time.perf_counter()
                st = bgt.run_simulation(cfg).stats
                best = min(best, time.perf_counter() - t0)
            key = "fp64" if prec == bgt.Precision.float64 else "fp32c"
            out[key]["shrinking_ms"] = best * 1e3
            out[key]["shrinking_leaves_equal"] = st.total_leaves == stats[prec].total_leaves
    finally:
        bgt.set_tree_shrink(prev)
    t0 = time.perf_counter()
    ref = orc.run_simulation(True, c.subjects, 1, pi, 6, alpha=c.alpha, h=c.h, variant="fast")
    out["cpu_port_ms"] = (time.perf_counter() - t0) * 1e3
    fields = ["correct_probability", "incorrect_probability", "false_positive_probability",
              "false_negative_probability", "unclassified_probability", "expected_stages", "expected_tests"]
    for prec, key in ((bgt.Precision.float64, "fp64"), (bgt.Precision.float32_compensated, "fp32c")):
        st = stats[prec]
        out[key]["max_abs_diff_vs_cpu"] = max(abs(getattr(st, f) - getattr(ref, f)) for f in fields)
        out[key]["leaves_equal_cpu"] = st.total_leaves == ref.total_leaves
    return out


def shrink_workload(repeats: int = 10) -> dict:
    """Projection of a 2^26-state binary lattice onto 24 of its subjects (two fully classified
    subjects summed out, project_posterior): per form, the median of `repeats` projections
    (host call to completion, the destination from the lattice cache) and its rate in
    algorithmic bytes (source read once, destination written once)."""
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import workload

    n, removed = 26, (3, 17)
    pi = workload.priors(7, n, 0.01, 0.10)
    classified = sum(1 << i for i in removed)
    plan = bgt.make_shrink_plan(n, n, 1, classified, 0)
    out = {"workload": f"binary non-dilution n={n} -> {plan.new_current_subjects} subjects "
                       f"(subjects {removed[0]} and {removed[1]} summed out), priors U[0.01,0.10] seed 7"}
    for prec, key, b in ((bgt.Precision.float64, "fp64", 8), (bgt.Precision.float32_compensated, "fp32c", 4)):
        lat = bgt.Lattice(bgt.LatticeType.replicated_non_dilution, n, pi, precision=prec)
        device_sync()
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            proj = lat.project(plan)
            device_sync()
            times.append(time.perf_counter() - t0)
            proj.close()
        ms = float(np.median(times)) * 1e3
        alg = (1 << n) * b + (1 << plan.base_atom_count) * b
        out[key] = {"ms": ms, "GBs_alg": alg / (ms * 1e-3) / 1e9, "alg_bytes": alg}
        lat.close()
    return out


def graphs_workload(steps: int = 24, lattices: int = 48) -> dict:
    """Batched steps of a fixed set of lattices of uneven shape (binary, n = 10..16 with and
    without dilution, FP32C), every lattice updated and re-selected each step, with the steps
    launched and with the steps replayed from captured hipGraphs (one capture per launch
    structure, then replays). Per-step host-call-to-results time, mean over the steps after the
    first two; the two runs follow the same responses and must select the same experiments."""
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import workload

    shapes = [(10 + (i % 7), (i // 7) % 2 == 1) for i in range(lattices)]
    ND, DL = bgt.LatticeType.replicated_non_dilution, bgt.LatticeType.replicated_dilution
    tables = [bgt.DilutionTable(n, 0.05, 0.9) if d else None for n, d in shapes]
    out = {"workload": f"{lattices} binary lattices, n = 10..16, dilution and not, FP32C; {steps} batched rounds "
                       "(update, classification, Op-BHA select) of every lattice, launched and replayed from hipGraphs"}
    picks = {}
    prev = bgt.set_batch_graphs(False)
    try:
        for use in (False, True):
            lats = [bgt.Lattice(DL if d else ND, n, workload.priors(300 + i, n, 0.005, 0.1),
                                precision=bgt.Precision.float32_compensated) for i, (n, d) in enumerate(shapes)]
            g = workload.SplitMix64(17)
            pending = [None] * lattices
            times, seq = [], []
            c0, r0 = bgt.batch_graph_counts()
            bgt.set_batch_graphs(use)
            for step in range(steps):
                resps = [None if e is None else g.below(2) for e in pending]
                t0 = time.perf_counter()
                sels, _ = bgt.batch_round(lats, pending, resps, bgt.SelectorType.op_bha, THRESHOLD, THRESHOLD, tables)
                times.append(time.perf_counter() - t0)
                pending = [s.experiment for s in sels]
                seq.append(tuple(pending))
            bgt.set_batch_graphs(False)
            c1, r1 = bgt.batch_graph_counts()
            key = "graphs" if use else "launched"
            out[key] = {"step_ms": float(np.mean(times[2:])) * 1e3, "captures": c1 - c0, "replays": r1 - r0}
            picks[key] = seq
            for lat in lats:
                lat.close()
    finally:
        bgt.set_batch_graphs(prev)
    out["identical_selections"] = picks["graphs"] == picks["launched"]
    out["speedup"] = out["launched"]["step_ms"] / out["graphs"]["step_ms"]
    return out


def bbpa_workload(rounds: int = 8, repeats: int = 3) -> dict:
    """Multinomial dilution lattice n = 8, k = 3 (2^24 states), DilutionTable(8, 0.05, 0.9),
    priors U[0.01, 0.10]: `rounds` rounds of Op-BBPA select -> update with responses sampled
    from a seeded true state. A round is the fused lattice round: the update and the
    histogram of the updated posterior in one pass (the feature tables on chip), then the
    tables' zeta and the inclusion-exclusion scoring. Algorithmic bytes per round: the
    posterior read and written once, 2 S b_p. Best of `repeats` runs of the rounds."""
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import workload

    c = workload.C3
    n, k = c.subjects, c.variants
    pi = workload.priors(c.seed, n * k, c.prior_lo, c.prior_hi)
    tab = bgt.DilutionTable(n, c.alpha, c.h)
    out = {"workload": "multinomial dilution n=8 k=3 (2^24 states), DilutionTable(8,0.05,0.9), priors U[0.01,0.10] "
                       "seed 3; round = update with a response sampled from a seeded true state + classification "
                       "+ Op-BBPA select (one histogram pass over the posterior)"}
    picks = {}
    states = 1 << (n * k)
    for prec, key, b in ((bgt.Precision.float64, "fp64", 8), (bgt.Precision.float32_compensated, "fp32c", 4)):
        best = float("inf")
        for _ in range(repeats):
            lat = bgt.Lattice(bgt.LatticeType.replicated_dilution, n, pi, variants=k, precision=prec)
            e = lat.select(bgt.SelectorType.op_bbpa).experiment
            sampler = workload.ResponseSampler(c.seed, n, k, workload.true_state(c.seed, pi), list(tab.values))
            sel = []
            device_sync()
            t0 = time.perf_counter()
            for _ in range(rounds):
                e = lat.round(e, sampler.draw(e), THRESHOLD, THRESHOLD, tab).experiment
                sel.append(e)
            device_sync()
            best = min(best, time.perf_counter() - t0)
            lat.close()
        picks[key] = sel
        ms = best / rounds * 1e3
        alg = 2 * states * b
        out[key] = {"ms_per_round": ms, "alg_bytes_per_round": alg, "round_GBs_alg": alg / (ms * 1e-3) / 1e9,
                    "round_frac": alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
    out["identical_selections"] = sum(a == b for a, b in zip(picks["fp64"], picks["fp32c"]))
    out["selections"] = rounds
    out["note"] = ("2^24 states (128 MiB FP64) stay in the 256 MiB Infinity Cache: cache-bound; reported against "
                   "HBM peak")
    return out


def serving_workload(sessions: int, rate: float, seed: int = 5) -> dict:
    """The serving mix: `sessions` adaptive sessions, n uniform in 10..24, k in {1, 2} with
    n*k <= 26, dilution 50/50, priors U[0.005, 0.10]; each session runs until classified or
    20 tests, with continuous batching and admission by HBM footprint (96 GiB). Two runs per
    form: Poisson arrivals at `rate` per second on the wall clock (latency under a load both
    forms sustain), and every session queued at once (capacity: sessions/s at saturation)."""
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import serving

    out = {"workload": f"{sessions} sessions, n in 10..24, k in {{1,2}}, n*k<=26, dilution 50/50, priors "
                       f"U[0.005,0.10] (seed {seed}), up to 20 tests each; 'at_rate': Poisson arrivals at "
                       f"{rate:g}/s; 'saturated': all sessions queued at time 0"}
    specs = serving.serving_mix(sessions, rate=rate, seed=seed)
    exps, ties = {}, {}
    for prec, key in ((bgt.Precision.float64, "fp64"), (bgt.Precision.float32_compensated, "fp32c")):
        out[key] = {}
        for mode, realtime in (("at_rate", True), ("saturated", False)):
            eng = serving.ServingEngine(precision=prec, hbm_budget=96 << 30)
            rep = eng.run(specs, realtime=realtime)
            results = rep.pop("results")
            if mode == "at_rate":
                exps[key] = {r.sid: r.experiments for r in results}
                ties[key] = {r.sid: (r.near_ties, r.masks) for r in results}
            out[key][mode] = rep
            log(f"serving {key} {mode}: {rep['sessions_per_s']:.1f} sessions/s p50 {rep['latency_ms']['p50']:.1f} ms "
                f"p99 {rep['latency_ms']['p99']:.1f} ms")
    out["identical_sessions"] = sum(exps["fp64"][s] == exps["fp32c"][s] for s in exps["fp64"])
    # each session whose FP32C experiments part from FP64's: the first differing test, and
    # whether either form reported that pick as a near tie (FP64 rounding decides it)
    div = []
    for sid, a in exps["fp64"].items():
        b = exps["fp32c"][sid]
        if a == b:
            continue
        i = next((j for j in range(min(len(a), len(b))) if a[j] != b[j]), min(len(a), len(b)))
        (n64, m64), (n32, m32) = ties["fp64"][sid], ties["fp32c"][sid]
        t64 = i < len(a) and n64[i]
        t32 = i < len(b) and n32[i]
        # a classification that differed first (a mass within FP32C's error of a threshold)
        cls = next((j for j in range(min(len(m64), len(m32), i + 1)) if m64[j] != m32[j]), None)
        div.append({"sid": sid, "first": i, "near_tie_fp64": bool(t64), "near_tie_fp32c": bool(t32),
                    "classification_first": cls is not None})
    out["divergent_sessions"] = len(div)
    out["divergent_at_reported_near_tie"] = sum(d["near_tie_fp64"] or d["near_tie_fp32c"] for d in div)
    out["divergent_after_classification"] = sum(
        d["classification_first"] and not (d["near_tie_fp64"] or d["near_tie_fp32c"]) for d in div)
    out["divergent_unexplained"] = [d for d in div if not (d["near_tie_fp64"] or d["near_tie_fp32c"]
                                                            or d["classification_first"])][:20]
    out["fp32c_vs_fp64_capacity"] = (out["fp32c"]["saturated"]["sessions_per_s"] /
                                     out["fp64"]["saturated"]["sessions_per_s"])
    # shrinking sessions (a session's lattice projected whenever a subject is fully
    # classified, as the reference's trees do): capacity at saturation, both forms
    for prec, key in ((bgt.Precision.float64, "fp64"), (bgt.Precision.float32_compensated, "fp32c")):
        eng = serving.ServingEngine(precision=prec, hbm_budget=96 << 30, shrink=True)
        rep = eng.run(specs, realtime=False)
        results = rep.pop("results")
        rep["sessions_shrunk"] = sum(r.shrinks > 0 for r in results)
        out[key]["saturated_shrinking"] = rep
        log(f"serving {key} saturated, shrinking: {rep['sessions_per_s']:.1f} sessions/s "
            f"p50 {rep['latency_ms']['p50']:.1f} ms p99 {rep['latency_ms']['p99']:.1f} ms")
    return out


def serving28_workload(sessions: int = 32, rate: float = 100.0) -> dict:
    """The serving loop on the headline lattice: `sessions` adaptive sessions of the n = 28
    binary non-dilution configuration (priors U[0.005, 0.05] seed 2, true state of session i
    drawn with seed 2000 + i), each until classified or 20 tests, batched per step; Poisson
    arrivals at `rate` per second and all queued at once, FP64 and FP32C, admission by HBM
    footprint (96 GiB)."""
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import serving, workload

    c = workload.C2
    pi = workload.priors(c.seed, c.subjects, c.prior_lo, c.prior_hi)
    g = workload.SplitMix64(c.seed)
    t = 0.0
    specs = []
    for i in range(sessions):
        t += -np.log(1.0 - g.uniform()) / rate
        specs.append(serving.SessionSpec(i, c.subjects, 1, False, list(pi), c.seed * 1000 + i, t))
    out = {"workload": f"{sessions} sessions of the n=28 binary non-dilution lattice (priors U[0.005,0.05] seed 2), "
                       f"each until classified or 20 tests; 'at_rate': Poisson arrivals at {rate:g}/s; 'saturated': "
                       "all sessions queued at time 0"}
    exps = {}
    for prec, key in ((bgt.Precision.float64, "fp64"), (bgt.Precision.float32_compensated, "fp32c")):
        out[key] = {}
        for mode, realtime in (("at_rate", True), ("saturated", False)):
            eng = serving.ServingEngine(precision=prec, hbm_budget=96 << 30)
            rep = eng.run(specs, realtime=realtime)
            results = rep.pop("results")
            if mode == "saturated":
                exps[key] = {r.sid: r.experiments for r in results}
            out[key][mode] = rep
            log(f"serving n=28 {key} {mode}: {rep['sessions_per_s']:.1f} sessions/s p50 "
                f"{rep['latency_ms']['p50']:.1f} ms p99 {rep['latency_ms']['p99']:.1f} ms")
    out["identical_sessions"] = sum(exps["fp64"][s] == exps["fp32c"][s] for s in exps["fp64"])
    out["fp32c_vs_fp64_capacity"] = (out["fp32c"]["saturated"]["sessions_per_s"] /
                                     out["fp64"]["saturated"]["sessions_per_s"])
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-aux", action="store_true", help="skip the tree, BBPA and serving workloads")
    ap.add_argument("--serving-sessions", type=int, default=4096)
    ap.add_argument("--serving-rate", type=float, default=300.0, help="Poisson arrival rate, sessions/s")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this provider runs one lattice per GPU and does not shard "
              "across devices; run with --gpus 1", file=sys.stderr)
        return 2

    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import workload

    if not bgt.hip_available():
        print("no gfx950 device visible", file=sys.stderr)
        return 1
    c = workload.C2
    n = c.subjects
    pi = workload.priors(c.seed, n, c.prior_lo, c.prior_hi)
    S = 1 << n
    W, K = args.warmup, args.steps

    tuned = tune_forms(n)
    log(f"n={n} states={S} steps={K} warmup={W}; transform form {tuned}")
    r64 = run_sessions(bgt.Precision.float64, n, pi, K, W, c.seed)
    post64 = r64["lattice"].posterior()
    m64 = r64["lattice"].atom_masses()
    r64["lattice"].close()
    log(f"fp64 {r64['seconds'] / K * 1e3:.3f} ms/round over {r64['sessions']} sessions")
    r32 = run_sessions(bgt.Precision.float32_compensated, n, pi, K, W, c.seed, drive=r64["plan"])
    post32 = r32["lattice"].posterior()
    m32 = r32["lattice"].atom_masses()
    r32["lattice"].close()
    log(f"fp32c {r32['seconds'] / K * 1e3:.3f} ms/round")
    # The timed runs bracket only pass 1 with events (the roofline kernel); every other kernel
    # and host phase is timed in a replay of the same rounds with the full clock, whose own
    # event records would otherwise add to the timed rounds.
    d64 = run_sessions(bgt.Precision.float64, n, pi, K, W, c.seed, drive=r64["plan"], clock=1)
    d64["lattice"].close()
    d32 = run_sessions(bgt.Precision.float32_compensated, n, pi, K, W, c.seed, drive=r64["plan"], clock=1)
    d32["lattice"].close()

    pk64, pk32 = r64["picks"], r32["picks"]
    rounds = len(pk64)
    same = sum(1 for a, b in zip(pk64, pk32) if a.experiment == b.experiment)
    resolved = [(a, b) for a, b in zip(pk64, pk32) if not (a.near_tie or b.near_tie)]
    big = post64 >= 1e-30 * post64.sum()
    post_err = float(np.max(np.abs(post32[big] - post64[big]) / post64[big]))
    l1_err = float(np.abs(post32 - post64).sum() / post64.sum())
    mass_err = float(np.max(np.abs(m32 - m64) / np.maximum(m64, 1e-300)))
    del post32, post64, big

    b32, b64 = 4, 8
    kt32, kt64 = d32["kernels"], d64["kernels"]
    # both passes timed in the timed rounds (one bracket per round, in turn)
    p1_32, p2_32 = r32["kernels"]["round_pass1"][0], r32["kernels"]["round_pass2"][0]
    p1_64, p2_64 = r64["kernels"]["round_pass1"][0], r64["kernels"]["round_pass2"][0]
    alg1 = S * (2 * b32 + b32)
    alg2 = S * b32
    round_alg32 = S * (2 * b32 + 2 * b32)
    round_alg64 = S * (2 * b64 + 2 * b64)
    # the dominant kernel: the longer of the two passes
    dom2 = p2_32 >= p1_32
    achieved = (alg2 / (p2_32 * 1e-3) if dom2 else alg1 / (p1_32 * 1e-3)) / 1e9
    traffic = load_traffic()
    dt32, dt64 = r32["seconds"], r64["seconds"]
    value = K / dt32
    lm32, lm64 = r32["lat_ms"], r64["lat_ms"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "rounds/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": dt32 / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32c",
        "data": "synthetic",
        "config": {
            "workload": "binary non-dilution lattice n=28 k=1 (2^28 states), priors U[0.005,0.05] seed 2; "
                        "round = update(response) + atom marginals + classification + Op-BHA select; adaptive "
                        "sessions with true states drawn from the prior (seeds 2000+i), each until classified or "
                        "20 tests; the FP32C run replays the (experiment, response) sequence chosen by the FP64 run",
            "subjects": n, "variants": 1, "states": S, "sessions": r64["sessions"],
            "form": "fp32c (fp32 storage, fp64 on-chip accumulation, exact fp64 re-decision of near ties)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": (traffic.get("pass2_bytes") if dom2 else traffic.get("pass1_bytes")) if traffic else None,
            "kernel": ("k_zeta_pass2<float,3,true> (bits 15..27 of the transform fused with the Op-BHA objective, "
                       "the top-2 and the error band; reads the intermediate once)") if dom2 else
                      "k_zeta_pass1<float,float,true,1> (update fused with bits 0..14 of the transform)",
            "alg_bytes_per_launch": alg2 if dom2 else alg1,
            "pass1_ms": p1_32,
            "pass1_GBs": alg1 / (p1_32 * 1e-3) / 1e9,
            "pass1_frac": alg1 / (p1_32 * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "pass2_ms": p2_32,
            "pass2_GBs": alg2 / (p2_32 * 1e-3) / 1e9,
            "pass2_frac": alg2 / (p2_32 * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "round_kernels_GBs": round_alg32 / ((p1_32 + p2_32) * 1e-3) / 1e9,
            "round_GBs": round_alg32 * value / 1e9,
            "round_frac": round_alg32 * value / 1e9 / PEAK_HBM_GBS,
            "transform_form": tuned,
            "traffic_detail": traffic,
            "timing_note": "pass1_ms / pass2_ms: HIP events around one pass per timed round (pass 1 and pass 2 "
                           "in turn, nothing else bracketed); the kernel/host breakdowns: a replay of the same "
                           "rounds with every phase bracketed",
        },
        "latency_ms": {"p50": float(np.percentile(lm32, 50)), "p99": float(np.percentile(lm32, 99)),
                       "note": "per fused round, host call to result (not the serving loop)"},
        "fp64": {
            "value": K / dt64,
            "ms_per_step": dt64 / K * 1e3,
            "pass1_ms": p1_64, "pass2_ms": p2_64,
            "round_GBs": round_alg64 * (K / dt64) / 1e9,
            "round_frac": round_alg64 * (K / dt64) / 1e9 / PEAK_HBM_GBS,
            "latency_ms": {"p50": float(np.percentile(lm64, 50)), "p99": float(np.percentile(lm64, 99))},
        },
        "fp32c_vs_fp64": {
            "fp32c_speedup": (K / dt32) / (K / dt64),
            "identical_selections": same, "selections": rounds,
            "resolved_selections": len(resolved),
            "identical_resolved": sum(1 for a, b in resolved if a.experiment == b.experiment),
            "near_ties_fp64": sum(1 for p in pk64 if p.near_tie),
            "near_ties_fp32c": sum(1 for p in pk32 if p.near_tie),
            "redecided_fp32c": sum(1 for p in pk32 if p.redecided),
            "band_overflow_fp32c": sum(1 for p in pk32 if p.redecided == 2),
            "rescored_candidates_fp32c_max": max(p.candidates for p in pk32),
            "same_classification": r64["classified"] == r32["classified"],
            "posterior_max_rel_err": post_err, "posterior_l1_rel_err": l1_err, "mass_max_rel_err": mass_err,
            "note": "errors of the last session's final posterior and masses, FP32C against FP64",
        },
        # device kernels (HIP events on the lattice's stream), host phases of a round and
        # error-band sizes, each averaged over the rounds that had them
        "kernels_ms_fp32c": {k: round(v[0], 4) for k, v in kt32.items() if not k.startswith(("n_", "t_", "host_"))},
        "host_ms_fp32c": {k: round(v[0], 4) for k, v in kt32.items() if k.startswith(("t_", "host_"))},
        "band_fp32c": {k[2:]: round(v[0], 1) for k, v in kt32.items() if k.startswith("n_")},
    }
    if not args.no_aux:
        log("tree (n=16 dilution, depth 6)")
        line["tree"] = tree_workload()
        log("bbpa (n=8, k=3)")
        line["bbpa"] = bbpa_workload()
        log("shrink (n=26 -> 24)")
        line["shrink"] = shrink_workload()
        log("batched steps: launched and graph replay")
        line["batch_graphs"] = graphs_workload()
        log(f"serving ({args.serving_sessions} sessions at {args.serving_rate:g}/s)")
        line["serving"] = serving_workload(args.serving_sessions, args.serving_rate)
        log("serving n=28 sessions")
        line["serving_n28"] = serving28_workload()
    if not args.no_cpu:
        # the FP64 run's first two rounds (internal experiment, response)
        steps = [(bgt.reverse_bits(p[1], n), p[2]) for p in r64["plan"][:2]]
        log("cpu baseline (two n=28 rounds, 1 thread)")
        cb = cpu_baseline(n, pi, steps)
        line["cpu_baseline"] = {
            "value": 1.0 / cb["round_s"], "unit": "rounds/s", "cores": 1, "kind": "port",
            "sample": f"two full n=28 rounds of the C restatement (-O3 -ffast-math), per-round mean: select "
                      f"{cb['select_s']:.2f} s, update {cb['update_s']:.2f} s, atom marginals {cb['classify_s']:.2f} s",
            "host_cpu": cpu_model(), "host_threads": os.cpu_count(),
        }
        log("cpu baseline (two n=28 rounds, all cores)")
        cm = cpu_baseline(n, pi, steps, all_cores=True)
        line["cpu_baseline_all_cores"] = {
            "value": 1.0 / cm["round_s"], "unit": "rounds/s", "cores": cm["threads"], "kind": "port",
            "sample": f"the same two rounds with the reference's OpenMP loops (select; update and marginals "
                      f"serial, as in the reference), per-round mean: select {cm['select_s']:.2f} s, update "
                      f"{cm['update_s']:.2f} s, atom marginals {cm['classify_s']:.2f} s",
        }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
