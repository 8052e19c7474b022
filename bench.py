"""Benchmark: Op-BHA select + update + atom marginals at n = 28 on one MI355X.

A step is one round of the adaptive loop on the 2^28-state binary non-dilution lattice
(priors U[0.005, 0.05], seed 2): apply the response to the current experiment, take the atom
marginals of the updated posterior and classify, and select the next experiment by Op-BHA --
the fused two-pass round of the HIP provider (bgt_hip_lattice_round). The headline value is
rounds/s of the compensated FP32 form. The plain FP64 form runs first and drives the
trajectory (its picks and the responses sampled for them); the FP32C form replays the same
(experiment, response) sequence, so the two forms' selections and final posteriors are
compared round by round.

Roofline: the dominant kernel is pass 1 (update fused with bits 0..14 of the transform):
it reads the posterior and writes it and the transform intermediate, S*(2 b_p + b_z) bytes
per launch (b = 4 for FP32C, 8 for FP64), timed with HIP events on the lattice's stream
over the timed rounds. A round moves S*(2 b_p + 2 b_z) bytes (the intermediate is reloaded
once by pass 2).

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


def enable_clock(lat, on: bool) -> None:
    from bayesiangrouptesting_amd import _native as nat
    nat.check(nat.lib().bgt_hip_lattice_kernel_times(lat.handle, 1 if on else 0, None, None, 0, None, 0))


def run_form(precision, n, pi, truth_pub, steps, warmup, seed, drive=None):
    """Rounds of one storage form. Without `drive` the responses are sampled for the form's
    own picks; with it, the given (experiment, response) sequence is replayed. Returns the
    lattice, the timed seconds, per-kernel times, the selections (initial one first) and the
    per-round host latencies (ms) of the timed rounds."""
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import _native as nat
    from bayesiangrouptesting_amd import workload

    lat = bgt.Lattice(bgt.LatticeType.replicated_non_dilution, n, pi, precision=precision)
    sampler = workload.ResponseSampler(seed, n, 1, truth_pub)
    sel = lat.select(bgt.SelectorType.op_bha)
    picks = [sel]
    e = sel.experiment
    lat_ms = []
    t0 = time.perf_counter()
    for i in range(warmup + steps):
        if i == warmup:
            nat.check(nat.lib().bgt_hip_device_synchronize())
            enable_clock(lat, True)
            t0 = time.perf_counter()
        if drive is not None:
            e, r = drive[i]
        else:
            r = sampler.draw(e)
        t1 = time.perf_counter()
        sel = lat.round(e, r, 0.01, 0.01)
        lat_ms.append((time.perf_counter() - t1) * 1e3)
        picks.append(sel)
        if drive is None:
            e = sel.experiment
    nat.check(nat.lib().bgt_hip_device_synchronize())
    dt = time.perf_counter() - t0
    kt = kernel_times(lat)
    enable_clock(lat, False)
    return lat, dt, kt, picks, lat_ms[warmup:]


def cpu_baseline(n, pi, e_int, r):
    """One full round of the C restatement at n = 28 (select, update, marginals), 1 thread."""
    from oracle import pyoracle as orc
    L = orc.lib("fast")
    # prior built with numpy (construction is not part of a round)
    post = np.ones(1, dtype=np.float64)
    for a in range(n):
        post = np.concatenate([post * pi[a], post * (1.0 - pi[a])])
    upset = np.empty(1 << n, dtype=np.float64)
    dptr = ctypes.POINTER(ctypes.c_double)
    t0 = time.perf_counter()
    L.orc_op_bha(post.ctypes.data_as(dptr), n, upset.ctypes.data_as(dptr))
    t1 = time.perf_counter()
    L.orc_update_in_place(0, n, 1, post.ctypes.data_as(dptr), e_int, r, None, 0)
    t2 = time.perf_counter()
    masses = np.empty(n, dtype=np.float64)
    L.orc_atom_masses(post.ctypes.data_as(dptr), n, masses.ctypes.data_as(dptr))
    t3 = time.perf_counter()
    del post, upset
    return {"select_s": t1 - t0, "update_s": t2 - t1, "classify_s": t3 - t2, "round_s": t3 - t0}


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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this provider runs one lattice per GPU and does not shard "
              "across devices (DESIGN.md describes how it would); run with --gpus 1", file=sys.stderr)
        return 2

    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import workload

    if not bgt.hip_available():
        print("no gfx950 device visible", file=sys.stderr)
        return 1
    c = workload.C2
    n = c.subjects
    pi = workload.priors(c.seed, n, c.prior_lo, c.prior_hi)
    truth_int = workload.true_state(c.seed, pi)
    truth_pub = bgt.reverse_bits(truth_int, n)
    S = 1 << n
    W, K = args.warmup, args.steps

    print(f"[bench] n={n} states={S} steps={K} warmup={W}", file=sys.stderr, flush=True)
    lat64, dt64, kt64, pk64, lm64 = run_form(bgt.Precision.float64, n, pi, truth_pub, K, W, c.seed)
    # the trajectory the FP64 form chose: (experiment applied, response) per round
    sampler = workload.ResponseSampler(c.seed, n, 1, truth_pub)
    drive = []
    for i in range(W + K):
        e = pk64[i].experiment
        drive.append((e, sampler.draw(e)))
    post64 = lat64.posterior()
    m64 = lat64.atom_masses()
    lat64.close()
    print(f"[bench] fp64 {dt64 / K * 1e3:.3f} ms/round", file=sys.stderr, flush=True)
    lat32, dt32, kt32, pk32, lm32 = run_form(bgt.Precision.float32_compensated, n, pi, truth_pub, K, W, c.seed,
                                             drive=drive)
    post32 = lat32.posterior()
    m32 = lat32.atom_masses()
    lat32.close()
    print(f"[bench] fp32c {dt32 / K * 1e3:.3f} ms/round", file=sys.stderr, flush=True)

    rounds = len(pk64)
    same = sum(1 for a, b in zip(pk64, pk32) if a.experiment == b.experiment)
    resolved = [(a, b) for a, b in zip(pk64, pk32) if not (a.near_tie or b.near_tie)]
    big = post64 >= 1e-30 * post64.sum()
    post_err = float(np.max(np.abs(post32[big] - post64[big]) / post64[big]))
    l1_err = float(np.abs(post32 - post64).sum() / post64.sum())
    mass_err = float(np.max(np.abs(m32 - m64) / np.maximum(m64, 1e-300)))
    del post32, post64, big

    b32, b64 = 4, 8
    p1_32 = kt32["round_pass1"][0]
    p2_32 = kt32["round_pass2"][0]
    p1_64 = kt64["round_pass1"][0]
    p2_64 = kt64["round_pass2"][0]
    alg1 = S * (2 * b32 + b32)
    round_alg32 = S * (2 * b32 + 2 * b32)
    round_alg64 = S * (2 * b64 + 2 * b64)
    achieved = alg1 / (p1_32 * 1e-3) / 1e9
    traffic = load_traffic()
    value = K / dt32
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
                        "round = update(response) + atom marginals + classification + Op-BHA select; the FP32C "
                        "run replays the (experiment, response) sequence chosen by the FP64 run",
            "subjects": n, "variants": 1, "states": S,
            "form": "fp32c (fp32 storage, fp64 on-chip accumulation, exact fp64 re-decision of near ties)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic.get("pass1_bytes") if traffic else None,
            "kernel": "k_zeta_pass1<float,float,true,1> (update fused with bits 0..14 of the transform)",
            "alg_bytes_per_launch": alg1,
            "pass1_ms": p1_32,
            "pass2_ms": p2_32,
            "pass2_GBs": S * b32 / (p2_32 * 1e-3) / 1e9,
            "round_kernels_GBs": round_alg32 / ((p1_32 + p2_32) * 1e-3) / 1e9,
            "round_GBs": round_alg32 * value / 1e9,
            "round_frac": round_alg32 * value / 1e9 / PEAK_HBM_GBS,
            "traffic_detail": traffic,
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
            "posterior_max_rel_err": post_err, "posterior_l1_rel_err": l1_err, "mass_max_rel_err": mass_err,
            "near_ties_fp32c": sum(1 for p in pk32 if p.near_tie),
            "redecided_fp32c": sum(1 for p in pk32 if p.redecided),
            "band_overflow_fp32c": sum(1 for p in pk32 if p.redecided == 2),
            "rescored_candidates_fp32c_max": max(p.candidates for p in pk32),
        },
        "kernels_ms_fp32c": {k: round(v[0], 4) for k, v in kt32.items()},
    }
    if not args.no_cpu:
        print("[bench] cpu baseline (one n=28 round, 1 thread)", file=sys.stderr, flush=True)
        e_int = bgt.reverse_bits(drive[0][0], n)
        cb = cpu_baseline(n, pi, e_int, drive[0][1])
        line["cpu_baseline"] = {
            "value": 1.0 / cb["round_s"], "unit": "rounds/s", "cores": 1, "kind": "port",
            "sample": f"one full n=28 round of the C restatement (-O3 -ffast-math): select {cb['select_s']:.2f} s, "
                      f"update {cb['update_s']:.2f} s, atom marginals {cb['classify_s']:.2f} s",
            "host_cpu": cpu_model(), "host_threads": os.cpu_count(),
        }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
