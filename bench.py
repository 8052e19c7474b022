"""Benchmark: Op-BHA select + update + atom marginals at n = 28 on one MI355X.

A step is one round of the adaptive loop on the 2^28-state binary non-dilution lattice
(priors U[0.005, 0.05], seed 2): apply the sampled response to the current experiment,
take the atom marginals of the updated posterior and classify, and select the next
experiment by Op-BHA -- the fused two-pass round of the HIP provider
(bgt_hip_lattice_round). The headline value is rounds/s of the compensated FP32 form;
the plain FP64 form runs the same rounds beside it.

Roofline: algorithmic bytes per round are S*(2*b_p + 2*b_z) (posterior read and
written once, the transform intermediate spilled and reloaded once), with b = 4 for
FP32C and 8 for FP64. The dominant kernel is pass 1 (update fused with the low-bit
transform): S*(2*b_p + b_z) bytes per launch, timed with HIP events on the lattice's
stream over the timed rounds.

CPU baseline: the C restatement of the reference CPU provider, built with the
reference's default flags, runs one full round at n = 28 on one host thread.
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


def run_form(precision, n, pi, truth_pub, steps, warmup, seed):
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import _native as nat
    from bayesiangrouptesting_amd import workload

    lat = bgt.Lattice(bgt.LatticeType.replicated_non_dilution, n, pi, precision=precision)
    sampler = workload.ResponseSampler(seed, n, 1, truth_pub)
    e = lat.select(bgt.SelectorType.op_bha).experiment
    trace = []
    for _ in range(warmup):
        r = sampler.draw(e)
        sel = lat.round(e, r, 0.01, 0.01)
        trace.append((e, r, sel.experiment, sel.near_tie))
        e = sel.experiment
    nat.check(nat.lib().bgt_hip_device_synchronize())
    enable_clock(lat, True)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = sampler.draw(e)
        sel = lat.round(e, r, 0.01, 0.01)
        trace.append((e, r, sel.experiment, sel.near_tie))
        e = sel.experiment
    nat.check(nat.lib().bgt_hip_device_synchronize())
    dt = time.perf_counter() - t0
    kt = kernel_times(lat)
    enable_clock(lat, False)
    return lat, dt, kt, trace


def cpu_baseline(n, pi, e_int, r):
    """One full round of the C restatement at n = 28 (select, update, classify), 1 thread."""
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the provider is single-GPU (see DESIGN.md for how the "
              "lattice would shard); run with --gpus 1", file=sys.stderr)
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

    print(f"[bench] n={n} states={S} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
    lat32, dt32, kt32, tr32 = run_form(bgt.Precision.float32_compensated, n, pi, truth_pub, args.steps,
                                       args.warmup, c.seed)
    post32 = lat32.posterior()
    m32 = lat32.atom_masses()
    lat32.close()
    print(f"[bench] fp32c {dt32 / args.steps * 1e3:.3f} ms/round", file=sys.stderr, flush=True)
    lat64, dt64, kt64, tr64 = run_form(bgt.Precision.float64, n, pi, truth_pub, args.steps, args.warmup, c.seed)
    post64 = lat64.posterior()
    m64 = lat64.atom_masses()
    lat64.close()
    print(f"[bench] fp64 {dt64 / args.steps * 1e3:.3f} ms/round", file=sys.stderr, flush=True)

    same = sum(1 for a, b in zip(tr32, tr64) if a[2] == b[2])
    first_div = next((i for i, (a, b) in enumerate(zip(tr32, tr64)) if a[2] != b[2]), None)
    big = post64 >= 1e-30 * post64.sum()
    post_err = float(np.max(np.abs(post32[big] - post64[big]) / post64[big])) if first_div is None else None
    l1_err = float(np.abs(post32 - post64).sum() / post64.sum()) if first_div is None else None
    mass_err = float(np.max(np.abs(m32 - m64) / np.maximum(m64, 1e-300))) if first_div is None else None
    del post32, post64

    def roof(kt, b):
        p1 = kt.get("round_pass1", (float("nan"), 0))[0]
        p2 = kt.get("round_pass2", (float("nan"), 0))[0]
        alg1 = S * (2 * b + b)
        alg_round = S * (2 * b + 2 * b)
        return {
            "pass1_ms": p1, "pass2_ms": p2,
            "pass1_GBs": alg1 / (p1 * 1e-3) / 1e9,
            "round_alg_bytes": alg_round,
            "round_GBs": alg_round / ((p1 + p2) * 1e-3) / 1e9,
        }

    r32 = roof(kt32, 4)
    r64 = roof(kt64, 8)
    value = args.steps / dt32
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "rounds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32c",
        "data": "synthetic",
        "config": {
            "workload": "binary non-dilution lattice n=28 k=1 (2^28 states), priors U[0.005,0.05] seed 2; "
                        "round = update(sampled response) + atom marginals + classification + Op-BHA select",
            "subjects": n, "variants": 1, "states": S, "form": "fp32c (fp32 storage, fp64 on-chip accumulation)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": r32["pass1_GBs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": r32["pass1_GBs"] / PEAK_HBM_GBS,
            "traffic": None,
            "kernel": "k_zeta_pass1<float,float,true> (update fused with bits 0..14 of the transform)",
            "alg_bytes_per_launch": S * 12,
            "round_GBs": r32["round_GBs"],
            "round_frac": r32["round_GBs"] / PEAK_HBM_GBS,
            "pass1_ms": r32["pass1_ms"],
            "pass2_ms": r32["pass2_ms"],
        },
        "fp64": {
            "value": args.steps / dt64,
            "ms_per_step": dt64 / args.steps * 1e3,
            "pass1_ms": r64["pass1_ms"], "pass2_ms": r64["pass2_ms"],
            "round_GBs": r64["round_GBs"], "round_frac": r64["round_GBs"] / PEAK_HBM_GBS,
            "fp32c_speedup": (args.steps / dt32) / (args.steps / dt64),
        },
        "fp32c_vs_fp64": {
            "identical_selections": same, "rounds": len(tr64), "first_divergence": first_div,
            "posterior_max_rel_err": post_err, "posterior_l1_rel_err": l1_err, "mass_max_rel_err": mass_err,
            "near_ties_fp32c": sum(1 for t in tr32 if t[3]),
        },
    }
    if not args.no_cpu:
        print("[bench] cpu baseline (one n=28 round, 1 thread)", file=sys.stderr, flush=True)
        e_int = bgt.reverse_bits(tr64[0][0], n)
        cb = cpu_baseline(n, pi, e_int, tr64[0][1])
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
