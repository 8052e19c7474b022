"""Short workloads for rocprofv3 runs (kernel traces and HBM counters).

    python -m bayesiangrouptesting_amd.profile_workloads round28 [fp32c|fp64] [rounds] [two|three]
        n = 28 binary non-dilution rounds of the benchmark's adaptive sessions (true state of
        session i drawn with seed 2000 + i, a session ends when classified), few rounds; the
        last argument pins the FP32C transform form (default: the tuned or built-in choice)
    python -m bayesiangrouptesting_amd.profile_workloads copy [MiB]
        one device-to-device copy of a known size, the calibration point for FETCH_SIZE /
        WRITE_SIZE (bytes the counters report per byte actually moved)
    python -m bayesiangrouptesting_amd.profile_workloads bbpa24 [fp32c|fp64] [rounds]
        n = 8, k = 3 dilution lattice (2^24 states): fused rounds (update, classification and
        Op-BBPA select from one histogram pass)
    python -m bayesiangrouptesting_amd.profile_workloads lookahead28 [fp32c|fp64] [calls]
        n = 28 binary lattice (the benchmark's priors) after 4 seeded updates: 5-step
        select_experiments calls
    python -m bayesiangrouptesting_amd.profile_workloads serving [fp32c|fp64] [sessions] [full|shrink]
        the serving mix (seed 5) with every session queued at once: batched steps until each
        session is classified or has had 20 tests (shrink: sessions project their lattice when a
        subject is fully classified)
"""
from __future__ import annotations

import ctypes
import sys
import time


def round28(form: str = "fp32c", rounds: int = 4, passes: str = "") -> None:
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import _native as nat
    from bayesiangrouptesting_amd import workload

    if passes:
        nat.check(nat.lib().bgt_hip_tune_force({"two": 0, "three": 1}[passes]))
    c = workload.C2
    n = c.subjects
    pi = workload.priors(c.seed, n, c.prior_lo, c.prior_hi)
    prec = bgt.Precision.float32_compensated if form == "fp32c" else bgt.Precision.float64
    lat = None
    session = -1
    done = 0
    elapsed = 0.0
    while done < rounds:
        if lat is None or lat.is_classified():
            if lat is not None:
                lat.close()
            session += 1
            seed = c.seed * 1000 + session
            lat = bgt.Lattice(bgt.LatticeType.replicated_non_dilution, n, pi, precision=prec)
            sampler = workload.ResponseSampler(seed, n, 1, bgt.reverse_bits(workload.true_state(seed, pi), n))
            e = lat.select(bgt.SelectorType.op_bha).experiment
        t0 = time.perf_counter()
        e = lat.round(e, sampler.draw(e), 0.01, 0.01).experiment
        nat.check(nat.lib().bgt_hip_device_synchronize())
        elapsed += time.perf_counter() - t0
        done += 1
    print(f"round28 {form}: {rounds} rounds over {session + 1} sessions, {elapsed / rounds * 1e3:.3f} ms/round",
          flush=True)
    lat.close()


def bbpa24(form: str = "fp32c", rounds: int = 4) -> None:
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import _native as nat
    from bayesiangrouptesting_amd import workload

    c = workload.C3
    n, k = c.subjects, c.variants
    pi = workload.priors(c.seed, n * k, c.prior_lo, c.prior_hi)
    table = bgt.DilutionTable(n, 0.05, 0.9)
    prec = bgt.Precision.float32_compensated if form == "fp32c" else bgt.Precision.float64
    lat = bgt.Lattice(bgt.LatticeType.replicated_dilution, n, pi, variants=k, precision=prec)
    truth = workload.true_state(c.seed, pi)
    sampler = workload.ResponseSampler(c.seed, n, k, truth, list(table.values))
    e = lat.select(bgt.SelectorType.op_bbpa).experiment
    t0 = time.perf_counter()
    for _ in range(rounds):
        e = lat.round(e, sampler.draw(e), 0.01, 0.01, table).experiment
    nat.check(nat.lib().bgt_hip_device_synchronize())
    print(f"bbpa24 {form}: {rounds} rounds, {(time.perf_counter() - t0) / rounds * 1e3:.3f} ms/round", flush=True)
    lat.close()


def lookahead28(form: str = "fp64", calls: int = 2) -> None:
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import _native as nat
    from bayesiangrouptesting_amd import workload

    c = workload.C2
    n = c.subjects
    pi = workload.priors(c.seed, n, c.prior_lo, c.prior_hi)
    prec = bgt.Precision.float32_compensated if form == "fp32c" else bgt.Precision.float64
    lat = bgt.Lattice(bgt.LatticeType.replicated_non_dilution, n, pi, precision=prec)
    seed = c.seed * 1000
    sampler = workload.ResponseSampler(seed, n, 1, bgt.reverse_bits(workload.true_state(seed, pi), n))
    for _ in range(4):
        e = lat.select(bgt.SelectorType.op_bha).experiment
        lat.update(e, sampler.draw(e))
    for i in range(calls):
        t0 = time.perf_counter()
        picks = lat.select_experiments(5)
        nat.check(nat.lib().bgt_hip_device_synchronize())
        print(f"lookahead28 {form}: call {i}, 5 steps, {(time.perf_counter() - t0) * 1e3:.3f} ms, picks {picks}",
              flush=True)
    lat.close()


def serving_mix(form: str = "fp32c", sessions: int = 512, mode: str = "full") -> None:
    import bayesiangrouptesting_amd as bgt
    from bayesiangrouptesting_amd import serving

    prec = bgt.Precision.float32_compensated if form == "fp32c" else bgt.Precision.float64
    specs = serving.serving_mix(sessions, rate=300.0, seed=5)
    eng = serving.ServingEngine(precision=prec, hbm_budget=96 << 30, shrink=(mode == "shrink"))
    rep = eng.run(specs, realtime=False)
    print(f"serving {form} {mode}: {rep['sessions']} sessions, {rep['sessions_per_s']:.1f} sessions/s, "
          f"{rep['step_ms']['steps']} steps ({rep['step_ms']['total_s'] * 1e3:.1f} ms in steps), "
          f"mean batch {rep['batch']['mean']:.1f}", flush=True)


def copy(mib: int = 1024) -> None:
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    hip_memcpy_d2d = 3
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), hip_memcpy_d2d) == 0
    assert hip.hipDeviceSynchronize() == 0
    print(f"copy: {mib} MiB device to device", flush=True)
    hip.hipFree(a)
    hip.hipFree(b)


def main(argv) -> int:
    if not argv:
        print(__doc__)
        return 2
    what = argv[0]
    if what == "round28":
        round28(argv[1] if len(argv) > 1 else "fp32c", int(argv[2]) if len(argv) > 2 else 4,
                argv[3] if len(argv) > 3 else "")
    elif what == "bbpa24":
        bbpa24(argv[1] if len(argv) > 1 else "fp32c", int(argv[2]) if len(argv) > 2 else 4)
    elif what == "lookahead28":
        lookahead28(argv[1] if len(argv) > 1 else "fp64", int(argv[2]) if len(argv) > 2 else 2)
    elif what == "copy":
        copy(int(argv[1]) if len(argv) > 1 else 1024)
    elif what == "serving":
        serving_mix(argv[1] if len(argv) > 1 else "fp32c", int(argv[2]) if len(argv) > 2 else 512,
                    argv[3] if len(argv) > 3 else "full")
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
