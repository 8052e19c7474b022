"""bgt_simulate: the reference's command-line entry point (examples/run_simulation.cpp) over
the HIP provider.

    python -m bayesiangrouptesting_amd.simulate --subjects N --prior P[,P...] [options]

The flags, their parsing, the summary line on stdout, the report and progress files and the
signal handling follow run_simulation.cpp:117-298:

* a single prior value is repeated for every atom (run_simulation.cpp:98-114);
* the first SIGINT/SIGTERM requests cancellation (reason user_signal), a second one exits at
  once with 128 + the signal number (run_simulation.cpp:17-26);
* --help prints the usage and returns 0; a missing --subjects or --prior prints it and
  returns 2; a failed run prints "bgt_simulate: <reason>" and returns 1.

Differences, each an addition: --provider also takes "hip" (and "cuda", the reference's name
for its GPU provider, selects the same device provider); --precision chooses the posterior
storage (float64, or float32_compensated: FP32 storage with FP64 accumulation). A malformed
command line is reported as a failed run (return 1) where the reference's parser throws out
of main.
"""
from __future__ import annotations

import os
import signal
import sys

_MODES = ("local_tree", "parallel_global_tree", "parallel_dynamic_tree", "parallel_hybrid_tree",
          "parallel_partial_tree", "parallel_fusion_tree")
_LATTICES = ("replicated_non_dilution", "replicated_dilution", "distributed_non_dilution", "distributed_dilution")
_POLICIES = ("all", "trimmed", "symmetric")


def usage(program: str) -> str:
    # run_simulation.cpp:117-134, plus the provider and precision additions
    return (f"Usage: {program} --subjects N --prior P[,P...] [options]\n"
            "Options:\n"
            "  --variants K\n"
            "  --lattice replicated_non_dilution|replicated_dilution|distributed_non_dilution|distributed_dilution\n"
            "  --mode local_tree|parallel_global_tree|parallel_dynamic_tree|parallel_hybrid_tree|"
            "parallel_partial_tree|parallel_fusion_tree\n"
            "  --search-depth N\n"
            "  --global-tree-depth N\n"
            "  --workload-granularity N\n"
            "  --provider auto|cpu|cuda|hip\n"
            "  --selector auto|op_bha|op_bbpa\n"
            "  --true-state-policy all|trimmed|symmetric\n"
            "  --evaluation-prior P[,P...]\n"
            "  --trim-percent P\n"
            "  --precision float64|float32_compensated\n"
            "  --write-csv --output-dir DIR --run-name NAME\n"
            "  --progress [--progress-interval SECONDS] [--progress-file PATH] [--no-progress-stderr]\n")


def parse_prior(value: str, expected_size: int) -> list:
    # run_simulation.cpp:98-114: comma-separated, empty items skipped, one value broadcast
    prior = [float(item) for item in value.split(",") if item]
    if len(prior) == 1 and expected_size > 1:
        prior = prior * expected_size
    return prior


def _choice(value: str, names, what: str) -> int:
    if value not in names:
        raise ValueError(f"unknown {what}: {value}")
    return names.index(value)


def parse_args(argv):
    """The reference's flag loop (run_simulation.cpp:165-237). Returns (config, prior text,
    evaluation prior text, help requested); the priors are expanded once --variants is known."""
    from . import api as bgt

    config = bgt.SimulationConfig()
    prior_arg = ""
    evaluation_prior_arg = ""
    i = 0

    def value(name):
        nonlocal i
        if i + 1 >= len(argv):
            raise ValueError(f"missing value for {name}")
        i += 1
        return argv[i]

    while i < len(argv):
        key = argv[i]
        if key in ("--help", "-h"):
            return None, "", "", True
        if key == "--subjects":
            config.subjects = int(value(key))
        elif key == "--variants":
            config.variants = int(value(key))
        elif key == "--prior":
            prior_arg = value(key)
        elif key == "--evaluation-prior":
            evaluation_prior_arg = value(key)
        elif key == "--lattice":
            config.lattice_type = bgt.LatticeType[_LATTICES[_choice(value(key), _LATTICES, "lattice")]]
        elif key == "--mode":
            config.mode = bgt.SimulationMode(_choice(value(key), _MODES, "mode"))
        elif key == "--search-depth":
            config.options.search_depth = int(value(key))
        elif key == "--global-tree-depth":
            config.global_tree_depth = int(value(key))
        elif key == "--workload-granularity":
            config.workload_granularity = int(value(key))
        elif key == "--provider":
            name = value(key)
            providers = {"auto": bgt.Provider.auto_select, "cpu": bgt.Provider.cpu, "cuda": bgt.Provider.hip,
                         "hip": bgt.Provider.hip}
            if name not in providers:
                raise ValueError(f"unknown provider: {name}")
            config.options.provider = providers[name]
        elif key == "--selector":
            name = value(key)
            selectors = {"auto": bgt.SelectorType.auto_select, "op_bha": bgt.SelectorType.op_bha,
                         "op_bbpa": bgt.SelectorType.op_bbpa}
            if name not in selectors:
                raise ValueError(f"unknown selector: {name}")
            config.options.selector = selectors[name]
        elif key == "--true-state-policy":
            config.true_state_policy = bgt.TrueStatePolicy(_choice(value(key), _POLICIES, "true-state policy"))
        elif key == "--trim-percent":
            config.trim_percent = float(value(key))
        elif key == "--threshold-up":
            config.options.threshold_up = float(value(key))
        elif key == "--threshold-lo":
            config.options.threshold_lo = float(value(key))
        elif key == "--branch-threshold":
            config.options.branch_threshold = float(value(key))
        elif key == "--dilution-alpha":
            config.dilution.alpha = float(value(key))
        elif key == "--dilution-h":
            config.dilution.h = float(value(key))
        elif key == "--precision":
            name = value(key)
            if name not in ("float64", "float32_compensated"):
                raise ValueError(f"unknown precision: {name}")
            config.options.precision = bgt.Precision[name]
        elif key == "--write-csv":
            config.report.write_csv = True
        elif key == "--output-dir":
            config.report.output_directory = value(key)
        elif key == "--run-name":
            config.report.run_name = value(key)
        elif key == "--progress":
            config.progress.enabled = True
            config.progress.print_stderr = True
        elif key == "--progress-interval":
            config.progress.interval_seconds = float(value(key))
        elif key == "--progress-file":
            config.progress.output_path = value(key)
        elif key == "--no-progress-stderr":
            config.progress.print_stderr = False
        else:
            raise ValueError(f"unknown argument: {key}")
        i += 1
    return config, prior_arg, evaluation_prior_arg, False


def summary_line(result) -> str:
    # run_simulation.cpp:257-266 (enums as their integer values, doubles as "%g")
    return (f"mode={int(result.mode)} provider={int(result.provider)} status={int(result.status)} "
            f"total_states={result.total_states} evaluated_states={result.evaluated_states} "
            f"evaluated_prior_mass={result.evaluated_prior_mass:g} leaves={result.stats.total_leaves} "
            f"correct={result.stats.correct_probability:g} expected_tests={result.stats.expected_tests:g}")


def main(argv=None, program: str = "bgt_simulate") -> int:
    from . import api as bgt

    argv = list(sys.argv[1:] if argv is None else argv)
    try:
        config, prior_arg, evaluation_prior_arg, want_help = parse_args(argv)
    except ValueError as e:
        print(f"bgt_simulate: {e}", file=sys.stderr)
        return 1
    if want_help:
        sys.stderr.write(usage(program))
        return 0
    if config.subjects <= 0 or not prior_arg:
        sys.stderr.write(usage(program))
        return 2
    size = config.subjects * config.variants
    try:
        config.prior = parse_prior(prior_arg, size)
        if evaluation_prior_arg:
            config.evaluation_prior = parse_prior(evaluation_prior_arg, size)
    except ValueError as e:
        print(f"bgt_simulate: {e}", file=sys.stderr)
        return 1

    count = [0]

    def on_signal(signum, _frame):
        count[0] += 1
        if count[0] == 1:
            bgt.request_simulation_cancellation(bgt.TerminationReason.user_signal)
            return
        os._exit(128 + signum)

    previous = {s: signal.getsignal(s) for s in (signal.SIGINT, signal.SIGTERM)}
    try:
        bgt.clear_simulation_cancellation()
        for s in previous:
            signal.signal(s, on_signal)
        result = bgt.run_simulation(config)
        if result.rank == 0:
            print(summary_line(result), flush=True)
            if result.report_path:
                print(f"report={result.report_path}", flush=True)
            if result.progress_path:
                print(f"progress={result.progress_path}", flush=True)
    except Exception as e:  # the reference reports what() of any exception the run throws
        msg = str(e)
        if isinstance(e, bgt.Error) and msg.startswith("[status "):
            msg = msg.split("] ", 1)[-1]
        print(f"bgt_simulate: {msg}", file=sys.stderr)
        return 1
    finally:
        for s, h in previous.items():
            signal.signal(s, h)
    return 0


if __name__ == "__main__":
    sys.exit(main())
