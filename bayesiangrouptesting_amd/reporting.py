"""Wire and disk formats of a simulation run: the CSV report, the progress snapshots
(latest JSON, JSONL history, stderr line) and the process-wide cancellation flag.

Field names, field order and number formatting follow the reference so its tooling can
read these files unchanged:

* CSV report: src/runtime/reporting.cpp:10-46. Doubles are written with the stream's
  default format (6 significant digits, the C "%g" form), booleans as true/false.
* progress JSON: src/runtime/progress.cpp:46-80 (doubles with 17 significant digits,
  the C "%.17g" form), written to a temporary file and renamed over the latest snapshot
  (progress.cpp:241-262); each snapshot is also appended to the .jsonl history.
* stderr line: progress.cpp:265-278.
* cancellation: progress.cpp:87-113 (the first reason requested wins until cleared).
"""
from __future__ import annotations

import os
import sys
import threading
import time

# include/bgt/simulation.hpp:13-42, in declaration order
MODE_NAMES = ("local_tree", "parallel_global_tree", "parallel_dynamic_tree", "parallel_hybrid_tree",
              "parallel_partial_tree", "parallel_fusion_tree")
STATUS_NAMES = ("completed", "interrupted", "running")
REASON_NAMES = ("none", "user_signal", "cancellation_requested")
STATUS_COMPLETED, STATUS_INTERRUPTED, STATUS_RUNNING = 0, 1, 2
REASON_NONE, REASON_USER_SIGNAL, REASON_CANCELLATION_REQUESTED = 0, 1, 2

_cancel_lock = threading.Lock()
_cancel_reason = REASON_NONE


def clear_cancellation() -> None:
    global _cancel_reason
    with _cancel_lock:
        _cancel_reason = REASON_NONE


def request_cancellation(reason: int = REASON_CANCELLATION_REQUESTED) -> None:
    global _cancel_reason
    if int(reason) == REASON_NONE:
        reason = REASON_CANCELLATION_REQUESTED
    with _cancel_lock:
        if _cancel_reason == REASON_NONE:
            _cancel_reason = int(reason)


def cancellation_reason() -> int:
    with _cancel_lock:
        return _cancel_reason


def mode_name(mode: int) -> str:
    return MODE_NAMES[int(mode)] if 0 <= int(mode) < len(MODE_NAMES) else "unknown"


def status_name(status: int) -> str:
    return STATUS_NAMES[int(status)] if 0 <= int(status) < len(STATUS_NAMES) else "unknown"


def reason_name(reason: int) -> str:
    return REASON_NAMES[int(reason)] if 0 <= int(reason) < len(REASON_NAMES) else "unknown"


def fmt6(v: float) -> str:
    """A double as an ostream with default flags writes it."""
    return "%g" % v


def fmt17(v: float) -> str:
    """A double as an ostream with setprecision(17) writes it."""
    return "%.17g" % v


def json_escape(text: str) -> str:
    # progress.cpp:16-44: backslash, quote, newline, carriage return and tab only
    out = []
    for ch in text:
        out.append({"\\": "\\\\", '"': '\\"', "\n": "\\n", "\r": "\\r", "\t": "\\t"}.get(ch, ch))
    return "".join(out)


def run_name(config) -> str:
    return config.report.run_name if config.report.run_name else mode_name(config.mode)


def report_text(config, result) -> str:
    """The CSV report body (reporting.cpp:21-44)."""
    s, t = result.stats, result.timings
    partial = int(result.status) == STATUS_INTERRUPTED
    coverage = 0.0 if result.total_states == 0 else result.evaluated_states / result.total_states
    rows = [
        ("status", status_name(result.status)),
        ("termination_reason", reason_name(result.termination_reason)),
        ("is_partial", "true" if partial else "false"),
        ("mode", mode_name(result.mode)),
        ("provider", str(int(result.provider))),
        ("subjects", str(config.subjects)),
        ("variants", str(config.variants)),
        ("world_size", str(result.world_size)),
        ("total_states", str(result.total_states)),
        ("evaluated_states", str(result.evaluated_states)),
        ("state_coverage_fraction", fmt6(coverage)),
        ("evaluated_prior_mass", fmt6(result.evaluated_prior_mass)),
        ("progress_path", result.progress_path),
        ("total_leaves", str(s.total_leaves)),
        ("correct_probability", fmt6(s.correct_probability)),
        ("incorrect_probability", fmt6(s.incorrect_probability)),
        ("false_positive_probability", fmt6(s.false_positive_probability)),
        ("false_negative_probability", fmt6(s.false_negative_probability)),
        ("unclassified_probability", fmt6(s.unclassified_probability)),
        ("expected_stages", fmt6(s.expected_stages)),
        ("expected_tests", fmt6(s.expected_tests)),
        ("setup_seconds", fmt6(t.setup_seconds)),
        ("tree_build_seconds", fmt6(t.tree_build_seconds)),
        ("evaluation_seconds", fmt6(t.evaluation_seconds)),
        ("total_seconds", fmt6(t.total_seconds)),
    ]
    return "field,value\n" + "".join(f"{k},{v}\n" for k, v in rows)


def write_report_if_requested(config, result) -> str:
    """reporting.cpp:10-46: <output_directory>/<run name>.csv, or .partial.csv when the run
    was interrupted; only rank 0 writes. Returns the path written, or ""."""
    if not config.report.write_csv or result.rank != 0:
        return ""
    os.makedirs(config.report.output_directory, exist_ok=True)
    suffix = ".partial.csv" if int(result.status) == STATUS_INTERRUPTED else ".csv"
    path = os.path.join(str(config.report.output_directory), run_name(config) + suffix)
    with open(path, "w") as f:
        f.write(report_text(config, result))
    return path


def snapshot_json(snap) -> str:
    """One progress snapshot as the reference's single-line JSON (progress.cpp:46-80)."""
    s, t = snap.stats, snap.timings
    return ("{"
            f"\"status\":\"{status_name(snap.status)}\","
            f"\"termination_reason\":\"{reason_name(snap.reason)}\","
            f"\"total_states\":{snap.total_states},"
            f"\"evaluated_states\":{snap.evaluated_states},"
            f"\"state_coverage_fraction\":{fmt17(snap.state_coverage_fraction)},"
            f"\"evaluated_prior_mass\":{fmt17(snap.evaluated_prior_mass)},"
            "\"stats\":{"
            f"\"total_leaves\":{s.total_leaves},"
            f"\"correct_probability\":{fmt17(s.correct_probability)},"
            f"\"incorrect_probability\":{fmt17(s.incorrect_probability)},"
            f"\"false_positive_probability\":{fmt17(s.false_positive_probability)},"
            f"\"false_negative_probability\":{fmt17(s.false_negative_probability)},"
            f"\"unclassified_probability\":{fmt17(s.unclassified_probability)},"
            f"\"expected_stages\":{fmt17(s.expected_stages)},"
            f"\"expected_tests\":{fmt17(s.expected_tests)}"
            "},"
            "\"timings\":{"
            f"\"setup_seconds\":{fmt17(t.setup_seconds)},"
            f"\"tree_build_seconds\":{fmt17(t.tree_build_seconds)},"
            f"\"evaluation_seconds\":{fmt17(t.evaluation_seconds)},"
            f"\"total_seconds\":{fmt17(t.total_seconds)}"
            "},"
            f"\"message\":\"{json_escape(snap.message)}\""
            "}")


def stderr_line(snap) -> str:
    line = (f"bgt progress: status={status_name(snap.status)} evaluated={snap.evaluated_states}/"
            f"{snap.total_states} coverage={fmt6(snap.state_coverage_fraction)} "
            f"prior_mass={fmt6(snap.evaluated_prior_mass)} elapsed={fmt6(snap.timings.total_seconds)}s")
    if snap.message:
        line += " " + snap.message
    return line


class ProgressTracker:
    """progress.cpp:116-279 for one process (rank 0). `make_snapshot(status, reason,
    message)` is supplied by the runtime, which owns the snapshot type."""

    def __init__(self, config, rank: int, total_states: int, total_start: float, make_snapshot):
        self._config = config
        self._rank = rank
        self.total_states = total_states
        self._total_start = total_start
        self._last_emit = total_start
        self._make = make_snapshot
        self.evaluated_states = 0
        self.evaluated_prior_mass = 0.0
        self.stats = None
        self.timings = None
        self.progress_path = self.latest_path() if self.enabled() else ""

    def enabled(self) -> bool:
        return bool(self._config.progress.enabled) and self._rank == 0

    def latest_path(self) -> str:
        p = self._config.progress
        if p.output_path:
            return str(p.output_path)
        return os.path.join(str(self._config.report.output_directory), run_name(self._config) + ".progress.json")

    def history_path(self) -> str:
        root, _ext = os.path.splitext(self.latest_path())
        return root + ".jsonl"

    def elapsed(self) -> float:
        return time.perf_counter() - self._total_start

    def replace_state(self, evaluated_states: int, prior_mass: float, stats) -> None:
        known = self.stats.total_leaves if self.stats is not None else 0
        self.evaluated_states = evaluated_states
        self.evaluated_prior_mass = prior_mass
        self.stats = stats
        if self.stats.total_leaves == 0:
            self.stats.total_leaves = known

    def should_stop(self) -> bool:
        return cancellation_reason() != REASON_NONE

    def maybe_emit(self, message: str = "", force: bool = False) -> bool:
        if not self.enabled():
            return False
        now = time.perf_counter()
        if not force and now - self._last_emit < self._config.progress.interval_seconds:
            return False
        reason = cancellation_reason()
        status = STATUS_RUNNING if reason == REASON_NONE else STATUS_INTERRUPTED
        self._write(self._make(status, reason, message))
        self._last_emit = now
        return True

    def finish(self, status: int, reason: int, message: str) -> None:
        if self.enabled():
            self._write(self._make(status, reason, message))

    def _write(self, snap) -> None:
        latest = self.latest_path()
        parent = os.path.dirname(latest)
        if parent:
            os.makedirs(parent, exist_ok=True)
        body = snapshot_json(snap)
        tmp = latest + ".tmp"
        with open(tmp, "w") as f:
            f.write(body + "\n")
        os.replace(tmp, latest)
        if self._config.progress.write_jsonl:
            history = self.history_path()
            hp = os.path.dirname(history)
            if hp:
                os.makedirs(hp, exist_ok=True)
            with open(history, "a") as f:
                f.write(body + "\n")
        if self._config.progress.print_stderr:
            print(stderr_line(snap), file=sys.stderr, flush=True)
