"""bgt_simulate and run_simulation's files on the device provider.

* CMakeLists.txt:519-531: the reference's CLI end-to-end cases; the scalar-prior case must
  report total_leaves,11;
* simulation_progress_test.cpp:32-66: a completed run with progress and a report;
* the report's statistics equal the oracle's TreeStats in the report's own number format.
The CLI runs in this process (main(argv)), not as a child process.
"""
import json
import os

import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import reporting, simulate
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def read_csv(path):
    lines = open(path).read().splitlines()
    assert lines[0] == "field,value"
    return dict(line.split(",", 1) for line in lines[1:])


def test_cli_scalar_prior_golden(tmp_path, capsys):
    print("bgt_simulate scalar prior", flush=True)
    rc = simulate.main(["--subjects", "3", "--variants", "1", "--prior", "0.02", "--lattice", "replicated_non_dilution",
                        "--mode", "local_tree", "--search-depth", "5", "--selector", "op_bbpa", "--threshold-up",
                        "0.01", "--threshold-lo", "0.01", "--branch-threshold", "0.001", "--write-csv",
                        "--output-dir", str(tmp_path), "--run-name", "bgt-simulate-scalar-prior"])
    out = capsys.readouterr().out
    assert rc == 0
    rows = read_csv(tmp_path / "bgt-simulate-scalar-prior.csv")
    assert rows["total_leaves"] == "11"
    assert rows["status"] == "completed" and rows["is_partial"] == "false"
    want = orc.run_simulation(False, 3, 1, [0.02] * 3, 5, 0.01, 0.01, 0.001, selector=2)
    for f in ("correct_probability", "incorrect_probability", "false_positive_probability",
              "false_negative_probability", "unclassified_probability", "expected_stages", "expected_tests"):
        assert rows[f] == reporting.fmt6(getattr(want, f)), f
    first = next(line for line in out.splitlines() if line.startswith("mode="))
    assert first.startswith("mode=0 provider=3 status=0 total_states=8 evaluated_states=8 evaluated_prior_mass=1 "
                            "leaves=11 ")
    assert f"report={tmp_path / 'bgt-simulate-scalar-prior.csv'}" in out


def test_cli_report_e2e(tmp_path, capsys):
    print("bgt_simulate report", flush=True)
    rc = simulate.main(["--subjects", "1", "--prior", "0.2", "--write-csv", "--output-dir", str(tmp_path), "--run-name",
                        "bgt-simulate-e2e"])
    assert rc == 0
    rows = read_csv(tmp_path / "bgt-simulate-e2e.csv")
    assert rows["total_states"] == "2" and rows["evaluated_states"] == "2" and rows["state_coverage_fraction"] == "1"
    assert rows["evaluated_prior_mass"] == "1" and rows["provider"] == str(int(bgt.Provider.hip))
    capsys.readouterr()


def test_completed_run_with_progress(tmp_path):
    # simulation_progress_test.cpp:32-66
    cfg = bgt.SimulationConfig(subjects=2, prior=[0.2, 0.3])
    cfg.options.search_depth = 2
    cfg.report.write_csv = True
    cfg.report.output_directory = str(tmp_path)
    cfg.report.run_name = "completed"
    cfg.progress.enabled = True
    cfg.progress.interval_seconds = 60.0
    bgt.clear_simulation_cancellation()
    print("run_simulation with progress", flush=True)
    r = bgt.run_simulation(cfg)
    assert r.status == bgt.SimulationRunStatus.completed
    assert r.termination_reason == bgt.TerminationReason.none
    assert r.total_states == 4 and r.evaluated_states == 4 and abs(r.evaluated_prior_mass - 1.0) <= 1e-12
    assert r.report_path == str(tmp_path / "completed.csv") and os.path.exists(r.report_path)
    assert r.progress_path == str(tmp_path / "completed.progress.json")
    latest = (tmp_path / "completed.progress.json").read_text()
    assert '"status":"completed"' in latest and '"evaluated_states":4' in latest
    history = (tmp_path / "completed.progress.jsonl").read_text().splitlines()
    assert [json.loads(h)["status"] for h in history] == ["running", "completed"]
    assert json.loads(history[0])["message"] == "starting HIP evaluation"
    final = json.loads(history[1])
    assert final["stats"]["total_leaves"] == r.stats.total_leaves
    assert final["stats"]["expected_tests"] == r.stats.expected_tests
    rows = read_csv(r.report_path)
    assert rows["status"] == "completed" and rows["is_partial"] == "false"
    assert rows["progress_path"] == r.progress_path
    assert r.timings.total_seconds >= r.timings.evaluation_seconds > 0.0


def test_progress_file_override_and_dilution(tmp_path, capsys):
    print("bgt_simulate dilution with --progress-file", flush=True)
    target = tmp_path / "deep" / "p.json"
    rc = simulate.main(["--subjects", "4", "--prior", "0.05,0.1,0.02,0.08", "--lattice", "replicated_dilution",
                        "--search-depth", "3", "--dilution-alpha", "0.05", "--dilution-h", "0.9", "--progress",
                        "--progress-file", str(target), "--precision", "float32_compensated"])
    cap = capsys.readouterr()
    assert rc == 0
    assert "bgt progress: status=completed evaluated=16/16" in cap.err
    assert f"progress={target}" in cap.out
    want = orc.run_simulation(True, 4, 1, [0.05, 0.1, 0.02, 0.08], 3, alpha=0.05, h=0.9)
    d = json.loads(target.read_text())
    assert d["stats"]["total_leaves"] == want.total_leaves
    assert abs(d["stats"]["expected_tests"] - want.expected_tests) <= 1e-9
    assert os.path.exists(tmp_path / "deep" / "p.jsonl")
