"""The phj_bench command line against the reference binary's observed behaviour
(tests/golden/cli_reference.json, written by tests/golden/make_cli_golden.py
from the reference's own phj-bench): the same options, the same rejections
with the same messages and exit codes, and (on a GPU) the same report and CSV
layout for the same run."""
import json
import os
import re
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(ROOT, "phj-bench_amd", "host", "phj_bench")
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_cli_golden as MG  # noqa: E402

with open(os.path.join(HERE, "golden", "cli_reference.json")) as f:
    REF = json.load(f)


def run(args, timeout=60):
    if not os.path.exists(BIN):
        raise AssertionError(f"{BIN} is not built (make -C phj-bench_amd)")
    return subprocess.run([BIN] + args, capture_output=True, text=True, timeout=timeout)


def test_help_lists_every_reference_option():
    p = run(["--help"])
    assert p.returncode == 0 and p.stderr == ""
    ours = set(re.findall(r"^\s+(--[a-z-]+|-h)\b", p.stdout, re.M))
    assert set(REF["options"]) <= ours
    assert ours - set(REF["options"]) == {"--device"}


@pytest.mark.parametrize("case", REF["errors"], ids=lambda c: " ".join(c["args"]) or "empty")
def test_rejections_match_reference(case):
    p = run(case["args"])
    assert (p.returncode, p.stderr) == (case["rc"], case["stderr"])
    assert p.stdout == ""


def test_accepted_forms_reach_the_device():
    # Parsing succeeds (the same forms the reference accepts, trailing comma
    # included); without a device the run then stops with status 1.
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present: the GPU test covers full runs")
    for args in (["--partition-bits-per-pass", "8,"], ["--build-payload-schema=u32,"], ["--partitions=1024",
                 "--passes=3"], ["--scheme", "phj-pure", "--check", "--csv", "/dev/null"]):
        p = run(args)
        assert p.returncode == 1 and p.stderr.startswith("[error] no usable HIP device"), (args, p.stderr)


@pytest.mark.gpu
def test_run_report_and_csv_match_reference(tmp_path):
    csv = tmp_path / "run.csv"
    args = REF["run"]["args"] + ["--csv", str(csv)]
    p = run(args, timeout=120)
    assert p.returncode == 0, p.stderr
    checks = [line for line in p.stderr.splitlines() if line.startswith("[check]")]
    assert checks == REF["run"]["check_lines"]
    assert MG.summary_tokens(p.stdout) == REF["run"]["summary"]
    header, rows = MG.csv_rows(csv.read_text())
    assert header == REF["run"]["csv_header"]
    assert rows == REF["run"]["csv_rows"]
    # a second run appends rows and writes no second header
    p = run(args, timeout=120)
    assert p.returncode == 0, p.stderr
    lines = csv.read_text().splitlines()
    assert len(lines) == 1 + 2 * len(REF["run"]["csv_rows"]) and lines.count(header) == 1
