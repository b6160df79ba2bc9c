"""Writes cli_reference.json: how the reference's own benchmark binary
(oracle/_ref/phj-bench, built from the reference's sources by oracle/Makefile)
answers a set of command lines. Inputs and observed outputs only:

  - "options": the option names its --help lists;
  - "errors": command lines it rejects, with exit code and stderr;
  - "run": one small run of every scheme with --check and --csv: the argv, the
    [check] lines, the CSV header, each CSV row with the timing columns replaced
    by "#" (blank columns stay blank), and the stdout summary as token lists
    with the numbers of the scheme tables dropped (wide numbers run into each
    other in the fixed-width cells, so only the labels are comparable).

    python tests/golden/make_cli_golden.py
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "..", "..", "oracle", "_ref", "phj-bench")

ERROR_CASES = [
    ["--scheme", "nope"],
    ["--build-rows", "12x"],
    ["--probe-rows", ""],
    ["--build-payload-schema", "u32,u7"],
    ["--probe-payload-schema", "i64"],
    ["--threads", "0"],
    ["--partitions", "3"],
    ["--partitions", "0"],
    ["--passes", "0"],
    ["--partitions", "16", "--passes", "5"],
    ["--partition-bits-per-pass", "0"],
    ["--partition-bits-per-pass", "17"],
    ["--partition-bits-per-pass", ",8"],
    ["--partition-bits-per-pass", "16,16,1"],
    ["--partitions", "1"],
    ["--reps", "0"],
    ["--seed", "-1"],
    ["--bep-budget-mib", "x"],
    ["--bogus", "1"],
    ["--seed"],
]

RUN_ARGS = ["--scheme", "all", "--reps", "2", "--threads", "2", "--build-rows", "3000", "--probe-rows", "20000",
            "--build-payload-schema", "u32,u128", "--probe-payload-schema", "u8,u64", "--partitions=64",
            "--passes=2", "--bep-budget-mib", "1", "--seed", "7", "--check"]

# CSV columns that carry measurements (they differ from run to run)
TIMED = {"build_wall_ms", "build_ns_per_row", "probe_wall_ms", "probe_ns_per_row", "build_shuffle_wall_ms",
         "build_shuffle_ns_per_row", "probe_shuffle_wall_ms", "probe_shuffle_ns_per_row",
         "eviction_overhead_wall_ms", "eviction_overhead_ns_per_row", "e2e_wall_ms", "peak_mem_mib",
         "bep_evictions", "bep_refinements", "bep_peak_mib", "bep_build_skip_retries"}

NUM = re.compile(r"^-?[\d.]+$")


def summary_tokens(stdout):
    """Lines as token lists; inside the scheme tables numbers are dropped."""
    out, in_tables = [], False
    for line in stdout.splitlines():
        if line.startswith("Scheme:"):
            in_tables = True
        toks = line.split()
        if in_tables:
            toks = [t for t in toks if not NUM.match(t)]
        out.append(toks)
    return out


def csv_rows(text):
    lines = text.splitlines()
    header = lines[0]
    cols = header.split(",")
    rows = []
    for line in lines[1:]:
        vals = split_csv(line)  # the two schema fields are quoted and hold commas
        assert len(vals) == len(cols), (line, len(vals), len(cols))
        rows.append(["#" if c in TIMED and v != "" else v for c, v in zip(cols, vals)])
    return header, rows


def split_csv(line):
    vals, cur, quoted = [], "", False
    for ch in line:
        if ch == '"':
            quoted = not quoted
            cur += ch
        elif ch == "," and not quoted:
            vals.append(cur)
            cur = ""
        else:
            cur += ch
    vals.append(cur)
    return vals


def main():
    assert os.path.exists(REF_BIN), "build oracle/_ref first (make -C oracle ref)"
    help_text = subprocess.run([REF_BIN, "--help"], capture_output=True, text=True, check=True).stdout
    options = sorted(set(re.findall(r"^\s+(--[a-z-]+|-h)\b", help_text, re.M)))
    errors = []
    for args in ERROR_CASES:
        p = subprocess.run([REF_BIN] + args, capture_output=True, text=True, timeout=60)
        errors.append({"args": args, "rc": p.returncode, "stderr": p.stderr})
    with tempfile.TemporaryDirectory() as d:
        csv = os.path.join(d, "run.csv")
        p = subprocess.run([REF_BIN] + RUN_ARGS + ["--csv", csv], capture_output=True, text=True, timeout=300,
                           check=True)
        text = open(csv).read()
    header, rows = csv_rows(text)
    fixture = {
        "options": options,
        "errors": errors,
        "run": {"args": RUN_ARGS, "check_lines": [l for l in p.stderr.splitlines() if l.startswith("[check]")],
                "csv_header": header, "csv_rows": rows, "summary": summary_tokens(p.stdout)},
    }
    with open(os.path.join(HERE, "cli_reference.json"), "w") as f:
        json.dump(fixture, f, indent=1)
    print(f"{len(options)} options, {len(errors)} error cases, {len(rows)} csv rows")


if __name__ == "__main__":
    main()
