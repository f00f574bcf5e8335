"""The sigmo_cli driver (the reference's flow, flags and report lines, written against
include/sigmo.hpp) run as a process, its stdout parsed by the line rules of the reference's
experiment parser (scripts/utils/output_analyzer.py:70-111, restated below), and its
'# Matches:' compared with the oracle's totals.

The parser rules restated: lines are stripped; '# Query Nodes', '# Query Graphs', '# Data
Nodes', '# Data Graphs' take the last space-separated field as an int; '# Matches:' and the
'# ... candidates:' lines take the last field with '.' thousands separators removed;
'[*] Refinement step N:' gives N; 'Find all' in the configs block is 'Yes' or 'No'; the GPU and
host stats blocks give 'Total time: T ms' with T the second-to-last field.
"""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

import pytest

from tests._util import GOLDEN

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "sigmo_amd" / "sigmo_cli"


def parse_report(out: str) -> dict:
    rep = {"refinement_steps": 0, "total_times": []}
    for line in (l.strip() for l in out.split("\n")):
        last = line.split(" ")[-1]
        if line.startswith("# Query Nodes"):
            rep["query_nodes"] = int(last)
        elif line.startswith("# Query Graphs"):
            rep["query_graphs"] = int(last)
        elif line.startswith("# Data Nodes"):
            rep["data_nodes"] = int(last)
        elif line.startswith("# Data Graphs"):
            rep["data_graphs"] = int(last)
        elif line.startswith("Find all"):
            rep["find_all"] = last == "Yes"
        elif line.startswith("# Total candidates:"):
            rep["total_candidates"] = int(last.replace(".", ""))
        elif line.startswith("# Zero candidates:"):
            rep["zero_candidates"] = int(last.replace(".", ""))
        elif line.startswith("# Matches:"):
            rep["matches"] = int(last.replace(".", ""))
        elif line.startswith("[*] Refinement step"):
            rep["refinement_steps"] = int(last.replace(":", ""))
        elif line.startswith("Total time:"):
            rep["total_times"].append(int(line.split(" ")[-2]))
    return rep


def run_cli(*args: str) -> str:
    print("sigmo_cli", " ".join(args), flush=True)
    p = subprocess.run([str(CLI), *args], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def test_cli_usage_and_argument_errors():
    """CPU: --help prints the reference's flags and exits 0; a missing data file is an error."""
    p = subprocess.run([str(CLI), "--help"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0
    for flag in ("--iterations", "--find-all", "--skip-join", "--candidates-domain", "--query-filter"):
        assert flag in p.stdout
    p = subprocess.run([str(CLI), "-Q", str(GOLDEN / "test_query.dat")], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0


@pytest.mark.gpu
@pytest.mark.parametrize("find_all", [False, True])
def test_cli_on_reference_fixtures(find_all):
    from sigmo_amd import capi
    from oracle import oracle as O

    qh, dh = capi.HostGraphs.load_dat(GOLDEN / "test_query.dat"), capi.HostGraphs.load_dat(GOLDEN / "test_data.dat")
    o = O.pipeline(qh.arrays(), dh.arrays(), 3, modes=[not find_all])
    args = ["-Q", str(GOLDEN / "test_query.dat"), "-D", str(GOLDEN / "test_data.dat"), "-i", "3"] + (["--find-all"] if find_all else [])
    rep = parse_report(run_cli(*args))
    assert rep["matches"] == o["total_all" if find_all else "total_first"]
    assert rep["query_nodes"] == qh.total_nodes and rep["query_graphs"] == qh.num_graphs
    assert rep["data_nodes"] == dh.total_nodes and rep["data_graphs"] == dh.num_graphs
    assert rep["refinement_steps"] == 3 and rep["find_all"] == find_all
    assert len(rep["total_times"]) == 2


@pytest.mark.gpu
@pytest.mark.parametrize("find_all", [False, True])
def test_cli_on_generated_config0(find_all):
    """2,000 molecules (seed 0) x 20 queries k in U[4,12], R = 6: '# Matches:' equals the committed
    oracle totals of config 0."""
    gold = json.loads((GOLDEN / "config0.json").read_text())
    args = ["--generate", "0:2000", "--generate-queries", "0:20:4:12", "-i", "6"] + (["--find-all"] if find_all else [])
    rep = parse_report(run_cli(*args))
    assert rep["matches"] == gold["matches_all" if find_all else "matches_first"]
    assert rep["data_graphs"] == 2000 and rep["query_graphs"] == 20
    assert rep["refinement_steps"] == 6
