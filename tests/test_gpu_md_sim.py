"""The md_sim driver of the GPU build (molecular-dynamics_amd/facade/md_sim, the reference's src/main.cpp over the
facade and md_run) against runs of the reference's own md_sim (tests/golden/md_sim_vectors.json, written by
tests/golden/make_md_sim_golden.py from oracle/_ref/md_sim_ref): in --deterministic mode the RESULT line (all
fields but the timings), the console lines (all but the timing line), the CSV rows, the g(r) file and the XYZ
frames are identical. The reference's bench/benchmark.py parses the RESULT line (benchmark.py:42-52); that parser
is restated in result_fields() and reads this driver's line unchanged.
"""
import hashlib
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "molecular-dynamics_amd", "facade", "md_sim")
GOLDEN = os.path.join(ROOT, "tests", "golden", "md_sim_vectors.json")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

from make_md_sim_golden import RESULT_VOLATILE, run, summarise  # noqa: E402


def result_fields(stdout: str) -> dict:
    """bench/benchmark.py's run_md parsing: the RESULT line's key=value fields, numbers where they parse."""
    line = next(l for l in stdout.splitlines() if l.startswith("RESULT "))
    out = {}
    for field in line.split()[1:]:
        k, v = field.split("=", 1)
        try:
            out[k] = float(v)
        except ValueError:
            out[k] = v
    return out


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)["scenarios"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cfg1", "nvt_melt", "xyz_25", "brute"])
def test_md_sim_matches_reference(golden, name):
    want = golden[name]
    print(f"md_sim {name}: {' '.join(want['args'])}", flush=True)
    console, result, texts = run(DRIVER, want["args"], want["files"])
    got = summarise(console, result, texts)
    assert got["result"] == {k: v for k, v in want["result"].items()} | {"device": "gpu"}
    assert got["console"] == want["console"]
    for kind in want["files"]:
        assert got[kind]["lines"] == want[kind]["lines"], kind
        assert got[kind]["head"] == want[kind]["head"], kind
        assert got[kind]["sha256"] == want[kind]["sha256"], kind


@pytest.mark.gpu
def test_md_sim_result_line_parses_like_the_reference_benchmark(tmp_path):
    proc = subprocess.run([DRIVER, "5", "--quiet", "--equil", "20", "--steps", "60", "--mixed"], capture_output=True,
                          text=True, check=True)
    r = result_fields(proc.stdout)
    for k in ("n", "threads", "dt", "steps", "seconds", "ms_per_step", "mean_T", "std_T", "max_drift", "skin",
              "cutoff", "rebuilds", "deterministic"):
        assert isinstance(r[k], float), k
    assert r["forces"] == "list" and len(r["state_hash"]) == 16 and r["n"] == 500 and r["steps"] == 60
    assert 0.0 < r["mean_T"] < 2.0 and r["max_drift"] < 1e-2


@pytest.mark.parametrize("args", [["--bogus"], ["0"], ["5", "--dt", "x"], ["5", "--steps", "0"], ["--seed", "-1"]])
def test_md_sim_rejects_bad_options(args):
    """The reference's parseOptions rules (main.cpp:55-126): unknown flags, non-numbers and out-of-range values
    print the usage and exit 1, before any GPU work."""
    proc = subprocess.run([DRIVER, *args], capture_output=True, text=True)
    assert proc.returncode == 1 and proc.stderr.startswith("Usage: md_sim")
    assert RESULT_VOLATILE  # the fixture's volatile fields are the timings and the thread count
