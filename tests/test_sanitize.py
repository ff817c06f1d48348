"""The engine's host-side C++ under AddressSanitizer and UBSan (CPU only).

tests/sanitize_driver.cpp includes the planner (dq_plan.hpp), the JSON
decoder, the gate library and the ChaCha8 stream -- all pure C++ -- and runs
them over circuits of the benchmark and parity families for every planner
configuration.  It is built here with g++ -fsanitize=address,undefined and
-fno-sanitize-recover, so any out-of-bounds access, use after free, leak,
signed overflow, bad shift or misaligned load ends the run with a report.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from dqsim_bosonic_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("san") / "sanitize_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", os.path.join(ROOT, "tests", "sanitize_driver.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_planner_and_host_clean_under_sanitizers(driver, tmp_path):
    circuits = {
        "random_blocks_30q_d4": W.random_blocks(30, 4, seed=42, measure_every=2),
        "random_blocks_16q_d12": W.random_blocks(16, 12, seed=3, remote_per_kind=4, measure_every=4),
        "brickwork_20q_d6": W.brickwork_u3(20, 6, seed=42),
        "mixed_13q": W.random_mixed(13, 300, seed=7),
        "mixed_9q": W.random_mixed(9, 200, seed=8),
        "hea_14q_d3": W.hea(14, 3, [[[0.3 * q + l, 0.7 * l - q] for q in range(14)] for l in range(3)]),
    }
    files = []
    for name, c in circuits.items():
        p = tmp_path / f"{name}.json"
        p.write_text(c.model_dump_json())
        files.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(driver), *files], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert r.stdout.count("qubits") == len(files)
