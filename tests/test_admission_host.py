"""Serving admission policy on the CPU (no GPU): the C++ unit test of
accelgraph_express_amd/csrc/ag_admit.h, the policy ag_serve.hip applies, run
against a mock byte counter.  Cases: full capacity when it fits; capacity
lowered in steps of 8; no room below 8 columns; idle batches of other graphs
evicted largest first and only while needed; busy batches never evicted; no
eviction when even evicting every idle batch leaves no room; stable order
among equal batches."""
import os
import subprocess

from conftest import ROOT


def test_admission_policy_with_mock_allocator():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "c"), "_bin/test_admission"],
                   check=True)
    r = subprocess.run([os.path.join(ROOT, "tests", "c", "_bin", "test_admission")],
                       capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 9 and "FAIL" not in r.stdout
