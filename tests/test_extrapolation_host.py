"""PPR extrapolation rule on the CPU (no GPU): the C++ unit test of
accelgraph_express_amd/csrc/ag_extrap.h, the rule both the batched and the
single-vector sweeps apply.  Cases: a steady residual ratio fires on the third
plain sweep with weight rho / (1 - rho); ratios outside [0.2, 0.95] and ratios
that wander by more than 2% never fire; a stepped sweep resets the state and
the next step waits three plain sweeps; and on a three-mode linear iteration
the steps halve the sweep count while both stops stay within the contraction
bound of a plain-sweep residual."""
import os
import subprocess

from conftest import ROOT


def test_extrapolation_rule():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "c"),
                    "_bin/test_extrapolation"], check=True)
    r = subprocess.run([os.path.join(ROOT, "tests", "c", "_bin", "test_extrapolation")],
                       capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 10 and "FAIL" not in r.stdout
