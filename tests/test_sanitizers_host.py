"""CPU code under AddressSanitizer + UndefinedBehaviorSanitizer (no GPU):
the oracle (every function on small seeded graphs and edge cases), the
library's host-only C++ (MT19937, generators, edge-list I/O on the
reference's fixture files, the CSR file check with malformed arrays), and
the serving admission and extrapolation policy tests.  Built by
tests/c/Makefile with gcc/g++ and -fsanitize=address,undefined."""
import os
import subprocess

import pytest

from conftest import ROOT

CDIR = os.path.join(ROOT, "tests", "c")


@pytest.mark.parametrize("name,args", [
    ("test_oracle_asan", []),
    ("test_host_asan", [os.path.join(ROOT, "tests", "golden", "reference_graphs")]),
    ("test_admission_asan", []),
    ("test_extrapolation_asan", []),
])
def test_under_sanitizers(name, args, tmp_path):
    subprocess.run(["make", "-s", "-C", CDIR, f"_bin/{name}"], check=True)
    extra = [str(tmp_path)] if name == "test_host_asan" else []
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(CDIR, "_bin", name)] + args + extra, capture_output=True,
                       text=True, timeout=300, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "FAIL" not in r.stdout and "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr
