"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/oracle_selftest.cpp drives every oracle entry point over degrees
1..7, both schemes, every volume / surface flux pair, walls and periodic
boundaries, shock capturing off and on and the three initial conditions,
on grids of 1, 2 and 5 elements per side; the sanitizer build
(`make -C oracle sanitize`) aborts on the first out-of-bounds access, use
after free, leak or undefined operation.  The GPU kernels have no sanitizer
on this hardware: they are checked by the oracle comparisons and the
run-twice determinism tests."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
BIN = os.path.join(ORACLE, "_san", "oracle_selftest")


def _sanitizers_available():
    if shutil.which("g++") is None:
        return False
    probe = subprocess.run(["g++", "-fsanitize=address,undefined", "-x", "c++", "-", "-o", os.devnull],
                           input="int main(){return 0;}", capture_output=True, text=True)
    return probe.returncode == 0


@pytest.mark.skipif(not _sanitizers_available(), reason="g++ without ASan/UBSan runtime")
def test_oracle_clean_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([BIN], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 failure(s)" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
