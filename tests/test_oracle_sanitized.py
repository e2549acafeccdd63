"""The oracle's CPU tests again, under AddressSanitizer and
UndefinedBehaviorSanitizer (no GPU): the sanitizer build of
oracle/gfmc_oracle.c is loaded in a child interpreter with the ASan runtime
preloaded; any out-of-bounds access, use after free or undefined operation
aborts the child."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_tests_pass_under_asan_and_ubsan():
    from oracle import build
    lib = build.build_sanitized()
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("the ASan runtime is not installed")
    env = dict(os.environ, PYGFMC_ORACLE_LIB=lib, LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    tests = ["tests/test_oracle.py", "tests/test_diag_oracle.py", "tests/test_offdiag_oracle.py",
             "tests/test_mlp_oracle.py"]
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                          "-p", "no:cacheprovider", *tests], cwd=ROOT, env=env,
                         capture_output=True, text=True, timeout=900)
    tail = (res.stdout + res.stderr)[-3000:]
    assert res.returncode == 0, tail
    assert "passed" in res.stdout, tail
