"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
entry point driven on small inputs (oracle/sanitize_check.c, built by
`make -C oracle sanitize`).  CPU only."""
import os
import subprocess

import pytest

from helpers import ROOT


def test_oracle_runs_clean_under_asan_and_ubsan():
    oracle = os.path.join(ROOT, "oracle")
    build = subprocess.run(["make", "-s", "-C", oracle, "sanitize"], capture_output=True, text=True)
    if build.returncode != 0 and "sanitize" in build.stderr and "cannot find" in build.stderr:
        pytest.skip("the compiler has no sanitizer runtime: " + build.stderr[-200:])
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    run = subprocess.run([os.path.join(oracle, "build", "sanitize_check")], capture_output=True, text=True, env=env,
                         timeout=600)
    assert run.returncode == 0, run.stderr[-4000:]
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr, run.stderr[-4000:]
    assert "every entry point ran clean" in run.stdout
