"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a
driver of every entry point (oracle/sanitize_check.c: fit, MVN, confined,
unconfined, step-capped, dry and from-a-well backtraces, a growing field, the
fixed-grid batch in both math modes on two threads, the math building blocks
at their edge inputs) built with -fsanitize=address,undefined and run with
leak detection; any report aborts it."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle")


def test_oracle_is_clean_under_address_and_undefined_sanitizers(tmp_path):
    if shutil.which("make") is None or shutil.which(os.environ.get("CC", "cc")) is None:
        pytest.skip("no C toolchain")
    build = subprocess.run(["make", "-C", ORACLE, "sanitize"], capture_output=True, text=True,
                           timeout=600)
    if build.returncode != 0 and "sanitize" in build.stderr and "cannot find" in build.stderr:
        pytest.skip("the compiler has no sanitizer runtime: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    run = subprocess.run([os.path.join(ORACLE, "_build", "sanitize_check")], capture_output=True,
                         text=True, timeout=600, env=env, cwd=str(tmp_path))
    assert run.returncode == 0, (run.stdout + run.stderr)[-4000:]
    assert run.stdout.strip() == "ok"
