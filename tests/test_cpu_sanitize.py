"""Sanitizer runs on the CPU (no GPU needed): the oracle with the engine's host
generator, and the engine library's host code, each built with AddressSanitizer and
UndefinedBehaviorSanitizer by tests/native/Makefile and run with leak detection on.
The drivers cross-check their results, so a run must be both clean and correct."""
import os
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
NATIVE = ROOT / "tests" / "native"


@pytest.fixture(scope="module")
def build_dir(tmp_path_factory):
    out = tmp_path_factory.mktemp("sanitize")
    r = subprocess.run(["make", "-s", "-j8", "-C", str(NATIVE), f"OUT={out}"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.mark.parametrize("prog", ["oracle_sanitize", "host_sanitize"])
def test_sanitized_run_is_clean(build_dir, prog):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([str(build_dir / prog)], capture_output=True, text=True, timeout=300, env=env)
    report = r.stdout[-2000:] + r.stderr[-4000:]
    assert r.returncode == 0, report
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert "all checks passed" in r.stdout
