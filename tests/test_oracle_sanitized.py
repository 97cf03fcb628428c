"""The FP64 oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile, target `asan`; host code only) runs the oracle's CPU tests:
golden fixtures, pins, boundaries, adjoint and registry.  Any out-of-bounds
access, use after free or undefined behaviour in oracle/refcpu.c aborts the
child run.  Skipped where gcc's sanitizer runtimes are absent."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = ["tests/test_oracle_golden.py", "tests/test_oracle_pins.py", "tests/test_oracle_boundaries.py",
         "tests/test_oracle_adjoint.py", "tests/test_materials.py"]


def _runtime(name):
    if not shutil.which("gcc"):
        return None
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_asan_and_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("gcc sanitizer runtimes not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    lib = os.path.join(ROOT, "oracle", "_asan", "librefcpu.so")
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", REFCPU_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    probe = subprocess.run([sys.executable, "-c", "from oracle import refcpu; print(refcpu.lib()._name)"], cwd=ROOT,
                           env=env, capture_output=True, text=True)
    assert probe.returncode == 0 and probe.stdout.strip() == lib, probe.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", *TESTS, "-q", "-x", "-p", "no:cacheprovider"], cwd=ROOT,
                       env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
