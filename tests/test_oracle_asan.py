"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle asan` builds the same C sources with
-fsanitize=address,undefined (every report aborts).  The CPU tests that drive
the oracle -- known-answer vectors, specification restatements, the search
pins, the loop-filter mask build, border extension, super-resolution and the
LR and CDEF searches, over all their frame geometries -- then run again in a
child process that preloads libasan and loads that build: any out-of-bounds
access, use after free, signed overflow, misaligned load or invalid shift in
the oracle fails this test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_TESTS = [
    "tests/test_oracle_kat.py",
    "tests/test_oracle_spec.py",
    "tests/test_oracle_search_pins.py",
    "tests/test_lfbuild.py",
    "tests/test_border.py",
    "tests/test_superres.py",
    "tests/test_lr_search.py",
    "tests/test_search.py",
]


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_make_asan_builds():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    assert os.path.exists(os.path.join(ROOT, "oracle", "_asan", "liborc.so"))


def test_oracle_tests_pass_under_sanitizers():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc's libasan runtime is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": asan,
        "ORC_LIB": os.path.join(ROOT, "oracle", "_asan", "liborc.so"),
        # Python's own allocations are not the subject: leaks are not reported
        "ASAN_OPTIONS": "detect_leaks=0:verify_asan_link_order=0:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
    })
    r = subprocess.run([sys.executable, "-m", "pytest", *ORACLE_TESTS, "-m", "not gpu", "-q", "-x",
                        "-p", "no:cacheprovider"], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
