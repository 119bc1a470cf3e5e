"""The oracle's C code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

`make -C oracle sanitize` builds the same source with -fsanitize=address,undefined; the
oracle's own test files then run against that build in a child Python started with
LD_PRELOAD=libasan (Python itself is not instrumented, so leak checking is off).  Any
out-of-bounds access, use after free or undefined behaviour in the checker aborts the
child, and the report is shown.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN = ROOT / "oracle" / "_san" / "liboracle_san.so"


def runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if r.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


def sanitized_env():
    asan = runtime("libasan.so")
    if asan is None:
        pytest.skip("libasan not found")
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], capture_output=True, text=True)
    assert r.returncode == 0 and SAN.exists(), r.stdout + r.stderr
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ORACLE_LIB=str(SAN), OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    return env


def test_oracle_tests_under_asan_and_ubsan():
    env = sanitized_env()
    files = ["tests/test_oracle.py", "tests/test_oracle_propagation.py", "tests/test_oracle_krylov.py",
             "tests/test_oracle_adjoint.py", "tests/test_oracle_periodic.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        *files], cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]


def test_sanitizer_catches_a_planted_overrun():
    """The build does check: the leaf run on an input one row short must be reported."""
    env = sanitized_env()
    code = (
        "import sys, numpy as np\n"
        "sys.path[:0] = ['oracle', 'octopus-debian-package_amd', 'tests']\n"
        "import oracle as O\n"
        "from helpers import small_sphere\n"
        "s = small_sphere(radius=2.0, spacing=0.3, natoms=0)\n"
        "fi = np.ones((s.np_part - 1, 4))\n"
        "O.operate_ri_vec(s.op, fi)\n"
        "print('no report')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode != 0 and "heap-buffer-overflow" in out, out[-2000:]
