"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer (gcc).

oracle.c is built with ``-fsanitize=address,undefined -fno-sanitize-recover=all`` into a
temporary directory, and the CPU oracle tests (the golden vectors, the closed-form and
element-scan pins, the spectra, windows, F-numbers and differentiation checks) run again
in a child process that preloads gcc's libasan and loads that build through
``FN_ORACLE_LIB``.  Any invalid access or undefined operation in oracle.c aborts the
child.  Leak detection is off: the Python interpreter keeps its own allocations to the
end.  A build with a planted write past the end of an output (``OR_PLANT_OOB``) shows
that the instrumented build reports such a mistake.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
SAN_CFLAGS = ("-O1 -g -std=c99 -fPIC -fopenmp -ffp-contract=off -Wall -Wextra -Wno-unused-parameter "
              "-fsanitize=address,undefined -fno-sanitize-recover=all -fno-omit-frame-pointer")
ORACLE_TESTS = ["tests/test_golden.py", "tests/test_oracle_aperture.py", "tests/test_oracle_das.py",
                "tests/test_oracle_diff.py", "tests/test_oracle_fnumbers.py", "tests/test_oracle_grid.py",
                "tests/test_oracle_spectra.py", "tests/test_oracle_windows.py"]


def _libasan():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    if r.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("gcc's libasan is not installed")
    return path


def _build(build_dir, plant=""):
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, f"BUILD={build_dir}", f"PLANT={plant}",
                    f"CFLAGS={SAN_CFLAGS}"], check=True)
    return os.path.join(build_dir, "liboracle.so")


def _env(lib):
    env = dict(os.environ)
    env.update(FN_ORACLE_LIB=lib, LD_PRELOAD=_libasan(), ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    return env


def test_oracle_tests_pass_under_sanitizers(tmp_path):
    lib = _build(str(tmp_path / "san"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        *ORACLE_TESTS], cwd=ROOT, env=_env(lib), capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail


def test_sanitized_build_reports_a_planted_overflow(tmp_path):
    lib = _build(str(tmp_path / "oob"), "-DOR_PLANT_OOB")
    code = ("import numpy as np\n"
            "from oracle import oracle as O\n"
            "O.fnumber_values(O.fnumber('constant', 1.5), np.linspace(0.5, 2.0, 37))\n"
            "print('no report')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=_env(lib), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, (r.stdout + r.stderr)[-3000:]
