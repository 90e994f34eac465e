"""The host code of the HIP library under AddressSanitizer and UndefinedBehaviorSanitizer.

``f_number_amd.build.build(host_sanitize=True)`` compiles the host side of every
translation unit with ``-fsanitize=address,undefined`` (hipcc, after ``-Xarch_host``,
``--cuda-host-only``: no gfx950 code) into a temporary directory.  The CPU tests of the C
ABI (argument and band checks, status identifiers, the host workload generator, the
line-scan parameters, the wide-array and scratch-size refusals) then run in a child
process that preloads clang's ASan runtime and loads that library through
``FN_LIB_PATH``; an invalid access or undefined operation in the host code aborts it.
Leak detection is off: the interpreter and the HIP runtime keep allocations to the end.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from f_number_amd import build as fb  # noqa: E402

HOST_TESTS = ["tests/test_host_abi.py", "tests/test_linescan_host.py"]


def test_host_abi_tests_pass_under_sanitizers(tmp_path):
    rt = fb.asan_runtime()
    if not (os.path.isabs(rt) and os.path.exists(rt)):
        pytest.skip("clang's ASan runtime is not installed")
    inst = [os.path.join(fb.OBJ, n) for n, _ in fb._units([]) if n.startswith("fn_beamform_")]
    if not all(os.path.exists(o) for o in inst):
        pytest.skip("the library's default build (whose kernel-instance objects this variant reuses) is missing")
    lib = fb.build(lib=str(tmp_path / "libfnumber_hip_asan.so"), obj_dir=str(tmp_path / "obj"),
                   host_sanitize=True)
    env = dict(os.environ)
    env.update(FN_LIB_PATH=lib, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    probe = ("import f_number_amd as fa\nfa.load()\n"
             "print([l.split()[-1] for l in open('/proc/self/maps') if 'libfnumber' in l][0])\n")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("libfnumber_hip_asan.so"), r.stdout + r.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        *HOST_TESTS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
