"""Build libfnumber_hip.so (gfx950) in-tree with hipcc.

Usage: ``python -m f_number_amd.build`` (or ``f_number_amd.build.build()``).
Objects go to f_number_amd/_build/, the library to f_number_amd/libfnumber_hip.so.
``python -m f_number_amd.build --host-sanitize`` builds the host-sanitizer variant
into f_number_amd/_build_asan/ (CPU checks of the host code only).
Every HIP translation unit is compiled with packed-FP32 instruction selection
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfnumber_hip.so")
ARCH = os.environ.get("FN_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall"]

# host_sanitize: AddressSanitizer and UndefinedBehaviorSanitizer on the host code, for
# CPU checks of the C ABI (argument checks, plans, host generators).  hipcc takes them
# after -Xarch_host, so the gfx950 code is compiled as usual; the plain C++ unit goes
# through clang++.  The beamforming instance units (kernels and their launch wrappers)
# are taken from the default build's objects as they are.  The library leaves the
# sanitizer runtime unresolved: a process loads it with clang's ASan runtime preloaded.
SAN_DIR = os.path.join(HERE, "_build_asan")
SAN_LIB = os.path.join(SAN_DIR, "libfnumber_hip_asan.so")
CLANGXX = "/opt/rocm/llvm/bin/clang++"
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize=vptr", "-fno-sanitize-recover=undefined",
             "-fno-omit-frame-pointer", "-g"]


def asan_runtime() -> str:
    """clang's dynamic ASan runtime (it holds the UBSan handlers too)."""
    r = subprocess.run([CLANGXX, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    return r.stdout.strip()


def _units(extra, host_sanitize=False):
    host_cc = [CLANGXX] + CXX_FLAGS + SAN_FLAGS if host_sanitize else [HIPCC, "-x", "c++"] + CXX_FLAGS
    if host_sanitize:
        extra = extra + [a for f in SAN_FLAGS for a in ("-Xarch_host", f)]
    units = [
        ("fn_host.o", host_cc + ["-c", os.path.join(CSRC, "fn_host.cpp")]),
        ("fn_core.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_core.hip")]),
        ("fn_synth.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_synth.hip")]),
        ("fn_adjoint.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_adjoint.hip")]),
        ("fn_windowed.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_windowed.hip")]),
        ("fn_signal.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_signal.hip")]),
    ]
    if host_sanitize:
        return units
    for mode in (0, 1):
        for ki in (8, 16, 32, 48, 64):
            for prec in (0, 1, 2, 3):
                units.append((f"fn_beamform_m{mode}_k{ki}_p{prec}.o",
                              [HIPCC, "-x", "hip"] + HIP_FLAGS + extra +
                              [f"-DINST_MODE={mode}", f"-DINST_KI={ki}", f"-DINST_PREC={prec}", "-c",
                               os.path.join(CSRC, "fn_beamform_inst.hip")]))
    return units


def _run(obj_dir, name, cmd):
    out = os.path.join(obj_dir, name)
    r = subprocess.run(cmd + ["-o", out], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile of {name} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(jobs: int | None = None, verbose: bool = False, extra_flags=(), lib: str = LIB,
          obj_dir: str = OBJ, host_sanitize: bool = False) -> str:
    """Compile every translation unit for gfx950 and link the shared library.

    extra_flags (e.g. -D tuning macros) and lib/obj_dir allow building variants
    next to the default library; host_sanitize builds the sanitizer variant (into
    _build_asan/ unless lib/obj_dir say otherwise)."""
    if host_sanitize and lib == LIB:
        lib, obj_dir = SAN_LIB, SAN_DIR
    os.makedirs(obj_dir, exist_ok=True)
    units = _units(list(extra_flags), host_sanitize)
    jobs = jobs or min(len(units), os.cpu_count() or 4, 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _run(obj_dir, *u), units))
    if host_sanitize:
        inst = [os.path.join(OBJ, n) for n, _ in _units([]) if n.startswith("fn_beamform_")]
        missing = [o for o in inst if not os.path.exists(o)]
        if missing:
            raise RuntimeError(f"host_sanitize reuses the default build's objects: build it first ({missing[0]})")
        objs += inst
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(verbose=True, host_sanitize="--host-sanitize" in sys.argv[1:])
    sys.exit(0)
