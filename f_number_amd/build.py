"""Build libfnumber_hip.so (gfx950) in-tree with hipcc.

Usage: ``python -m f_number_amd.build`` (or ``f_number_amd.build.build()``).
Objects go to f_number_amd/_build/, the library to f_number_amd/libfnumber_hip.so.
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


def _units(extra):
    units = [
        ("fn_host.o", [HIPCC, "-x", "c++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, "fn_host.cpp")]),
        ("fn_core.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_core.hip")]),
        ("fn_synth.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_synth.hip")]),
        ("fn_adjoint.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_adjoint.hip")]),
        ("fn_windowed.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_windowed.hip")]),
        ("fn_signal.o", [HIPCC, "-x", "hip"] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "fn_signal.hip")]),
    ]
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
          obj_dir: str = OBJ) -> str:
    """Compile every translation unit for gfx950 and link the shared library.

    extra_flags (e.g. -D tuning macros) and lib/obj_dir allow building variants
    next to the default library."""
    os.makedirs(obj_dir, exist_ok=True)
    units = _units(list(extra_flags))
    jobs = jobs or min(len(units), os.cpu_count() or 4, 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _run(obj_dir, *u), units))
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
    build(verbose=True)
    sys.exit(0)
