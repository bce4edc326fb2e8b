"""Build the gfx950 library and executable in-tree.

    python -m vlasolve_amd.build          # -> vlasolve_amd/lib/libvlasolve_hip.so
                                          #    vlasolve_amd/bin/vlasolve_hip

Device code is compiled by hipcc for gfx950 only; host C++ by ROCm's clang++.
Floating-point contraction is off everywhere: every kernel keeps the
reference's rounding (no FMA), which the parity tests check bit for bit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libvlasolve_hip.so"
EXE = PKG / "bin" / "vlasolve_hip"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
CLANGXX = str(ROCM / "llvm" / "bin" / "clang++")
ARCH = os.environ.get("VLS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["vls_kernels.hip", "vls_capi.hip", "vls_adjoint.hip"]
HOST_SOURCES = ["host/params.cpp", "host/grid.cpp", "host/solver.cpp", "host/driver.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                      "-packed-fp32-ops", "-munsafe-fp-atomics",
                      # the register line solvers are fully unrolled loops of up to ~200
                      # steps whose bodies exceed the default limit of a forced unroll
                      "-mllvm", "-pragma-unroll-threshold=1000000"]
HOST_FLAGS = COMMON + ["-fopenmp"]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return res


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    EXE.parent.mkdir(exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = BUILD / (Path(s).stem + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", str(CSRC / s), "-o", str(o)])
    for s in HOST_SOURCES:
        o = BUILD / (Path(s).stem + ".o")
        jobs.append([CLANGXX] + HOST_FLAGS + ["-c", str(CSRC / s), "-o", str(o)])
    main_o = BUILD / "main.o"
    jobs.append([CLANGXX] + HOST_FLAGS + ["-c", str(CSRC / "host/main.cpp"), "-o", str(main_o)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr)
    objs = [str(BUILD / (Path(s).stem + ".o")) for s in HIP_SOURCES + HOST_SOURCES]
    tmp = str(LIB) + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    tmp = str(EXE) + ".tmp"
    _run([HIPCC, "-fopenmp", "-o", tmp, str(main_o), f"-L{LIB.parent}", "-lvlasolve_hip",
          "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, EXE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
