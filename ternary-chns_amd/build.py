"""Build libchns_amd.so in-tree for gfx950.

    python ternary-chns_amd/build.py

Compiles the host tables with g++ (-ffp-contract=off: the ETD2RK coefficients
must round exactly like the reference), the kernels with hipcc for gfx950 and
links one shared library into ``chns_amd/libchns_amd.so``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "chns_amd", "libchns_amd.so")
REF_OUT = os.path.join(HERE, "chns_amd", "libchns_ref.so")
DRIVER = os.path.join(HERE, "chns_amd", "chns2d")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are disabled for every HIP translation unit
NO_PK = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = {
        "tables": os.path.join(BUILD, "chns_tables.o"),
        "api": os.path.join(BUILD, "chns_api.o"),
        "diag": os.path.join(BUILD, "chns_api_diag.o"),
        "device": os.path.join(BUILD, "chns_device.o"),
        "batch": os.path.join(BUILD, "chns_batch.o"),
        "ensemble": os.path.join(BUILD, "chns_ensemble.o"),
        "adjapi": os.path.join(BUILD, "chns_adjoint_api.o"),
        "adjdev": os.path.join(BUILD, "chns_adjoint.o"),
        "filter": os.path.join(BUILD, "chns_filter.o"),
        "noalloc": os.path.join(BUILD, "chns_noalloc.o"),
        "sched": os.path.join(BUILD, "chns_batch_sched.o"),
    }
    steps = [
        ["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I" + INC,
         "-c", os.path.join(CSRC, "chns_tables.cpp"), "-o", objs["tables"]],
        [HIPCC, "-O2", "-fPIC", "-std=c++17", "-I" + INC, "-I/opt/rocm/include", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
         "-c", os.path.join(CSRC, "chns_api.cpp"), "-o", objs["api"]],
        [HIPCC, "-O2", "-fPIC", "-std=c++17", "-I" + INC, "-I/opt/rocm/include", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
         "-c", os.path.join(CSRC, "chns_api_diag.cpp"), "-o", objs["diag"]],
        [HIPCC, "-O2", "-fPIC", "-std=c++17", "-I" + INC, "-I/opt/rocm/include", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
         "-c", os.path.join(CSRC, "chns_ensemble.cpp"), "-o", objs["ensemble"]],
        [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC, *NO_PK,
         "-c", os.path.join(CSRC, "chns_device.hip"), "-o", objs["device"]],
        [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC, *NO_PK,
         "-c", os.path.join(CSRC, "chns_batch.hip"), "-o", objs["batch"]],
        [HIPCC, "-O2", "-fPIC", "-std=c++17", "-I" + INC, "-I/opt/rocm/include", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
         "-c", os.path.join(CSRC, "chns_adjoint.cpp"), "-o", objs["adjapi"]],
        [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC, *NO_PK,
         "-c", os.path.join(CSRC, "chns_adjoint.hip"), "-o", objs["adjdev"]],
        [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC, *NO_PK,
         "-c", os.path.join(CSRC, "chns_filter.hip"), "-o", objs["filter"]],
        [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC, *NO_PK,
         "-c", os.path.join(CSRC, "chns_noalloc.hip"), "-o", objs["noalloc"]],
        [HIPCC, "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC, *NO_PK,
         "-c", os.path.join(CSRC, "chns_batch_sched.hip"), "-o", objs["sched"]],
    ]
    with ThreadPoolExecutor(8) as ex:
        list(ex.map(_run, steps))
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", OUT, objs["device"], objs["batch"],
          objs["adjdev"], objs["filter"], objs["noalloc"], objs["sched"], objs["api"], objs["diag"], objs["ensemble"], objs["adjapi"], objs["tables"]])
    # the reference-named entry points (chns_ref.h) over the C ABI, and the
    # time-loop driver on them (plain C), found next to the library via rpath
    _run(["gcc", "-O2", "-std=c11", "-D_GNU_SOURCE", "-Wall", "-Wextra", "-fPIC", "-shared", "-I" + INC,
          os.path.join(HERE, "host", "chns_ref.c"), "-o", REF_OUT, "-L" + os.path.dirname(OUT), "-lchns_amd",
          "-Wl,-rpath,$ORIGIN", "-lm"])
    _run(["gcc", "-O2", "-std=c11", "-D_GNU_SOURCE", "-Wall", "-I" + INC, os.path.join(HERE, "host", "chns2d.c"),
          "-o", DRIVER, "-L" + os.path.dirname(OUT), "-lchns_ref", "-lchns_amd", "-Wl,-rpath,$ORIGIN", "-lm"])
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
