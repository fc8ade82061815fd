"""Builds the HIP library ``libinlm.so`` in-tree for gfx950.

``python -m microtomodenoise_amd.build`` compiles every HIP/C++ source of the
package with hipcc into ``microtomodenoise_amd/libinlm.so``.  The kernels are
compiled with ``-ffp-contract=off`` (the parity kernels rely on unfused
multiply/add) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libinlm.so")

SOURCES = ["inlm_kernels.hip", "inlm_mfma.hip", "inlm_learn.hip", "inlm_noalloc.hip", "inlm_capi.cpp",
           "inlm_tables.cpp", "inlm_io.cpp", "inlm_slabs.cpp", "inlm_rgb.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
    "-ffp-contract=off", "-fopenmp",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libinlm.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


FACADE_SRC = os.path.join(ROOT, "tests", "cpp", "facade_main.cpp")
FACADE_BIN = os.path.join(ROOT, "tests", "cpp", "facade_main")


def build_facade_test(verbose: bool = False) -> str:
    """Compiles the C++ driver of the header-only drop-in facade
    (include/iternlm_gpu.h) against libinlm.so, for the GPU tests."""
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), FACADE_SRC, "-L" + HERE, "-linlm",
           "-Wl,-rpath," + HERE, "-o", FACADE_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building the facade driver")
    return FACADE_BIN


CLI_SRC = os.path.join(CSRC, "inlm_main.cpp")
CLI_BIN = os.path.join(HERE, "inlm")


def build_cli(verbose: bool = False) -> str:
    """Compiles the command-line program (the reference's main.cpp flags over
    inlm_denoise_files) against libinlm.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), CLI_SRC, "-L" + HERE, "-linlm",
           "-Wl,-rpath," + HERE, "-o", CLI_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building the command-line program")
    return CLI_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_facade_test(verbose=True))
    print(build_cli(verbose=True))
