"""Build the HIP library in-tree for gfx950 (hipcc, no JIT cache)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "edlut_amd.hip")
OUT = os.path.join(HERE, "libedlut_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off, IEEE division/sqrt and preserved denormals keep the float
# results identical to the reference's CPU models; packed FP32 ops are off.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


DRIVER_SRC = os.path.join(HERE, "host", "plugin_driver.cpp")
DRIVER = os.path.join(HERE, "host", "plugin_driver")


def build_driver(verbose: bool = False) -> str:
    """The C++ plug-in driver (host code only, g++), linked to the HIP library."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-o", DRIVER, DRIVER_SRC, "-L", HERE, "-ledlut_amd",
           "-Wl,-rpath,$ORIGIN/..", "-I", os.path.join(ROOT, "include")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for " + DRIVER_SRC)
    if verbose:
        print("built", DRIVER)
    return DRIVER


def build(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-o", OUT + ".tmp", SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for " + SRC)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    build_driver(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
