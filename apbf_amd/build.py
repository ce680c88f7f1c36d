"""Build recipe of the HIP library (libapbf_hip.so, gfx950).

Every HIP source is compiled for gfx950 with floating-point contraction off,
IEEE division/square root and denormals kept, so device float results match
the host bit for bit where the algorithm is the same.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libapbf_hip.so")
SOURCES = ["primitives.hip", "plain_kernels.hip", "rows.hip", "walk.hip", "adaptive.hip", "scene.hip",
           "capi_lists.hip", "mlp.hip", "graph.hip", "bsearch.hip", "particle_info.hip",
           "snapshot.hip", "noalloc.hip", "lists.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
    "-Wno-unused-value", "-Wno-unused-result",
]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


HOST_CHECK_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "host_api_check.cpp")
HOST_CHECK = os.path.join(os.path.dirname(HERE), "tests", "cpp", "host_api_check")


def build_host_check() -> str:
    """The C++ driver of include/apbf.hpp used by the tests (links the library)."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(os.path.dirname(HERE), "include"), "-o", HOST_CHECK,
           HOST_CHECK_SRC, "-L", HERE, "-lapbf_hip", "-Wl,-rpath,$ORIGIN/../../apbf_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for the host interface check")
    return HOST_CHECK


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_host_check())
