"""Compile libhtrip_gpu.so (gfx950) in-tree with hipcc.

Usage: python -m htrip_amd.build   (or htrip_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhtrip_gpu.so")
SOURCES = ["kernels_build.hip", "kernels_query.hip", "runtime.hip", "server.hip"]
ARCH = os.environ.get("HTRIP_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are switched off for all device code.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-pthread",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stderr)
        raise RuntimeError("hipcc failed building libhtrip_gpu.so")
    os.replace(tmp, LIB)
    return LIB


def build_dropin_test(out: str) -> str:
    """The C++20 drop-in test program (tests/cpp/test_dropin.cpp) against the built library."""
    repo = os.path.dirname(HERE)
    src = os.path.join(repo, "tests", "cpp", "test_dropin.cpp")
    cmd = [
        "g++", "-std=c++20", "-O2", "-pthread", "-o", out, src,
        "-I", os.path.join(HERE, "include"), "-I", os.path.join(repo, "include"),
        "-L", HERE, "-lhtrip_gpu", f"-Wl,-rpath,{HERE}",
    ]
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
