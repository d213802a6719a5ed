"""Build recipe for libppf_amd.so (gfx950) -- run by __graft_entry__.build().

hipcc compiles every HIP source of csrc/ into one shared library next to this
file.  Numerics flags: no FMA contraction, no fast-math, packed FP32 off; HIP's
default correctly rounded fp32 division/sqrt and IEEE denormals are kept.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libppf_amd.so")
SOURCES = ["api.hip", "batch.hip", "server.hip", "preprocess.hip", "host_io.hip", "kernels.hip", "radix_sort.hip", "tune.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", CSRC, "-o", tmp, *srcs, "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx",
           "-Wl,-rpath,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed building libppf_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
