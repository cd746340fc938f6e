"""Builds librefeff_fms.so (host C++ + gfx950 HIP kernels) in-tree with hipcc.

Also builds librefeff_fms_test.so: the same sources with REFEFF_FMS_TEST_HOOKS,
which adds the kernel-level test entry points (single-matrix LU and GEMM)
that the product library does not export."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librefeff_fms.so")
TEST_LIB = os.path.join(HERE, "librefeff_fms_test.so")
SOURCES = ["fms_host.cpp", "fms_device.hip", "fms_debye.hip", "fms_files.cpp", "fms_io.cpp", "fms_ff2x.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of the generated code on purpose
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = []
    for out, extra in ((LIB, []), (TEST_LIB, ["-DREFEFF_FMS_TEST_HOOKS"])):
        cmd = [hipcc(), *HIPCC_FLAGS, *extra, "-o", out + ".tmp", *srcs]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((out, subprocess.Popen(cmd)))
    for out, proc in jobs:
        if proc.wait() != 0:
            raise subprocess.CalledProcessError(proc.returncode, f"hipcc -> {out}")
    for out, _ in jobs:
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
