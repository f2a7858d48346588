"""Build libufvm.so (HIP, gfx950) in-tree with hipcc.

The library is one shared object: the gfx950 kernels, the host orchestration
and the extern "C" ABI of include/ufvm.h.  Built in place next to this file so
that it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libufvm.so")

SOURCES = ["api.hip", "kernels.hip", "kernels_batch.hip", "simple_api.hip", "simple_kernels.hip",
           "pcg.hip", "server.hip", "noalloc.hip", "hostalg.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # every multiply and add rounds separately, as the reference arithmetic does
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


# kernels of known traffic that calibrate the profiler's memory counters
# (measurement tooling loaded by bench.py only; not part of the solver)
CALIB_LIB = os.path.join(HERE, "libufvm_calib.so")


def _compile(sources, out, verbose):
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *sources, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    _compile([os.path.join(CSRC, s) for s in SOURCES], LIB, verbose)
    _compile([os.path.join(CSRC, "calibrate.hip")], CALIB_LIB, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
