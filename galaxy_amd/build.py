"""Builds libgalhip.so in-tree for gfx950 with hipcc.

Flags:
  -ffp-contract=off       every product and sum rounded separately, as the
                          reference's FP32 arithmetic (cell indices and CIC
                          weights must be bit-identical)
  -fhip-fp32-correctly-rounded-divide-sqrt
                          IEEE FP32 division and square root
  -target-feature -packed-fp32-ops
                          no packed FP32 instructions
  -amdgpu-atomic-optimizer-strategy=None
                          atomics to one shared address stay single
                          instructions: the kernels reduce within the
                          wavefront themselves (DPP, ballots) and issue
                          block-level atomics from one lane, where the
                          compiler's default rewrite added a loop over the
                          active lanes to every such atomic
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgalhip.so")
SOURCES = ["galhip_api.hip", "galhip_push.hip", "galhip_solve.hip", "galhip_p3d.hip",
           "galhip_p3d_api.hip", "galhip_adj.hip", "galhip_diag.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]


DRIVER = os.path.join(HERE, "galhip_galaxy")


def build(verbose=False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


def build_driver(verbose=False) -> str:
    """The native host driver (csrc/galhip_galaxy.cpp): plain C++ against the
    C ABI, linked to the in-tree libgalhip.so."""
    tmp = DRIVER + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "galhip_galaxy.cpp"), "-L", HERE, "-lgalhip",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
