"""Build the HIP shared library ``laml_amd/liblaml_pmm.so`` for gfx950.

Run ``python -m laml_amd.build`` (or ``__graft_entry__.build()``).  hipcc
cross-compiles on a machine without a GPU; the library is built in-tree so it
travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblaml_pmm.so")

SOURCES = ["pmm_kernels.hip", "pmm_capi.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 instructions are kept out of every HIP object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose=True):
    """Compile every HIP source into the in-tree shared library."""
    cmd = [hipcc_path()] + HIPCC_FLAGS + [
        "-I", os.path.join(ROOT, "include"),
        "-I", CSRC,
    ] + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
