"""Build libpsy_gpu.so (HIP, gfx950) in-tree with hipcc.

Run as ``python -m psychopath_cpp_amd.build`` or through
``__graft_entry__.build()``.  The library is always rebuilt.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsy_gpu.so")

SOURCES = ["psy_kernels.hip", "psy_capi.cpp", "psy_render.cpp", "psy_parser.cpp", "psy_film.cpp", "psy_subdiv.cpp"]

# -ffp-contract=off and IEEE div/sqrt: results must match the reference's
# uncontracted SSE float arithmetic bit for bit.  Packed FP32 instructions
# are disabled for correctness on this hardware.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose=True):
    cmd = [hipcc()] + HIPCC_FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
