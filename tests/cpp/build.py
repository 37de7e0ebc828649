"""TEST INFRASTRUCTURE: builds tests/cpp/test_cpp_api with g++ against
include/psychopath_gpu.hpp and libpsy_gpu.so (the binary stays in-tree,
git-ignored, so GPU runs use it as built here)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "test_cpp_api.cpp")
BIN = os.path.join(HERE, "test_cpp_api")


def build():
    libdir = os.path.join(ROOT, "psychopath_cpp_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), SRC, "-o",
           BIN + ".tmp", "-L" + libdir, "-lpsy_gpu", "-Wl,-rpath," + libdir, "-L/opt/rocm/lib",
           "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ build of test_cpp_api failed:\n" + r.stderr[-4000:])
    os.replace(BIN + ".tmp", BIN)
    return BIN


if __name__ == "__main__":
    print(build())
