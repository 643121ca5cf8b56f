"""Builds the native library of the search path in-tree.

    python -m pca_amd.build        # -> pca_amd/libpcagpu.so + CLIs

hipcc compiles the gfx950 kernels and the C++ host driver into one shared
library with a C ABI (include/pcagpu.h), plus the three drop-in command
lines.  Nothing here needs a GPU: hipcc cross-compiles for gfx950.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpcagpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "value.hip", "noalloc.hip", "pcagpu.cpp", "server.cpp", "complete.cpp", "host_core.cpp", "host_dfs.cpp"]
# the three drop-in command lines, one source built in three modes
CLIS = {"gpu_hybrid": 0, "gpu_enumerate_oep": 1, "gpu_wavefront": 2}

# -packed-fp32-ops: no packed FP32 VALU forms in any kernel of this library.
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", LIB] + objs)
    for exe, mode in CLIS.items():
        _run(["g++", "-O2", "-std=c++17", "-DPCA_CLI_MODE=%d" % mode, "-I" + os.path.join(ROOT, "include"),
              os.path.join(CSRC, "cli.cpp"), "-o", os.path.join(HERE, exe), "-L" + HERE, "-lpcagpu",
              "-Wl,-rpath,$ORIGIN"])
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
