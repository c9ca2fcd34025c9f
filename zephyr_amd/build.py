"""Build the zfv HIP library (libzfv.so) for gfx950, in-tree.

The library is compiled with -ffp-contract=off: no multiply-add fusion, so
the device arithmetic is the same sequence of correctly rounded IEEE
operations as the reference's x86-64 build (CMakeLists.txt:16-19, no -march).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libzfv.so")

SOURCES = [
    os.path.join(HERE, "csrc", "zfv_kernels.hip"),
    os.path.join(HERE, "csrc", "zfv_amr.hip"),
    os.path.join(HERE, "csrc", "zfv_capi.cpp"),
]
HEADERS = [
    os.path.join(HERE, "csrc", "zfv_physics.hpp"),
    os.path.join(HERE, "csrc", "zfv_internal.hpp"),
    os.path.join(ROOT, "include", "zfv.h"),
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ZFV_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def build(verbose=False):
    """Compile libzfv.so; always recompiles (the build check)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libzfv.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
