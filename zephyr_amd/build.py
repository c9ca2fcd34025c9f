"""Build the zfv HIP library (libzfv.so) for gfx950, in-tree.

The library is compiled with -ffp-contract=off: no multiply-add fusion, so
the device arithmetic is the same sequence of correctly rounded IEEE
operations as the reference's x86-64 build (CMakeLists.txt:16-19, no -march).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libzfv.so")

SOURCES = [
    os.path.join(HERE, "csrc", "zfv_kernels.hip"),
    os.path.join(HERE, "csrc", "zfv_amr.hip"),
    os.path.join(HERE, "csrc", "zfv_adjoint.hip"),
    os.path.join(HERE, "csrc", "zfv_mm.hip"),
    os.path.join(HERE, "csrc", "zfv_order.hip"),
    os.path.join(HERE, "csrc", "zfv_tile.hip"),
    os.path.join(HERE, "csrc", "zfv_transfer.hip"),
    os.path.join(HERE, "csrc", "zfv_3d.hip"),
    os.path.join(HERE, "csrc", "zfv_3d.cpp"),
    os.path.join(HERE, "csrc", "zfv_capi.cpp"),
    os.path.join(HERE, "csrc", "zfv_io.cpp"),
]
HEADERS = [
    os.path.join(HERE, "csrc", "zfv_physics.hpp"),
    os.path.join(HERE, "csrc", "zfv_internal.hpp"),
    os.path.join(HERE, "csrc", "zfv_adjoint_ops.hpp"),
    os.path.join(HERE, "csrc", "zfv_step_common.hpp"),
    os.path.join(HERE, "csrc", "zfv_step_faces.hpp"),
    os.path.join(HERE, "csrc", "zfv_mm.hpp"),
    os.path.join(HERE, "csrc", "zfv_3d.hpp"),
    os.path.join(ROOT, "include", "zfv.h"),
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ZFV_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
# Sources compile one per process, then link; the translation units share no
# device symbols, so no relocatable device code is needed.
JOBS = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))


EXAMPLES = {
    # problem files written against the host C++ facade (zephyr_amd/host/zfv.hpp)
    "smf_quadrants": os.path.join(HERE, "host", "examples", "smf_quadrants.cpp"),
}
EXAMPLE_PATHS = {k: os.path.join(BUILD_DIR, k) for k in EXAMPLES}


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(what)


def build(verbose=False):
    """Compile libzfv.so and the host examples; always recompiles (the build check)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = [os.path.join(BUILD_DIR, os.path.basename(src) + ".o") for src in SOURCES]

    def compile_one(i):
        _run([HIPCC] + FLAGS + ["-c", SOURCES[i], "-o", objs[i]], "hipcc failed compiling " + SOURCES[i], verbose)

    with ThreadPoolExecutor(JOBS) as pool:
        list(pool.map(compile_one, range(len(SOURCES))))
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, "hipcc failed linking libzfv.so", verbose)
    os.replace(tmp, LIB_PATH)
    for name, src in EXAMPLES.items():
        _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "host"),
              src, "-o", EXAMPLE_PATHS[name], "-L" + BUILD_DIR, "-lzfv", "-Wl,-rpath,$ORIGIN"],
             f"g++ failed building {name}", verbose)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
