"""Builds libcudaem_amd.so (gfx950) in-tree with hipcc.

Usage: python -m cudaem_amd.build [-v]
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcudaem_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["em_kernels.hip", "em_host.cpp", "em_capi.cpp", "em_simulators.cpp"]

# Packed FP32 instructions stay off (-packed-fp32-ops); gfx950 only.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]


# The passes prune line slots that hold known zeros (band-limit skip): the
# kernels' file may drop "x + 0", which differs from x only in the sign of a
# zero.
EXTRA = {"em_kernels.hip": ["-fno-signed-zeros"]}


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    cmd = [HIPCC] + FLAGS + EXTRA.get(src, []) + ["-x", "hip", "-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_cpp_test(verbose)
    return LIB


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_simulators.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_simulators")


def build_cpp_test(verbose=False):
    """The C++ drop-in test program: plain host C++ against the simulator
    surface, linked to the library (rpath relative to the binary)."""
    if not os.path.exists(CPP_TEST_SRC):
        return None
    tmp = CPP_TEST_BIN + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), CPP_TEST_SRC, "-o", tmp,
           "-L", HERE, "-lcudaem_amd", "-Wl,-rpath,$ORIGIN/../../cudaem_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ test program failed to build:\n{r.stderr}")
    os.replace(tmp, CPP_TEST_BIN)
    if verbose:
        print("built", CPP_TEST_BIN)
    return CPP_TEST_BIN


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
