"""Build recipe of the HIP library: hipcc for gfx950, in-tree output.

    python -m anitowave_adaptive_amd.build        # -> anitowave_adaptive_amd/libanitowave.so
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libanitowave.so")
ARCH = os.environ.get("AW_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["aw_kernels.hip", "aw_runtime.cpp"]

# -packed-fp32-ops: keep the compiler from emitting packed FP32 arithmetic.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src):
    obj = os.path.join(CSRC, "_build", os.path.splitext(src)[0] + ".o")
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


EXAMPLE_SRC = os.path.join(HERE, "examples", "msbg_main.cpp")
EXAMPLE = os.path.join(HERE, "examples", "msbg_main")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_example(verbose=False):
    """The C++ caller written against include/msbg_compat.hpp, compiled by the
    host compiler alone (no HIP headers) and linked to libanitowave.so."""
    cxx = os.environ.get("CXX", "g++")
    tmp = EXAMPLE + ".tmp"
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", INCLUDE, EXAMPLE_SRC, "-o", tmp, "-L", HERE,
           "-lanitowave", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"example build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, EXAMPLE)
    if verbose:
        print("built", EXAMPLE)
    return EXAMPLE


def build(verbose=False):
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_example(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
