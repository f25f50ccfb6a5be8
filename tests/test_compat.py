"""The C++ drop-in surface (include/msbg_compat.hpp) builds with the host
compiler alone, and the example caller written against it
(anitowave_adaptive_amd/examples/msbg_main.cpp) links to the C ABI library;
without arguments it prints its usage, and without a GPU it fails cleanly
at the context (no compute call on the CPU)."""
import os
import subprocess

import numpy as np
import pytest

import anitowave_adaptive_amd as aw
from anitowave_adaptive_amd import build as B


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "msbg_compat.hpp"\nint main() { msbg::Bands b; return b.num_levels == 3 ? 0 : 1; }\n')
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", aw.INCLUDE_DIR, str(src), "-o", str(exe),
                    "-L", os.path.dirname(aw.LIB_PATH), "-lanitowave"], check=True)


def test_example_links_and_prints_usage():
    assert os.path.exists(B.EXAMPLE), "build() builds the example caller"
    r = subprocess.run([B.EXAMPLE], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_example_fails_cleanly_without_a_device(tmp_path):
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([1], np.int64).tofile(f)
        np.full(6, 0.5, np.float32).tofile(f)
    r = subprocess.run([B.EXAMPLE, str(inp), str(tmp_path / "out.bin"), "1", "1", "1", "1", "0", "0", "1", "1e-6"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "aw_ctx_create" in r.stderr
