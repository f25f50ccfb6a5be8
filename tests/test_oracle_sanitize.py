"""The FP64 oracle under sanitizers and across thread counts.

* `make -C oracle sanitize` builds the oracle with a driver of its entry
  points (topology, P2G with and without densities, PCG and adaptive
  stationary solves, the FP64 step and its adjoint, remap, resample) under
  AddressSanitizer and UndefinedBehaviorSanitizer and runs it with four
  OpenMP threads: no report, the driver's own checks pass.
* ThreadSanitizer is not usable with this toolchain's libgomp (its barriers
  are not instrumented, so every parallel region reports the runtime's own
  stack hand-off); instead one step with 1 and with 8 OpenMP threads agrees
  to round-off (a racy accumulation would not).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from anitowave_adaptive_amd import scenes

ORACLE_DIR = os.path.dirname(os.path.abspath(O.__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with the sanitizer runtimes")
def test_asan_ubsan_driver_runs_clean():
    r = subprocess.run(["make", "-s", "-C", ORACLE_DIR, "sanitize"], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "sanitize_driver:" in r.stdout and "ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_thread_count_independent():
    pos, vel = scenes.pool_sphere(n=32, radius=0.15)
    out = []
    nt = O.num_threads()
    try:
        for threads in (1, 8):
            O.set_threads(threads)
            g = O.OracleGrid((4, 4, 4), num_levels=3, h=1.0 / 32, band=(0, 1, 0, 0))
            g.set_params()
            ref = g.step(pos, vel, 1 / 240, (0, -9.81, 0), 0.95, tol=1e-10, max_iters=400)
            assert ref["status"] == 0
            out.append((ref["pos"], ref["vel"], g.cell("p")))
    finally:
        O.set_threads(nt)
    (p1, v1, q1), (p8, v8, q8) = out
    np.testing.assert_allclose(p8, p1, rtol=0, atol=1e-12)
    np.testing.assert_allclose(v8, v1, rtol=0, atol=1e-10 * np.abs(v1).max())
    np.testing.assert_allclose(q8, q1, rtol=0, atol=1e-10 * np.abs(q1).max())
