"""Edge cases of the whole step against the oracle on the 64^3 3-level pool
grid: no particles (every block coarsest, nothing to solve), one particle, a
5000-particle clump inside one block, and the pool with particles on and
beyond the tank's faces (x = 1, y < 0, z = 1.2, x one ulp below 1), which bin
into the nearest boundary block as on the CPU.  Both precisions.  Bars as for
the full-step parity: levels and per-block counts bit-exact, positions within
1e-5, velocities within 1e-3 of max |v|.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
from helpers import DT, GRAV, GpuParticles, gpu_grid, oracle_grid, particles, step_checked

pytestmark = pytest.mark.gpu


def _case(name):
    rng = np.random.default_rng(9)
    if name == "empty":
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)
    if name == "one":
        return np.array([[0.51, 0.52, 0.53]], np.float32), np.array([[0.1, -0.2, 0.3]], np.float32)
    if name == "clump":
        p = (np.array([0.4, 0.4, 0.4]) + rng.random((5000, 3)) * (1 / 8)).astype(np.float32)
        return p, (0.1 * rng.standard_normal((5000, 3))).astype(np.float32)
    pos, vel = particles("pool64")
    pos = pos.copy()
    sel = rng.choice(len(pos), 2000, replace=False)
    pos[sel[:500], 0] = 1.0
    pos[sel[500:1000], 1] = -0.01
    pos[sel[1000:1500], 2] = 1.2
    pos[sel[1500:], 0] = np.nextafter(np.float32(1), np.float32(0))
    return pos, vel


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["empty", "one", "clump", "outside"])
def test_step_edge_cases(ctx, name, precision):
    pos, vel = _case(name)
    print(f"[edges] {name}: {len(pos)} particles, precision {precision}", flush=True)
    og = oracle_grid("pool64")
    ref = og.step(pos, vel, DT, GRAV, 0.95, tol=1e-6)
    assert ref["status"] == 0
    gg = gpu_grid(ctx, "pool64")
    sv = aw.VCycleSolver(gg, precision=precision)
    gp = GpuParticles(ctx, pos, vel)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6)
    st = step_checked(step, gp, precision)
    print(f"[edges] {name}: iters gpu {st.solve.iterations} oracle {ref['iters']}", flush=True)
    np.testing.assert_array_equal(gg.levels(), og.levels())
    _, cnt = og.bin(pos)
    np.testing.assert_array_equal(gg.particle_counts(), cnt)
    if name == "empty":
        assert np.all(gg.levels() == 2)
        assert st.solve.iterations == 0
        assert not np.any(gg.cell("p"))
        return
    P, V = gp.unpermuted()
    assert np.abs(P - ref["pos"]).max() < 1e-5
    vmax = max(np.abs(ref["vel"]).max(), 1e-6)
    assert np.abs(V - ref["vel"]).max() / vmax < 1e-3
