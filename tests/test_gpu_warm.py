"""Warm start of the pressure solve (aw_step_params.warm_start): each step's
PCG starts from the previous step's pressure, remapped cell by cell onto the
new layout when the topology changed.  No reference counterpart (the
reference starts from p = 0, main.cu:300).

Both runs solve to the same relative residual, so a rollout with warm starts
must track the zero-start rollout to within the solver tolerance: positions
within 2e-5 absolute and velocities within 1e-3 of max |v| after 6 steps of
the 64^3 pool with a sphere falling at 8 m/s, whose block levels change at
steps 2 and 3 (the oracle's topology of the same rollout, run on the CPU,
changes there), so two warm starts go through the remap onto a new layout.
The warm-started solves must take fewer PCG iterations in total; the same
holds for mixed precision to 1e-8, whose FP64 iterative refinement starts from
the remapped pressure.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
from anitowave_adaptive_amd import scenes
from helpers import DT, GRAV, GpuParticles, gpu_grid, step_checked

pytestmark = pytest.mark.gpu

STEPS = 6


def _run(ctx, warm, precision=0, tol=1e-6):
    pos, vel = scenes.pool_sphere(n=64, seed=1234, vy=-8.0)
    gg = gpu_grid(ctx, "pool64")
    sv = aw.VCycleSolver(gg, precision=precision)
    gp = GpuParticles(ctx, pos, vel)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=tol, warm_start=warm, max_iters=400)
    iters, changed = [], []
    for _ in range(STEPS):
        st = step_checked(step, gp, precision)
        iters.append(st.solve.iterations)
        changed.append(st.topology_changed)
    P, V = gp.unpermuted()
    return P, V, iters, changed


def test_warm_start_tracks_zero_start(ctx):
    P0, V0, it0, ch0 = _run(ctx, False)
    P1, V1, it1, ch1 = _run(ctx, True)
    print(f"[warm] PCG iterations zero start {it0}, warm start {it1}; topology changes {ch0}", flush=True)
    assert ch0 == ch1 and sum(ch0[1:]) >= 2, (ch0, ch1)
    assert it1[0] == it0[0], "the first step has no previous pressure"
    assert sum(it1[1:]) < sum(it0[1:]), (it0, it1)
    assert np.abs(P1 - P0).max() < 2e-5
    vmax = max(np.abs(V0).max(), 1e-6)
    assert np.abs(V1 - V0).max() / vmax < 1e-3


def test_warm_start_mixed_precision(ctx):
    """The FP64 iterative refinement starts from the remapped pressure too:
    to 1e-8, fewer inner iterations in total and the same particles."""
    P0, V0, it0, ch0 = _run(ctx, False, precision=1, tol=1e-8)
    P1, V1, it1, ch1 = _run(ctx, True, precision=1, tol=1e-8)
    print(f"[warm] mixed 1e-8 inner iterations zero start {it0}, warm start {it1}", flush=True)
    assert ch0 == ch1, (ch0, ch1)
    assert sum(it1[1:]) < sum(it0[1:]), (it0, it1)
    assert np.abs(P1 - P0).max() < 2e-5
    vmax = max(np.abs(V0).max(), 1e-6)
    assert np.abs(V1 - V0).max() / vmax < 1e-3


def _graded(rng, D, L):
    """Random native levels with 2:1 grading over the 26-neighbourhood."""
    from scipy.ndimage import minimum_filter
    lv = rng.integers(0, L, (D[2], D[1], D[0]))
    while True:
        new = np.minimum(lv, minimum_filter(lv, size=3, mode="nearest") + 1)
        if np.array_equal(new, lv):
            return lv.reshape(-1).astype(np.uint8)
        lv = new


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_remap_cells_bit_exact(ctx, seed):
    """aw_remap_cells (the warm start's pressure_remap_kernel) against the
    oracle's remap by containment, bit-exact, between random graded layouts
    with all four native levels; an unchanged layout copies the field."""
    import oracle as O
    D = (6, 5, 4)
    rng = np.random.default_rng(seed)
    prev, cur = _graded(rng, D, 4), _graded(rng, D, 4)
    print(f"[warm] remap seed {seed}: levels {np.bincount(prev, minlength=4)} -> {np.bincount(cur, minlength=4)}",
          flush=True)
    gg = aw.MSBGManager(ctx, D, 1.0 / 64, num_levels=4)
    og = O.OracleGrid(D, num_levels=4, h=1.0 / 64)
    gg.set_levels(prev)
    nprev = int(gg.layout().total_cells)
    src = rng.standard_normal(nprev).astype(np.float32)
    dsrc = aw.DeviceArray.from_numpy(ctx, src)
    dlev = aw.DeviceArray.from_numpy(ctx, prev)
    gg.set_levels(cur)
    ncur = int(gg.layout().total_cells)
    dout = aw.DeviceArray(ctx, 4 * ncur)
    gg.remap_cells(dlev.ptr, dsrc.ptr, dout.ptr)
    got = dout.download(np.float32, ncur)
    og.set_levels(cur)
    want = og.remap_cells(prev, src.astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(got, want)
    gg.set_levels(prev)
    dcopy = aw.DeviceArray(ctx, 4 * nprev)
    gg.remap_cells(dlev.ptr, dsrc.ptr, dcopy.ptr)
    np.testing.assert_array_equal(dcopy.download(np.float32, nprev), src)
