"""The smoother's launches leave the all-air blocks out (the smoother writes
nothing on air cells): per multigrid level, aw_level_info.smoothed_blocks is
the number of blocks holding a liquid cell, counted here from the level's own
liquid mask and cell offsets (levels that keep the full lists report every
block).  After further steps, with the liquid mask changed under the same or
a new topology, the counts follow the new masks, and the step still matches
the oracle's FP64 step on the same topology (positions 1e-5 absolute,
velocities 1e-3 of max |v|).  The
smoother's values against the oracle's on every level are checked in
test_gpu_parity.py (test_residual_and_smoother_parity).
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
from helpers import DT, GRAV, GpuParticles, gpu_grid, oracle_grid, particles, step_checked

pytestmark = pytest.mark.gpu


def _liquid_blocks(ctx, sv, k):
    L = sv.level_info(k)
    nb = int(L.num_blocks)
    coff = aw.download(ctx, L.d_cell_off, np.uint32, nb + 1).astype(np.int64)
    liq = sv.level_get(k, "liquid").astype(np.int64)
    per_block = np.add.reduceat(liq, coff[:-1]) if nb else np.zeros(0, np.int64)
    return nb, int(L.smoothed_blocks), int((per_block > 0).sum())


def _check_levels(ctx, sv, tag):
    out = []
    for k in range(sv.num_levels()):
        nb, sm, live = _liquid_blocks(ctx, sv, k)
        print(f"[lists] {tag} level {k}: {nb} blocks, {live} with liquid, smoothed {sm}", flush=True)
        assert sm == live or sm == nb, (k, nb, sm, live)
        out.append((nb, sm, live))
    # level 0 always runs the work-list smoother
    nb0, sm0, live0 = out[0]
    assert sm0 == live0
    return out


@pytest.mark.parametrize("name", ["dam32", "pool64", "double64"])
def test_smoothed_blocks_are_the_liquid_blocks(ctx, name):
    pos, vel = particles(name)
    gg = gpu_grid(ctx, name)
    sv = aw.VCycleSolver(gg)
    gp = GpuParticles(ctx, pos, vel)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6)
    step_checked(step, gp)
    lv = _check_levels(ctx, sv, f"{name} step 1")
    # the scenes leave air above the liquid: some blocks are left out
    assert lv[0][1] < lv[0][0]


def test_lists_follow_the_mask_over_steps(ctx):
    name = "pool64"
    pos, vel = particles(name)
    gg = gpu_grid(ctx, name)
    sv = aw.VCycleSolver(gg)
    gp = GpuParticles(ctx, pos, vel)
    # topology kept between steps: the lists change with the liquid mask alone
    keep = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6, rebuild_topology=False)
    rebuild = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6)
    step_checked(rebuild, gp)
    for i, stepper in enumerate([keep, keep, rebuild, keep]):
        p1, v1, ids = gp.get()
        lv_before = gg.levels().copy()
        st = step_checked(stepper, gp)
        _check_levels(ctx, sv, f"step {i + 2}")
        og = oracle_grid(name)
        ref = og.step64(p1, v1, DT, GRAV, 0.95, tol=1e-10, levels=None if stepper is rebuild else lv_before)
        assert ref["status"] == 0
        P, V = gp.unpermuted()
        Pr = np.empty_like(ref["pos"])
        Vr = np.empty_like(ref["vel"])
        Pr[ids] = ref["pos"]
        Vr[ids] = ref["vel"]
        assert np.abs(P - Pr).max() < 1e-5
        assert np.abs(V - Vr).max() / max(np.abs(Vr).max(), 1e-6) < 1e-3
