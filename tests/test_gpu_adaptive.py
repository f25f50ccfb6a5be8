"""The reference's adaptive smoothing on the HIP path (aw_solver_params
adaptive = 1; smoothActiveBlocks / activateBlock / solveLevel,
Smoother.cuh:97-291) against the oracle's, which test_oracle_adaptive.py pins.

* One smoothing call (one iteration: block colour 0, then 1) on levels 0-2
  of the 3-level pool from identical FP32 inputs: the per-block sweep counts
  and next-active flags equal the oracle's on at least 99% of the blocks (an
  early-exit or activation test can fall either side of its threshold within
  FP32 rounding), and x within 1e-5 on the blocks whose decisions, and those
  of the colour-0 face neighbours they read, agree.
* theta = 0 with one sweep per block is the plain smoother (local_sweeps 1).
* A whole step with stationary adaptive V-cycles inside FP64 iterative
  refinement to 1e-8 matches the oracle's step (PCG to 1e-10): positions
  1e-5 absolute, velocities 1e-3 of max |v|.
* Conjugate gradients with the adaptive smoother is rejected.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
import oracle as O
from helpers import DT, GRAV, GpuParticles, gpu_grid, oracle_grid, particles, step_checked

pytestmark = pytest.mark.gpu

NAME = "pool64"


def _setup(ctx, **kw):
    pos, vel = particles(NAME)
    og = oracle_grid(NAME)
    og.build_topology(pos)
    og.p2g(pos, vel, DT, GRAV)
    og.divergence()
    og.mg_build()
    gg = gpu_grid(ctx, NAME)
    gp = GpuParticles(ctx, pos, vel)
    gg.build_topology(gp.pos.ptr, gp.n)
    aw.P2GTransfer.bin(gg, gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n)
    aw.P2GTransfer.run(gg, gp.pos.ptr, gp.vel.ptr, gp.n, DT, GRAV)
    aw.G2PManager.divergence(gg)
    sv = aw.VCycleSolver(gg, method=1, prolong_scale=0.75, **kw)
    sv.setup()
    return og, gg, gp, sv


def _blocks(og, k):
    info = og.mg_info(k)
    D, bs = info["dims"], info["bs"]
    e, coff = og.mg_e(k)
    blk = np.repeat(np.arange(info["nb"]), np.diff(coff.astype(np.int64)))
    Bc = np.array([[b % D[0], (b // D[0]) % D[1], b // (D[0] * D[1])] for b in range(info["nb"])])
    nbrs = []
    for b in range(info["nb"]):
        lst = []
        for a in range(3):
            for sd in (-1, 1):
                q = Bc[b].copy()
                q[a] += sd
                if 0 <= q[a] < D[a]:
                    lst.append(q[0] + D[0] * (q[1] + D[1] * q[2]))
        nbrs.append(lst)
    return blk, Bc.sum(1) & 1, nbrs


@pytest.mark.parametrize("k", [0, 1, 2])
def test_one_call_against_the_oracle(ctx, k):
    theta = 0.03
    og, gg, gp, sv = _setup(ctx, adaptive=1, theta=theta)
    og.set_adaptive(True, theta=theta)
    n = og.mg_info(k)["ncells"]
    rng = np.random.default_rng(20 + k)
    liq = og.mg_get(k, "liquid") > 0
    # x = 0 on air cells, the solver's invariant (Dirichlet): the GPU stencil reads them
    x0 = np.where(liq, rng.standard_normal(n) * 0.1, 0.0).astype(np.float32)
    b = (rng.standard_normal(n) * np.where(rng.random(n) < 0.3, 5.0, 1.0)).astype(np.float32)
    og.mg_set(k, "x", x0.astype(np.float64))
    og.mg_set(k, "b", b.astype(np.float64))
    og.smooth(k, 1, False)
    sv.level_set(k, "x", x0)
    sv.level_set(k, "b", b)
    sv.smooth(k, 1, False)
    sw_o, act_o = og.mg_adaptive(k)
    sw_g, act_g = sv.level_get(k, "block_sweeps"), sv.level_get(k, "active")
    agree = sw_o == sw_g
    f_sw, f_act = agree.mean(), (act_o == act_g).mean()
    print(f"[adaptive] level {k}: {len(sw_o)} blocks, sweeps {np.bincount(sw_o)} (oracle), agree {f_sw:.4f}, "
          f"active agree {f_act:.4f}", flush=True)
    assert f_sw >= 0.99 and f_act >= 0.99
    assert len(np.unique(sw_o[sw_o > 0])) >= 2
    blk, col, nbrs = _blocks(og, k)
    clean = agree.copy()
    for q in np.nonzero(col == 1)[0]:
        clean[q] = clean[q] and all(agree[m] for m in nbrs[q])
    cells = clean[blk]
    xo, xg = og.mg_get(k, "x"), sv.level_get(k, "x").astype(np.float64)
    e = np.abs(xg - xo)[cells].max() / np.abs(xo).max()
    print(f"[adaptive] level {k}: x on {cells.mean():.3f} of the cells within {e:.1e}", flush=True)
    assert e < 1e-5


def test_theta_zero_one_sweep_is_the_plain_smoother(ctx):
    outs = []
    for adaptive in (0, 1):
        og, gg, gp, sv = _setup(ctx, adaptive=adaptive, theta=0.0, max_block_sweeps=1, local=1)
        n = og.mg_info(0)["ncells"]
        rng = np.random.default_rng(5)
        liq = og.mg_get(0, "liquid") > 0
        sv.level_set(0, "x", np.where(liq, rng.standard_normal(n) * 0.1, 0.0).astype(np.float32))
        sv.level_set(0, "b", rng.standard_normal(n).astype(np.float32))
        sv.smooth(0, 3, False)
        outs.append(sv.level_get(0, "x").astype(np.float64))
    assert np.abs(outs[1] - outs[0]).max() <= 1e-6 * np.abs(outs[0]).max()


def test_adaptive_stationary_step(ctx):
    pos, vel = particles(NAME)
    og = oracle_grid(NAME)
    ref = og.step(pos, vel, DT, GRAV, 0.95, tol=1e-10, max_iters=400)
    assert ref["status"] == 0
    gg = gpu_grid(ctx, NAME)
    sv = aw.VCycleSolver(gg, method=1, prolong_scale=0.75, precision=1, adaptive=1)
    gp = GpuParticles(ctx, pos, vel)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-8, max_iters=400)
    st = step_checked(step, gp, precision=1)
    print(f"[adaptive] step: {st.solve.outer_iterations} refinements, {st.solve.iterations} V-cycles, "
          f"residual {st.solve.rel_residual:.1e}", flush=True)
    P, V = gp.unpermuted()
    assert np.abs(P - ref["pos"]).max() < 1e-5
    assert np.abs(V - ref["vel"]).max() / max(np.abs(ref["vel"]).max(), 1e-6) < 1e-3


def test_pcg_with_adaptive_smoothing_rejected(ctx):
    gg = gpu_grid(ctx, NAME)
    with pytest.raises(aw.AWError) as e:
        aw.VCycleSolver(gg, method=0, adaptive=1)
    assert e.value.status == aw.AW_ERR_INVALID
