"""GPU adjoint of the step against the FP64 CPU adjoint (oracle), stage by
stage, and a checkpointed rollout gradient against the FP64 oracle's reverse
pass and central finite differences of the FP64 forward.

The oracle adjoint is itself pinned by central differences
(tests/test_oracle_adjoint.py).  Bars:
* stage fields and particle partials: 1e-4 of the field's max magnitude;
  the position partials of G2P + RK3 are piecewise smooth (trilinear weights
  kink on lattice planes), so a particle whose FP32 stage position lies within
  rounding of a kink may take the other one-sided derivative: such a particle
  must match the oracle's partial at the position shifted by -/+3e-7 along
  (1, 1, 1) instead, and at most 1 in 10^4 particles (at least 2) may do so;
* rollout gradient dL/dv0: 1e-3 relative (max norm) against the oracle's;
* directional derivatives against central differences (eps 1e-3, 32 seeded
  directions, topology and liquid masks frozen at the GPU's): 1e-3 relative.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
import oracle as O
from anitowave_adaptive_amd import scenes
from helpers import DT, GRAV, GpuParticles, gpu_grid, oracle_grid, particles, rel_err

pytestmark = pytest.mark.gpu

ALPHA = 0.95
STAGE_TOL = 1e-4


KINK_DELTA = 3e-7  # absolute shift, 2e-5 cells at h = 1/64: beyond FP32 rounding, below the smooth change


def _one_sided(og, pos, vel, xo, vo):
    """Oracle G2P + RK3 position partials at pos shifted by -/+KINK_DELTA along
    (1, 1, 1).  Call before the base-point adjoint: adj_g2p overwrites the
    oracle's ubar fields."""
    p64, v64 = pos.astype(np.float64), vel.astype(np.float64)
    return [og.adj_g2p(p64 + sgn * KINK_DELTA, v64, xo, vo, ALPHA, DT)[0] for sgn in (-1.0, 1.0)]


def _kink_aware(got, base, alts, tol, what):
    """Per-particle check of (n, 3) partials against the oracle's base point,
    with the one-sided alternatives for the few particles on a kink; returns
    the indices that needed an alternative."""
    got = np.asarray(got, np.float64)
    scale = max(np.abs(base).max(), 1e-30)
    err = np.abs(got - base).max(axis=1) / scale
    bad = np.flatnonzero(err >= tol)
    for i in bad:
        alt = min(np.abs(got[i] - a[i]).max() for a in alts) / scale
        assert alt < tol, (what, int(i), float(err[i]), float(alt))
    assert len(bad) <= max(2, len(got) // 10000), (what, len(bad))
    return bad


def _liquid_rel(a, b, liq):
    a = np.asarray(a, np.float64)[liq]
    b = np.asarray(b, np.float64)[liq]
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("name", ["dam32", "pool64", "slosh32", "double64"])
def test_step_adjoint_stages_match_oracle(ctx, name):
    print(f"[adjoint] stages {name}", flush=True)
    pos, vel = particles(name)
    rng = np.random.default_rng(21)
    vel = (vel + 0.3 * rng.standard_normal(vel.shape)).astype(np.float32)
    n = len(pos)
    og = oracle_grid(name)
    r = og.step(pos, vel, DT, GRAV, ALPHA, tol=1e-10, max_iters=400)
    assert r["status"] == 0
    gg = gpu_grid(ctx, name)
    sv = aw.VCycleSolver(gg, precision=1)
    gp = GpuParticles(ctx, pos, vel)
    stepper = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=ALPHA, tol=1e-10, max_iters=400)
    stepper(gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, n)
    np.testing.assert_array_equal(gg.cell("liquid"), og.cell("liquid").astype(np.uint8))

    xo = rng.standard_normal((n, 3))
    vo = rng.standard_normal((n, 3))
    dxb = aw.DeviceArray.from_numpy(ctx, xo.astype(np.float32))
    dvb = aw.DeviceArray.from_numpy(ctx, vo.astype(np.float32))
    adj = aw.Adjoint(gg, sv)

    # G2P + RK3
    adj.g2p_advect(dxb.ptr, dvb.ptr, ALPHA, DT)
    alts = _one_sided(og, pos, vel, xo, vo)
    xi, vi = og.adj_g2p(pos, vel, xo, vo, ALPHA, DT)
    for f in ("ubar_new", "ubar_old"):
        for a in range(3):
            e = rel_err(adj.face(f, a), og.face(f, a))
            assert e < STAGE_TOL, (f, a, e)
    gx = dxb.download(np.float32, 3 * n).reshape(-1, 3)
    gv = dvb.download(np.float32, 3 * n).reshape(-1, 3)
    kinks = _kink_aware(gx, xi, alts, STAGE_TOL, "g2p dL/dx")
    print(f"[adjoint] {name}: g2p dL/dx err {rel_err(gx, xi):.2e}, {len(kinks)} particle(s) on a kink", flush=True)
    assert rel_err(gv, vi) < STAGE_TOL
    keep = np.ones(n, bool)
    keep[kinks] = False
    gx_before, gv_before = gx.astype(np.float64), gv.astype(np.float64)

    # projection: pbar, adjoint solve, D^T
    st = adj.project(tol=1e-10, max_iters=400)
    og.adj_project(tol=1e-12)
    liq = og.cell("liquid") > 0
    assert _liquid_rel(adj.cell("pbar"), og.cell("pbar"), liq) < STAGE_TOL
    e_rhs = rel_err(adj.cell("rhsbar"), og.cell("rhsbar"))
    print(f"[adjoint] {name}: adjoint solve {st.iterations} iterations, rel res {st.rel_residual:.2e}, "
          f"rhsbar err {e_rhs:.2e}", flush=True)
    assert e_rhs < STAGE_TOL
    for a in range(3):
        assert rel_err(adj.face("ubar_star", a), og.face("ubar_star", a)) < STAGE_TOL, a

    # P2G: fold adjoint and deposit partials, alone (the oracle's P2G adjoint
    # accumulated onto the GPU's own G2P partials) and for the whole step
    adj.p2g(dxb.ptr, dvb.ptr)
    og.adj_p2g(pos, vel, gx_before, gv_before)
    og.adj_p2g(pos, vel, xi, vi)
    for f in ("Mbar", "mbar"):
        for a in range(3):
            e = rel_err(adj.face(f, a), og.face(f, a))
            assert e < STAGE_TOL, (f, a, e)
    gx = dxb.download(np.float32, 3 * n).reshape(-1, 3)
    gv = dvb.download(np.float32, 3 * n).reshape(-1, 3)
    sx, sv_ = rel_err(gx, gx_before), rel_err(gv, gv_before)
    ex, ev = rel_err(gx[keep], xi[keep]), rel_err(gv, vi)
    print(f"[adjoint] {name}: p2g stage err x {sx:.2e} v {sv_:.2e}; step dL/dx err {ex:.2e}, dL/dv err {ev:.2e}",
          flush=True)
    assert sx < STAGE_TOL and sv_ < STAGE_TOL
    assert ex < STAGE_TOL and ev < STAGE_TOL


def _com_loss(x, target):
    c = x.mean(axis=0) - target
    return float(c @ c), np.broadcast_to(2.0 * c / len(x), x.shape).copy()


def test_rollout_gradient_matches_oracle_and_central_differences(ctx):
    """Sloshing pool 32^3, 2 native levels, 8 steps, checkpoint every 4, loss
    |COM(x_8) - x*|^2, gradient w.r.t. v0 (mixed-precision solves to 1e-10)."""
    n_cells, levels, band = 32, 2, (0, 1, 0, 0)
    pos, vel = scenes.slosh(n=n_cells, ppc=2, seed=3)
    rng = np.random.default_rng(4)
    vel = (vel + 0.2 * rng.standard_normal(vel.shape)).astype(np.float32)
    n = len(pos)
    steps, every = 8, 4
    target = np.array([0.6, 0.35, 0.5])
    print(f"[adjoint] rollout: {n} particles, {steps} steps", flush=True)
    gg = aw.MSBGManager(ctx, (n_cells // 8,) * 3, 1.0 / n_cells, num_levels=levels, band=band)
    sv = aw.VCycleSolver(gg, precision=1)
    prm = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=ALPHA, tol=1e-10, max_iters=400).params
    gp = GpuParticles(ctx, pos, vel)
    ro = aw.Rollout(gg, sv, n, steps, checkpoint_every=every, keep_masks=True)
    ro.forward(prm, gp.pos.ptr, gp.vel.ptr, gp.ids.ptr)
    xK, _ = gp.unpermuted()
    L, xbar = _com_loss(xK.astype(np.float64), target)
    dxb = aw.DeviceArray.from_numpy(ctx, xbar.astype(np.float32))
    dvb = aw.DeviceArray.from_numpy(ctx, np.zeros((n, 3), np.float32))
    st = ro.backward(dxb.ptr, dvb.ptr)
    g_gpu = dvb.download(np.float32, 3 * n).reshape(-1, 3).astype(np.float64)
    ratio = (st.recompute_ms + st.backward_ms) / st.forward_ms
    print(f"[adjoint] forward {st.forward_ms:.1f} ms, recompute {st.recompute_ms:.1f} ms, backward "
          f"{st.backward_ms:.1f} ms, adjoint/forward {ratio:.2f}", flush=True)
    # the recomputation replays the forward pass: same levels and liquid masks
    # every step, and each segment ends on the checkpoint stored there up to
    # the round-off of arrival-order float sums (positions and velocities)
    print(f"[adjoint] replay: {st.replay_mismatched_steps} mismatched steps, max state difference "
          f"{st.replay_max_state_diff:.1e}", flush=True)
    assert st.replay_mismatched_steps == 0
    assert st.replay_max_state_diff < 1e-4 * max(np.abs(vel).max(), 1.0)
    masks = [ro.masks(k) for k in range(steps)]

    og = O.OracleGrid((n_cells // 8,) * 3, num_levels=levels, h=1.0 / n_cells, band=band)
    og.set_params()
    states, _ = O.rollout64(og, pos, vel, steps, DT, GRAV, ALPHA, tol=1e-12, frozen=masks)
    _, xbar64 = _com_loss(states[-1][0], target)
    _, g_or = O.rollout_grad64(og, states, masks, xbar64, np.zeros_like(xbar64), DT, GRAV, ALPHA, tol=1e-12)
    e = rel_err(g_gpu, g_or)
    print(f"[adjoint] rollout dL/dv0: GPU vs FP64 oracle adjoint {e:.2e}", flush=True)
    assert e < 1e-3

    eps = 1e-3
    worst = 0.0
    for k in range(32):
        d = np.random.default_rng(100 + k).standard_normal(vel.shape)
        d /= np.linalg.norm(d)
        xp = O.rollout64(og, pos, vel + eps * d, steps, DT, GRAV, ALPHA, tol=1e-12, frozen=masks)[0][-1][0]
        xm = O.rollout64(og, pos, vel - eps * d, steps, DT, GRAV, ALPHA, tol=1e-12, frozen=masks)[0][-1][0]
        fd = (_com_loss(xp, target)[0] - _com_loss(xm, target)[0]) / (2 * eps)
        ad = float((g_gpu * d).sum())
        worst = max(worst, abs(fd - ad) / abs(fd))
    print(f"[adjoint] 32 directions: worst |fd - adj| / |fd| = {worst:.2e}", flush=True)
    assert worst < 1e-3
