"""The five workloads at their own sizes on the HIP path, against the FP64
oracle on the same inputs.

* dam64 (64^3, 1 level, 524,288 particles): 50 steps; every step keeps the
  solver's convergence contract, and at steps 0, 25 and 49 one oracle step
  from the GPU's current particles matches the GPU's step (positions 1e-5
  absolute, velocities 1e-3 of max |v|).
* pool256 (256^3 effective, 3 levels, 40.9 M particles; the workload text
  says ~16 M, 8 particles per fine liquid cell give 40.9 M): the first step
  and the step at impact of the falling sphere (step 30, the splash phase)
  against one oracle step from the same particles, same bars; topology
  bit-exact on both.
* pool256 mixed precision to 1e-8 and 1e-10: the FP64 solution's residual
  evaluated with the oracle's independently assembled FP64 operator on the
  GPU's own right-hand side is at the tolerance; pressure, u_new and
  position errors against the oracle's FP64 step are reported and bounded.
* slosh128 (128^3 effective, 2 levels, 6.7 M particles): the 16-step
  rollout gradient dL/dv0 with checkpoints every 4 steps against the FP64
  oracle's adjoint along the GPU's own states (5e-4) and along the FP64
  trajectory (2e-3: FP32 drift across weight kinks), directional derivatives
  along 32 seeded directions against the FP64 adjoint's, and central
  differences (eps 1e-3) of the FP64 forward with the GPU's frozen masks
  along 3 of them, 1e-3 (the 32-direction central-difference check runs at
  32^3 in test_gpu_adjoint.py: each direction costs two 16-step FP64
  rollouts here).
* double512 (512^3 effective, 4 levels, 116 M particles): one plain FP32
  step and one mixed-precision step to 1e-8 against the oracle's step, the
  topology bit-exact and 2:1 graded.
"""
import time

import numpy as np
import pytest

import anitowave_adaptive_amd as aw
import oracle as O
from anitowave_adaptive_amd import scenes
from helpers import DT, GRAV, GpuParticles, rel_err, step_checked

pytestmark = pytest.mark.gpu

ALPHA = 0.95


def log(msg):
    print(msg, flush=True)


def setup(ctx, name, precision=0, tol=1e-6, max_iters=200):
    w = scenes.WORKLOADS[name]
    t = time.perf_counter()
    pos, vel = w["scene"](n=w["n"])
    log(f"[configs] {name}: {len(pos)} particles seeded in {time.perf_counter() - t:.1f}s")
    gg = aw.MSBGManager(ctx, w["dims"], 1.0 / w["n"], num_levels=w["levels"], band=w["band"])
    sv = aw.VCycleSolver(gg, precision=precision)
    og = O.OracleGrid(w["dims"], num_levels=w["levels"], h=1.0 / w["n"], band=w["band"])
    og.set_params()
    gp = GpuParticles(ctx, pos, vel)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=ALPHA, tol=tol, max_iters=max_iters)
    return w, pos, vel, gg, sv, og, gp, step


def check_step_against_oracle(og, gg, gp, step, tol, precision=0, what=""):
    """One GPU step from the particles' current state and one oracle step
    from the same particles (same order); compared by particle id."""
    p, v, ids = gp.get()
    t = time.perf_counter()
    ref = og.step(p, v, DT, GRAV, ALPHA, tol=tol, max_iters=400)
    t_or = time.perf_counter() - t
    assert ref["status"] == 0
    st = step_checked(step, gp, precision)
    np.testing.assert_array_equal(gg.levels(), og.levels())
    P, V = gp.unpermuted()
    Pr = np.empty_like(ref["pos"])
    Vr = np.empty_like(ref["vel"])
    Pr[ids] = ref["pos"]
    Vr[ids] = ref["vel"]
    dx = float(np.abs(P - Pr).max())
    vmax = max(float(np.abs(Vr).max()), 1e-6)
    dv = float(np.abs(V - Vr).max()) / vmax
    log(f"[configs] {what}: iterations gpu {st.solve.iterations} oracle {ref['iters']}, true rel. residual "
        f"{st.solve.rel_residual:.2e}, max |dx| {dx:.2e}, max |dv| / max |v| {dv:.2e} (oracle {t_or:.1f}s)")
    assert dx < 1e-5, dx
    assert dv < 1e-3, dv
    return st


def test_dam64_50_steps(ctx):
    w, pos, vel, gg, sv, og, gp, step = setup(ctx, "dam64")
    n = len(pos)
    floors = 0
    for k in range(w["steps"]):
        if k in (0, 25, w["steps"] - 1):
            st = check_step_against_oracle(og, gg, gp, step, 1e-6, what=f"dam64 step {k}")
        else:
            st = step_checked(step, gp)
        floors += step.status != aw.AW_OK
        assert st.total_cells == 262144
    P, V = gp.unpermuted()
    assert P.shape == (n, 3) and np.all(np.isfinite(P)) and np.all(np.isfinite(V))
    assert P.min() >= 0.0 and P.max() <= 1.0
    # the column has started to collapse: mean x moves right, mean y drops
    assert P[:, 0].mean() > pos[:, 0].mean() and P[:, 1].mean() < pos[:, 1].mean()
    log(f"[configs] dam64: {w['steps']} steps, {floors} at the FP32 floor")


def test_pool256_first_step_and_impact(ctx):
    w, pos, vel, gg, sv, og, gp, step = setup(ctx, "pool256")
    check_step_against_oracle(og, gg, gp, step, 1e-6, what="pool256 step 0")
    lv0 = np.bincount(gg.levels(), minlength=3)
    for k in range(1, 30):
        step_checked(step, gp)
    st = check_step_against_oracle(og, gg, gp, step, 1e-6, what="pool256 step 30 (impact)")
    lv = np.bincount(gg.levels(), minlength=3)
    log(f"[configs] pool256 blocks by level: step 0 {lv0.tolist()}, step 30 {lv.tolist()}")
    assert st.liquid_cells > 0 and lv[0] > 0


@pytest.mark.parametrize("tol", [1e-8, 1e-10])
def test_pool256_mixed_precision(ctx, tol):
    w, pos, vel, gg, sv, og, gp, step = setup(ctx, "pool256", precision=1, tol=tol, max_iters=400)
    ref = og.step(pos, vel, DT, GRAV, ALPHA, tol=min(tol, 1e-10), max_iters=400)
    assert ref["status"] == 0
    p_or = og.cell("p")
    un_or = [og.face("u_new", a) for a in range(3)]
    st = step_checked(step, gp, precision=1)
    assert step.status == aw.AW_OK
    liq = gg.cell("liquid") > 0
    np.testing.assert_array_equal(liq, og.cell("liquid") > 0)
    # the FP64 solution's residual with the oracle's FP64 operator, on the GPU's right-hand side
    x64 = sv.level_get(0, "x64")
    b = gg.cell("rhs").astype(np.float64)
    r = np.where(liq, b - og.apply(0, x64), 0.0)
    rel = float(np.linalg.norm(r) / np.linalg.norm(b[liq]))
    ep = rel_err(x64[liq], p_or[liq])
    eu = max(rel_err(gg.face("u_new", a), un_or[a]) for a in range(3))
    P, V = gp.unpermuted()
    ex = float(np.abs(P - ref["pos"]).max())
    log(f"[configs] pool256 mixed {tol:g}: {st.solve.outer_iterations} refinements, {st.solve.iterations} inner "
        f"iterations; residual by the oracle's operator {rel:.2e} (reported {st.solve.rel_residual:.2e}); vs the "
        f"FP64 oracle step: pressure {ep:.2e}, u_new {eu:.2e}, positions {ex:.2e}")
    assert rel <= 1.05 * tol, rel
    assert ep < 1e-4 and eu < 1e-4 and ex < 1e-5


def _com_loss(x, target):
    c = x.mean(axis=0) - target
    return float(c @ c), np.broadcast_to(2.0 * c / len(x), x.shape).copy()


def _quadratic_basis(pos):
    """The 10 monomials of degree <= 2 in the particles' initial positions
    (centred on the domain)."""
    x, y, z = (pos.astype(np.float64) - 0.5).T
    return np.column_stack([np.ones_like(x), x, y, z, x * x, y * y, z * z, x * y, y * z, z * x])


def _smooth_direction(basis, seed):
    """A seeded random unit direction in initial-velocity space that varies
    smoothly over the particles: per component, a random quadratic
    polynomial of the initial position."""
    d = basis @ np.random.default_rng(seed).standard_normal((basis.shape[1], 3))
    return d / np.linalg.norm(d)


def test_slosh128_rollout_gradient(ctx):
    """The GPU gradient is checked three ways.  (1) Against the FP64 oracle
    adjoint evaluated along the GPU's own forward states (each step's
    topology and liquid mask recomputed by the oracle from the GPU state,
    bit-exact): the adjoint implementations compared on identical inputs.
    (2) Against the FP64 adjoint along the FP64 oracle's own trajectory (GPU
    masks frozen): this adds the FP32 forward's drift (about 1e-7 in
    position), which moves a few particles across kinks of the trilinear
    weights (a one-sided derivative changes there), so this bar is 2e-3.
    (3) Directional derivatives along 32 seeded random directions that vary
    smoothly over the particles against the FP64 adjoint's, and against
    central differences of the FP64 forward along 3 of them (eps 1e-3):
    1e-3 relative.  White-noise directions, whose projections cancel down to
    the kink noise, are compared against (1)'s gradient at 5e-3."""
    w = scenes.WORKLOADS["slosh128"]
    pos, vel = w["scene"](n=w["n"])
    n = len(pos)
    steps, every = 16, 4
    target = np.array([0.6, 0.35, 0.5])
    log(f"[configs] slosh128: {n} particles, {steps} steps, checkpoint every {every}")
    gg = aw.MSBGManager(ctx, w["dims"], 1.0 / w["n"], num_levels=w["levels"], band=w["band"])
    sv = aw.VCycleSolver(gg, precision=1)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=ALPHA, tol=1e-10, max_iters=400)
    prm = step.params
    # the GPU's forward states, step by step (the rollout's steps are the same aw_step calls)
    gp = GpuParticles(ctx, pos, vel)
    states_gpu = [(pos.astype(np.float64), vel.astype(np.float64))]
    for k in range(steps):
        step_checked(step, gp, precision=1)
        P, V = gp.unpermuted()
        states_gpu.append((P.astype(np.float64), V.astype(np.float64)))
    gp = GpuParticles(ctx, pos, vel)
    ro = aw.Rollout(gg, sv, n, steps, checkpoint_every=every, keep_masks=True)
    ro.forward(prm, gp.pos.ptr, gp.vel.ptr, gp.ids.ptr)
    xK, _ = gp.unpermuted()
    _, xbar = _com_loss(xK.astype(np.float64), target)
    dxb = aw.DeviceArray.from_numpy(ctx, xbar.astype(np.float32))
    dvb = aw.DeviceArray.from_numpy(ctx, np.zeros((n, 3), np.float32))
    st = ro.backward(dxb.ptr, dvb.ptr)
    g_gpu = dvb.download(np.float32, 3 * n).reshape(-1, 3).astype(np.float64)
    ratio = (st.recompute_ms + st.backward_ms) / st.forward_ms
    log(f"[configs] slosh128 GPU: forward {st.forward_ms:.1f} ms, recompute {st.recompute_ms:.1f} ms, backward "
        f"{st.backward_ms:.1f} ms, adjoint/forward {ratio:.2f}; replay: {st.replay_mismatched_steps} mismatched "
        f"steps, max state difference {st.replay_max_state_diff:.1e}")
    masks = [ro.masks(k) for k in range(steps)]
    og = O.OracleGrid(w["dims"], num_levels=w["levels"], h=1.0 / w["n"], band=w["band"])
    og.set_params()
    t = time.perf_counter()
    _, xbar_g = _com_loss(states_gpu[-1][0], target)
    _, g_traj = O.rollout_grad64(og, states_gpu, [(None, None)] * steps, xbar_g, np.zeros_like(xbar_g), DT, GRAV,
                                 ALPHA, tol=1e-12)
    e1 = rel_err(g_gpu, g_traj)
    log(f"[configs] slosh128 (1) dL/dv0 vs the FP64 adjoint along the GPU states: {e1:.2e} "
        f"({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    states, _ = O.rollout64(og, pos, vel, steps, DT, GRAV, ALPHA, tol=1e-12, frozen=masks)
    _, xbar64 = _com_loss(states[-1][0], target)
    _, g_or = O.rollout_grad64(og, states, masks, xbar64, np.zeros_like(xbar64), DT, GRAV, ALPHA, tol=1e-12)
    e2 = rel_err(g_gpu, g_or)
    drift = max(float(np.abs(a[0] - b[0]).max()) for a, b in zip(states_gpu, states))
    log(f"[configs] slosh128 (2) dL/dv0 vs the FP64 adjoint along the FP64 trajectory: {e2:.2e} (trajectory "
        f"drift {drift:.1e}; the two FP64 adjoints differ by {rel_err(g_traj, g_or):.2e}) "
        f"({time.perf_counter() - t:.1f}s)")
    worst_w = 0.0
    for k in range(32):
        d = np.random.default_rng(100 + k).standard_normal(vel.shape)
        d /= np.linalg.norm(d)
        a, b = float((g_gpu * d).sum()), float((g_traj * d).sum())
        worst_w = max(worst_w, abs(a - b) / abs(b))
    worst_s, dirs = 0.0, []
    basis = _quadratic_basis(pos)
    for k in range(32):
        d = _smooth_direction(basis, 200 + k)
        dirs.append(d)
        a, b = float((g_gpu * d).sum()), float((g_or * d).sum())
        worst_s = max(worst_s, abs(a - b) / abs(b))
    log(f"[configs] slosh128 (3) 32 white-noise directions vs (1): worst {worst_w:.2e}; 32 smooth directions vs "
        f"the FP64 trajectory's adjoint: worst {worst_s:.2e}")
    eps = 1e-3
    worst_fd = 0.0
    for k in range(3):
        d = dirs[k]
        xp = O.rollout64(og, pos, vel + eps * d, steps, DT, GRAV, ALPHA, tol=1e-12, frozen=masks)[0][-1][0]
        xm = O.rollout64(og, pos, vel - eps * d, steps, DT, GRAV, ALPHA, tol=1e-12, frozen=masks)[0][-1][0]
        fd = (_com_loss(xp, target)[0] - _com_loss(xm, target)[0]) / (2 * eps)
        ad = float((g_gpu * d).sum())
        worst_fd = max(worst_fd, abs(fd - ad) / abs(fd))
        log(f"[configs] slosh128 direction {k}: central difference {fd:.6e}, GPU adjoint {ad:.6e}")
    log(f"[configs] slosh128 central differences: worst {worst_fd:.2e}")
    # (1) and the white-noise projections keep a kink term of their own: the
    # rollout's backward pass recomputes its segments, and P2G's LDS float sums
    # are not order-deterministic, so the recomputed states differ from the
    # recorded ones by rounding (measured 1.8e-4 and 1.8e-3)
    assert e1 < 5e-4
    assert e2 < 2e-3
    assert worst_w < 5e-3
    assert worst_s < 1e-3
    assert worst_fd < 1e-3
    assert ratio < 3.0


def test_double512_plain_and_mixed(ctx):
    w, pos, vel, gg, sv, og, gp, step = setup(ctx, "double512")
    check_step_against_oracle(og, gg, gp, step, 1e-6, what="double512 plain step 0")
    lv = gg.levels().reshape(64, 64, 64).astype(int)
    for ax in range(3):
        assert np.abs(np.diff(lv, axis=ax)).max() <= 1
    counts = np.bincount(lv.reshape(-1), minlength=4)
    log(f"[configs] double512 blocks by level {counts.tolist()}")
    assert np.all(counts > 0)
    assert int(gg.particle_counts().sum()) == len(pos)
    # mixed precision to 1e-8 from the state after the plain step
    sv2 = aw.VCycleSolver(gg, precision=1)
    step2 = aw.FlipStep(gg, sv2, dt=DT, gravity=GRAV, alpha=ALPHA, tol=1e-8, max_iters=400)
    st = check_step_against_oracle(og, gg, gp, step2, 1e-8, precision=1, what="double512 mixed 1e-8 step 1")
    assert step2.status == aw.AW_OK and st.solve.rel_residual <= 1e-8
