"""Two-phase / variable-density coefficients on the HIP path against the
FP64 oracle (aw_step_params.d_density; beta = 1 / rho from the density
deposit, coarsened through the multigrid levels).

* One step with densities per particle on the 3-level pool (a heavy sphere
  falling into a light pool) and on the 4-level double dam (random
  densities): the level-0 beta within 1e-5 of the oracle's, the operator of
  every multigrid level (applied to a random vector on its liquid cells)
  within 1e-5 (this is the coefficient coarsening, level by level, the tail
  levels included), and the particles after the step (positions 1e-5
  absolute, velocities 1e-3 of max |v|).
* Stratified column at rest (2 x 2 x 2 lattice particles per cell, density
  2 under 0.5): mixed precision to 1e-10 gives the closed-form pressure
  steps rho dt |g| per cell in each layer (1e-5) and u_new = 0.
* A step without d_density after one with it gives the uniform-beta step
  (the grid and the solver's coarse levels return to beta = 1).
* The adjoint and rollouts reject a variable-density state.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
import oracle as O
from anitowave_adaptive_amd import scenes
from helpers import DT, GRAV, GpuParticles, gpu_grid, oracle_grid, particles, rel_err, step_checked

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _densities(name, pos):
    if name == "pool64":
        c = np.array([0.5, 0.7, 0.5])
        sphere = ((pos - c) ** 2).sum(-1) < 0.1 ** 2
        return np.where(sphere, 3.0, 0.8).astype(np.float32)
    return np.random.default_rng(2).uniform(0.5, 2.5, len(pos)).astype(np.float32)


def _liquid_rel(a, b, liq):
    a, b = np.asarray(a, np.float64)[liq], np.asarray(b, np.float64)[liq]
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("name", ["pool64", "double64"])
def test_density_step_parity(ctx, name):
    pos, vel = particles(name)
    rho = _densities(name, pos)
    print(f"[density] {name}: {len(pos)} particles, densities {rho.min():.2f}..{rho.max():.2f}", flush=True)
    og = oracle_grid(name)
    ref = og.step(pos, vel, DT, GRAV, 0.95, tol=1e-6, density=rho.astype(np.float64))
    assert ref["status"] == 0
    gg = gpu_grid(ctx, name)
    sv = aw.VCycleSolver(gg)
    gp = GpuParticles(ctx, pos, vel)
    dr = aw.DeviceArray.from_numpy(ctx, rho)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6, d_density=dr.ptr)
    st = step_checked(step, gp)
    print(f"[density] {name}: iterations gpu {st.solve.iterations} oracle {ref['iters']}", flush=True)
    np.testing.assert_array_equal(gg.levels(), og.levels())
    for a in range(3):
        e = rel_err(gg.face("beta", a), og.face("beta", a))
        assert e < TOL, (a, e)
    assert sv.num_levels() == og.num_mg_levels()
    rng = np.random.default_rng(4)
    for k in range(sv.num_levels()):
        liq = og.mg_get(k, "liquid") > 0
        np.testing.assert_array_equal(sv.level_get(k, "liquid") > 0, liq)
        if not liq.any():
            continue
        x = np.where(liq, rng.standard_normal(len(liq)), 0.0)
        dx = aw.DeviceArray.from_numpy(ctx, x.astype(np.float32))
        dy = aw.DeviceArray(ctx, 4 * len(liq))
        sv.apply(k, dx.ptr, dy.ptr)
        y = dy.download(np.float32, len(liq))
        e = _liquid_rel(y, og.apply(k, x.astype(np.float32).astype(np.float64)), liq)
        print(f"[density] {name}: level {k} operator {e:.1e}", flush=True)
        assert e < TOL, (k, e)
    P, V = gp.unpermuted()
    assert np.abs(P - ref["pos"]).max() < 1e-5
    assert np.abs(V - ref["vel"]).max() / max(np.abs(ref["vel"]).max(), 1e-6) < 1e-3


def _lattice(n, top):
    q = (np.arange(2) + 0.5) / 2
    cells = np.stack(np.meshgrid(np.arange(n), np.arange(top), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    sub = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    pos = ((cells[:, None, :] + sub[None, :, :]).reshape(-1, 3) / n).astype(np.float32)
    return pos, np.zeros_like(pos)


def test_stratified_column_at_rest(ctx):
    n = 32
    pos, vel = _lattice(n, 16)
    rho = np.where(pos[:, 1] < 0.25, 2.0, 0.5).astype(np.float32)
    gg = aw.MSBGManager(ctx, (4, 4, 4), 1.0 / n, num_levels=1)
    sv = aw.VCycleSolver(gg, precision=1)
    gp = GpuParticles(ctx, pos, vel)
    dr = aw.DeviceArray.from_numpy(ctx, rho)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-10, max_iters=400, d_density=dr.ptr)
    step_checked(step, gp, precision=1)
    p = sv.level_get(0, "x64")
    P = np.zeros((n, n, n))
    k = 0
    for bz in range(4):
        for by in range(4):
            for bx in range(4):
                blk = p[k:k + 512].reshape(8, 8, 8)   # z, y, x inside the block
                P[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = blk
                k += 512
    d = P[:, :-1, :] - P[:, 1:, :]
    g = -GRAV[1]
    np.testing.assert_allclose(d[:, 0:7, :], 2.0 * DT * g, rtol=TOL)
    np.testing.assert_allclose(d[:, 8:15, :], 0.5 * DT * g, rtol=TOL)
    np.testing.assert_allclose(d[:, 7, :], 1.25 * DT * g, rtol=TOL)
    for a in range(3):
        assert np.abs(gg.face("u_new", a)).max() < 1e-5


def test_density_off_again_and_adjoint_rejects(ctx):
    name = "pool64"
    pos, vel = particles(name)
    rho = _densities(name, pos)
    gg = gpu_grid(ctx, name)
    sv = aw.VCycleSolver(gg)
    gp = GpuParticles(ctx, pos, vel)
    dr = aw.DeviceArray.from_numpy(ctx, rho)
    step_d = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6, d_density=dr.ptr)
    step_checked(step_d, gp)
    adj = aw.Adjoint(gg, sv)
    xb = aw.DeviceArray.from_numpy(ctx, np.zeros((gp.n, 3), np.float32))
    vb = aw.DeviceArray.from_numpy(ctx, np.zeros((gp.n, 3), np.float32))
    with pytest.raises(aw.AWError) as e:
        adj.step(step_d.params, xb.ptr, vb.ptr)
    assert e.value.status == aw.AW_ERR_UNSUPPORTED
    # the next step without densities: beta back to 1 on every level
    p1, v1, ids = gp.get()
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6)
    step_checked(step, gp)
    for a in range(3):
        bt = gg.face("beta", a)
        assert set(np.unique(bt)) <= {0.0, 1.0}
    og = oracle_grid(name)
    ref = og.step(p1, v1, DT, GRAV, 0.95, tol=1e-6)
    P, V = gp.unpermuted()
    Pr = np.empty_like(ref["pos"])
    Pr[ids] = ref["pos"]
    assert np.abs(P - Pr).max() < 1e-5
    rng = np.random.default_rng(1)
    for k in range(sv.num_levels()):
        liq = og.mg_get(k, "liquid") > 0
        if not liq.any():
            continue
        x = np.where(liq, rng.standard_normal(len(liq)), 0.0).astype(np.float32)
        dx = aw.DeviceArray.from_numpy(ctx, x)
        dy = aw.DeviceArray(ctx, 4 * len(liq))
        sv.apply(k, dx.ptr, dy.ptr)
        e = _liquid_rel(dy.download(np.float32, len(liq)), og.apply(k, x.astype(np.float64)), liq)
        assert e < TOL, (k, e)
    ro = aw.Rollout(gg, sv, gp.n, 2)
    with pytest.raises(aw.AWError) as e:
        ro.forward(step_d.params, gp.pos.ptr, gp.vel.ptr, gp.ids.ptr)
    assert e.value.status == aw.AW_ERR_UNSUPPORTED
