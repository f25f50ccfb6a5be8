"""Error paths of the step and the solver (each leaves the objects usable):

* a NaN particle velocity: the right-hand side is not finite, aw_step returns
  AW_ERR_NAN before any solver iteration; the next step with clean particles
  matches the oracle's;
* an iteration cap below what the tolerance needs: AW_ERR_NOT_CONVERGED with
  the iterations capped and the true relative residual above the tolerance,
  for conjugate gradients and stationary V-cycles, FP32 and FP64-refined;
* a memory limit below what a grid needs (aw_set_memory_limit): AW_ERR_OOM,
  the library's held bytes back where they were once the objects are
  destroyed, and the same work succeeding after the limit is lifted; an
  aw_malloc past the limit fails the same way.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
from helpers import DT, GRAV, GpuParticles, gpu_grid, oracle_grid, particles, step_checked

pytestmark = pytest.mark.gpu

NAME = "pool64"


def test_nan_velocity_gives_err_nan(ctx):
    pos, vel = particles(NAME)
    bad = vel.copy()
    bad[len(bad) // 2, 1] = np.nan
    gg = gpu_grid(ctx, NAME)
    sv = aw.VCycleSolver(gg)
    gp = GpuParticles(ctx, pos, bad)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6)
    with pytest.raises(aw.AWError) as e:
        step(gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n)
    print(f"[errors] NaN velocity: {e.value}", flush=True)
    assert e.value.status == aw.AW_ERR_NAN
    assert step.status == aw.AW_ERR_NAN
    # the same grid and solver, clean particles: the oracle's step
    gp2 = GpuParticles(ctx, pos, vel)
    step_checked(step, gp2)
    ref = oracle_grid(NAME).step(pos, vel, DT, GRAV, 0.95, tol=1e-6)
    P, V = gp2.unpermuted()
    assert np.abs(P - ref["pos"]).max() < 1e-5
    assert np.abs(V - ref["vel"]).max() / np.abs(ref["vel"]).max() < 1e-3


@pytest.mark.parametrize("method,precision", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_iteration_cap_gives_not_converged(ctx, method, precision):
    pos, vel = particles(NAME)
    gg = gpu_grid(ctx, NAME)
    kw = dict(prolong_scale=0.75) if method == 1 else {}
    sv = aw.VCycleSolver(gg, method=method, precision=precision, **kw)
    gp = GpuParticles(ctx, pos, vel)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-10, max_iters=2)
    with pytest.raises(aw.AWError) as e:
        step(gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n)
    assert e.value.status == aw.AW_ERR_NOT_CONVERGED
    st = step(gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n, allow_not_converged=True)
    print(f"[errors] cap 2, method {method} precision {precision}: {st.solve.iterations} iterations, "
          f"residual {st.solve.rel_residual:.1e}", flush=True)
    assert step.status == aw.AW_ERR_NOT_CONVERGED
    assert 0 < st.solve.iterations <= 2
    assert 1e-10 < st.solve.rel_residual < 1.0
    # with room to converge the same objects reach the tolerance
    step2 = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-5 if precision == 0 else 1e-8,
                        max_iters=400)
    step_checked(step2, gp, precision=precision)


def test_memory_limit_gives_err_oom(ctx):
    pos, vel = particles(NAME)
    gp = GpuParticles(ctx, pos, vel)
    held = aw.memory_stats()[0]
    try:
        aw.set_memory_limit(held + (1 << 20))
        objs = []
        with pytest.raises(aw.AWError) as e:
            gg = gpu_grid(ctx, NAME)
            objs.append(gg)
            sv = aw.VCycleSolver(gg)
            objs.append(sv)
            step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6)
            step(gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n)
        print(f"[errors] limit held + 1 MiB: {e.value}", flush=True)
        assert e.value.status == aw.AW_ERR_OOM
        assert aw.memory_stats()[0] <= held + (1 << 20)
        for o in reversed(objs):
            o.close()
        assert aw.memory_stats()[0] == held
        with pytest.raises(aw.AWError) as e:
            aw.DeviceArray(ctx, 2 << 20)
        assert e.value.status == aw.AW_ERR_OOM
        assert aw.memory_stats()[0] == held
    finally:
        aw.set_memory_limit(0)
    gg = gpu_grid(ctx, NAME)
    sv = aw.VCycleSolver(gg)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6)
    step_checked(step, gp)
