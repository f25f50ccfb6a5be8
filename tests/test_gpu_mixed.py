"""FP32 V-cycles inside FP64 iterative refinement, against the FP64 oracle.

The pressure solve reaches relative residuals of 1e-8 and 1e-10 in the FP64
sense (residual of the FP64 solution under the oracle's FP64 operator), which
the plain FP32 path cannot (its floor is near 1e-5 on these problems).
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
from helpers import DT, GRAV, GpuParticles, gpu_grid, oracle_grid, particles, rel_err, step_checked

pytestmark = pytest.mark.gpu


def _setup(ctx, name, precision):
    pos, vel = particles(name)
    og = oracle_grid(name)
    og.build_topology(pos)
    og.p2g(pos, vel, DT, GRAV)
    og.divergence()
    og.mg_build()
    gg = gpu_grid(ctx, name)
    gp = GpuParticles(ctx, pos, vel)
    gg.build_topology(gp.pos.ptr, gp.n)
    aw.P2GTransfer.bin(gg, gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n)
    aw.P2GTransfer.run(gg, gp.pos.ptr, gp.vel.ptr, gp.n, DT, GRAV)
    aw.G2PManager.divergence(gg)
    sv = aw.VCycleSolver(gg, precision=precision)
    sv.setup()
    return og, gg, gp, sv


@pytest.mark.parametrize("name", ["dam32", "pool64", "slosh32", "double64"])
@pytest.mark.parametrize("tol", [1e-8, 1e-10])
def test_mixed_precision_reaches_fp64_tolerance(ctx, name, tol):
    print(f"[mixed] {name} tol {tol}", flush=True)
    og, gg, gp, sv = _setup(ctx, name, 1)
    n = len(og.cell("rhs"))
    gg.set_cell("p", np.zeros(n, np.float32))
    st = sv.solve(tol=tol, max_iters=400)
    assert st.rel_residual <= tol
    assert st.outer_iterations >= 2
    x64 = sv.level_get(0, "x64")
    b_gpu = gg.cell("rhs").astype(np.float64)
    liq = gg.cell("liquid") > 0
    r = b_gpu - og.apply(0, x64)
    true_rel = np.linalg.norm(r[liq]) / np.linalg.norm(b_gpu[liq])
    assert true_rel < 10 * tol, true_rel
    # against the oracle's FP64 solution of its own FP64 right-hand side
    og.set_cell("p", np.zeros(n))
    og.solve(tol=1e-12, max_iters=300)
    p_or = og.cell("p")
    err = np.abs(x64 - p_or).max() / np.abs(p_or).max()
    print(f"[mixed] {name} tol {tol}: outer {st.outer_iterations} inner {st.iterations} "
          f"true rel {true_rel:.2e} p err vs oracle {err:.2e}", flush=True)
    assert err < 1e-5


def test_fp32_floor_is_above_mixed(ctx):
    og, gg, gp, sv = _setup(ctx, "pool64", 0)
    n = len(og.cell("rhs"))
    gg.set_cell("p", np.zeros(n, np.float32))
    st32 = sv.solve(tol=1e-6, max_iters=200, allow_not_converged=True)
    # the FP32 floor lies above 1e-6 here: the contract reports it as not converged
    assert sv.status == aw.AW_ERR_NOT_CONVERGED and st32.rel_residual > 1e-6 >= st32.recurrence_rel_residual
    p32 = gg.cell("p").astype(np.float64)
    b_gpu = gg.cell("rhs").astype(np.float64)
    liq = gg.cell("liquid") > 0
    r32 = np.linalg.norm((b_gpu - og.apply(0, p32))[liq]) / np.linalg.norm(b_gpu[liq])
    sv2 = aw.VCycleSolver(gg, precision=1)
    sv2.setup()
    gg.set_cell("p", np.zeros(n, np.float32))
    sv2.solve(tol=1e-10, max_iters=400)
    x64 = sv2.level_get(0, "x64")
    r64 = np.linalg.norm((b_gpu - og.apply(0, x64))[liq]) / np.linalg.norm(b_gpu[liq])
    print(f"[mixed] true residual FP32 {r32:.2e} vs FP64-IR {r64:.2e}", flush=True)
    assert r64 < 1e-3 * r32


@pytest.mark.parametrize("name", ["pool64"])
def test_mixed_step_errors_vs_oracle(ctx, name):
    pos, vel = particles(name)
    og = oracle_grid(name)
    r = og.step(pos, vel, DT, GRAV, 0.95, tol=1e-10, max_iters=300)
    out = {}
    for prec, tol in [(0, 1e-6), (1, 1e-8), (1, 1e-10)]:
        gg = gpu_grid(ctx, name)
        sv = aw.VCycleSolver(gg, precision=prec)
        gp = GpuParticles(ctx, pos, vel)
        st = step_checked(aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=tol, max_iters=400), gp, prec)
        P, V = gp.unpermuted()
        ex = np.abs(P - r["pos"]).max()
        ev = np.abs(V - r["vel"]).max() / np.abs(r["vel"]).max()
        ep = rel_err(gg.cell("p"), og.cell("p"))
        out[(prec, tol)] = (ex, ev, ep)
        print(f"[mixed] step precision {prec} tol {tol:g}: iters {st.solve.iterations} "
              f"outer {st.solve.outer_iterations} |dx| {ex:.2e} rel |dv| {ev:.2e} rel |dp| {ep:.2e}", flush=True)
    # the pressure error against the FP64 oracle falls with the solve
    # tolerance until the FP32 right-hand side (FP32 P2G) bounds it near 1e-7
    # (measured: 1.1e-6 at FP32 1e-6, 1.4e-7 mixed 1e-10); particle errors
    # sit at the FP32 floor of the transfers for every solve precision
    assert out[(1, 1e-10)][2] < 0.5 * out[(0, 1e-6)][2] and out[(1, 1e-10)][2] < 1e-6
    for k, (ex, ev, ep) in out.items():
        assert ex < 1e-5 and ev < 1e-5, (k, ex, ev)
