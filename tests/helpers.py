"""Shared set-up for parity tests: identical particles on the oracle and the HIP path."""
import numpy as np

import anitowave_adaptive_amd as aw
from anitowave_adaptive_amd import scenes
import oracle as O

SCENES = {
    # name: (fine cells per axis, native levels, bands in blocks, scene fn)
    "dam32": (32, 1, (0, 0, 0, 0), scenes.dam_break),
    "pool64": (64, 3, (1, 2, 0, 0), scenes.pool_sphere),
    "slosh32": (32, 2, (1, 0, 0, 0), scenes.slosh),
    "double64": (64, 4, (0, 1, 1, 0), scenes.double_dam),  # all four native levels, r = 8, 4, 2, 1
}

DT = 1.0 / 240.0
GRAV = (0.0, -9.81, 0.0)


def particles(name, seed=1234):
    n, L, band, fn = SCENES[name]
    pos, vel = fn(n=n, seed=seed)
    return pos, vel


def oracle_grid(name):
    n, L, band, _ = SCENES[name]
    g = O.OracleGrid((n // 8,) * 3, num_levels=L, h=1.0 / n, band=band)
    g.set_params()
    return g


def gpu_grid(ctx, name):
    n, L, band, _ = SCENES[name]
    return aw.MSBGManager(ctx, (n // 8,) * 3, 1.0 / n, num_levels=L, band=band)


class GpuParticles:
    def __init__(self, ctx, pos, vel):
        self.n = len(pos)
        self.pos = aw.DeviceArray.from_numpy(ctx, np.ascontiguousarray(pos, np.float32))
        self.vel = aw.DeviceArray.from_numpy(ctx, np.ascontiguousarray(vel, np.float32))
        self.ids = aw.DeviceArray.from_numpy(ctx, np.arange(self.n, dtype=np.int32))

    def get(self):
        p = self.pos.download(np.float32, 3 * self.n).reshape(-1, 3)
        v = self.vel.download(np.float32, 3 * self.n).reshape(-1, 3)
        i = self.ids.download(np.int32, self.n)
        return p, v, i

    def unpermuted(self):
        p, v, i = self.get()
        P = np.empty_like(p)
        V = np.empty_like(v)
        P[i] = p
        V[i] = v
        return P, V


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / den


# Bound on the true relative residual an FP32 pressure reaches on these scenes
# (rounding p to FP32 alone leaves about eps32 |A| |p| / |b|).
FP32_FLOOR = 5e-5


def check_solve_contract(status, st, tol, precision=0):
    """The solver's convergence contract: AW_OK only with the true relative
    residual at or below tol; an FP32 solve may stop at its floor instead,
    with AW_ERR_NOT_CONVERGED, its recurrence residual at or below tol and
    the true one above tol but below FP32_FLOOR."""
    if status == aw.AW_OK:
        assert st.rel_residual <= tol, (st.rel_residual, tol)
        return
    assert status == aw.AW_ERR_NOT_CONVERGED, status
    assert precision == 0, "a mixed-precision solve must reach its tolerance"
    assert st.rel_residual > tol, (st.rel_residual, tol)
    assert st.recurrence_rel_residual <= tol, st.recurrence_rel_residual
    assert st.rel_residual < FP32_FLOOR, st.rel_residual


def step_checked(stepper, gp, precision=0):
    """One aw_step on GpuParticles under the convergence contract."""
    st = stepper(gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n, allow_not_converged=True)
    check_solve_contract(stepper.status, st.solve, stepper.params.rel_tol, precision)
    return st
