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
