"""Gradient-check harness: the checkpointed GPU rollout wrapped in a
torch.autograd.Function (the only place torch autograd touches the library),
checked against the FP64 CPU adjoint and central finite differences of the
FP64 forward with the GPU's topology and liquid masks frozen.

Workload: sloshing pool on a 32^3-effective 2-level grid, 6 steps,
checkpoint every 2; loss |COM(x_K) - x*|^2 built in torch; gradients with
respect to the initial velocities and positions.  The torch tensors live on
the host: the Function moves them through the library's own device buffers.

Run as its own process (tests/test_gpu_autograd.py starts it): torch is
imported before the library loads, so the process holds a single HIP
runtime.  torch links its bundled HIP/HSA runtimes under names that do not
match the SONAMEs of the system ones, so importing torch after the library
has loaded would bring a second runtime stack into the process.
"""
import os
import sys

import torch  # noqa: F401  (first: see above)
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), os.path.join(os.path.dirname(HERE), "oracle"), HERE]

import anitowave_adaptive_amd as aw  # noqa: E402
import oracle as O  # noqa: E402
from anitowave_adaptive_amd import scenes  # noqa: E402
from helpers import DT, GRAV, rel_err  # noqa: E402

ALPHA = 0.95


def _rollout_function():
    import torch

    class FlipRollout(torch.autograd.Function):
        """x_K = rollout(x_0, v_0): forward through aw_rollout_forward, backward
        through aw_rollout_backward.  Tensors are (n, 3) float32 host tensors in
        particle-id order; device buffers belong to the library."""

        @staticmethod
        def forward(ctx, x0, v0, rollout, params):
            n = x0.shape[0]
            dev = rollout.grid.ctx
            pos = aw.DeviceArray.from_numpy(dev, x0.detach().numpy().astype(np.float32))
            vel = aw.DeviceArray.from_numpy(dev, v0.detach().numpy().astype(np.float32))
            ids = aw.DeviceArray.from_numpy(dev, np.arange(n, dtype=np.int32))
            rollout.forward(params, pos.ptr, vel.ptr, ids.ptr)
            p = pos.download(np.float32, 3 * n).reshape(-1, 3)
            order = ids.download(np.int32, n)
            xK = np.empty_like(p)
            xK[order] = p
            ctx.rollout = rollout
            return torch.from_numpy(xK)

        @staticmethod
        def backward(ctx, gxK):
            n = gxK.shape[0]
            dev = ctx.rollout.grid.ctx
            xbar = aw.DeviceArray.from_numpy(dev, gxK.detach().numpy().astype(np.float32))
            vbar = aw.DeviceArray.from_numpy(dev, np.zeros((n, 3), np.float32))
            ctx.rollout.backward(xbar.ptr, vbar.ptr)
            gx = torch.from_numpy(xbar.download(np.float32, 3 * n).reshape(-1, 3))
            gv = torch.from_numpy(vbar.download(np.float32, 3 * n).reshape(-1, 3))
            return gx, gv, None, None

    return FlipRollout


def autograd_rollout_gradient():
    n_cells, levels, band, steps = 32, 2, (0, 1, 0, 0), 6
    pos, vel = scenes.slosh(n=n_cells, ppc=2, seed=9)
    rng = np.random.default_rng(2)
    vel = (vel + 0.2 * rng.standard_normal(vel.shape)).astype(np.float32)
    n = len(pos)
    target = torch.tensor([0.6, 0.35, 0.5], dtype=torch.float64)
    print(f"[autograd] {n} particles, {steps} steps", flush=True)

    ctx = aw.Context(0)
    grid = aw.MSBGManager(ctx, (n_cells // 8,) * 3, 1.0 / n_cells, num_levels=levels, band=band)
    solver = aw.VCycleSolver(grid, precision=1)
    params = aw.FlipStep(grid, solver, dt=DT, gravity=GRAV, alpha=ALPHA, tol=1e-10, max_iters=400).params
    rollout = aw.Rollout(grid, solver, n, steps, checkpoint_every=2, keep_masks=True)

    x0 = torch.tensor(pos, requires_grad=True)
    v0 = torch.tensor(vel, requires_grad=True)
    xK = _rollout_function().apply(x0, v0, rollout, params)
    loss = ((xK.double().mean(dim=0) - target) ** 2).sum()
    loss.backward()
    gx = x0.grad.double().numpy()
    gv = v0.grad.double().numpy()
    st = rollout.stats
    print(f"[autograd] loss {loss.item():.6e}; forward {st.forward_ms:.1f} ms, recompute {st.recompute_ms:.1f} ms, "
          f"backward {st.backward_ms:.1f} ms", flush=True)

    # FP64 oracle reverse pass on the GPU's frozen topology and masks
    masks = [rollout.masks(k) for k in range(steps)]
    og = O.OracleGrid((n_cells // 8,) * 3, num_levels=levels, h=1.0 / n_cells, band=band)
    og.set_params()
    states, _ = O.rollout64(og, pos, vel, steps, DT, GRAV, ALPHA, tol=1e-12, frozen=masks)
    tgt = target.numpy()
    c = states[-1][0].mean(axis=0) - tgt
    xbar = np.broadcast_to(2.0 * c / n, (n, 3)).copy()
    ox, ov = O.rollout_grad64(og, states, masks, xbar, np.zeros_like(xbar), DT, GRAV, ALPHA, tol=1e-12)
    ev, ex = rel_err(gv, ov), rel_err(gx, ox)
    print(f"[autograd] vs FP64 adjoint: dL/dv0 {ev:.2e}, dL/dx0 {ex:.2e}", flush=True)
    assert ev < 1e-3 and ex < 1e-3

    def L64(x, v):
        xe = O.rollout64(og, x, v, steps, DT, GRAV, ALPHA, tol=1e-12, frozen=masks)[0][-1][0]
        d = xe.mean(axis=0) - tgt
        return float(d @ d)

    eps = 1e-3
    for k in range(4):
        r = np.random.default_rng(200 + k)
        dv = r.standard_normal(vel.shape)
        dv /= np.linalg.norm(dv)
        fd = (L64(pos, vel + eps * dv) - L64(pos, vel - eps * dv)) / (2 * eps)
        ad = float((gv * dv).sum())
        print(f"[autograd] direction {k}: fd {fd:.6e} adjoint {ad:.6e}", flush=True)
        assert abs(fd - ad) <= 1e-3 * abs(fd)


if __name__ == "__main__":
    autograd_rollout_gradient()
    print("[autograd] ok", flush=True)
