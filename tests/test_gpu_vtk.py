"""The uniform resample kernel (aw_resample_cells; sampleMSBG,
vtx_visualization.cuh:8-62) against the oracle, bit-exact, on the pressure a
real step leaves (the oracle resamples the same FP32 values): output sizes
equal to the fine grid, finer, coarser (centres on cell faces), odd, and a
single voxel.  The exporters' files read back to the resampled values and to
the grid's levels.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
from helpers import DT, GRAV, GpuParticles, gpu_grid, oracle_grid, particles, step_checked
from test_vtk import _read_vtk

pytestmark = pytest.mark.gpu

DIMS = [(64, 64, 64), (128, 96, 80), (32, 32, 32), (37, 11, 53), (1, 1, 1)]


@pytest.mark.parametrize("name", ["pool64", "double64"])
def test_resample_bit_exact(ctx, name, tmp_path):
    pos, vel = particles(name)
    gg = gpu_grid(ctx, name)
    sv = aw.VCycleSolver(gg)
    gp = GpuParticles(ctx, pos, vel)
    step_checked(aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6), gp)
    p = gg.cell("p")
    assert np.any(p)
    og = oracle_grid(name)
    og.set_levels(gg.levels())
    lay = gg.layout()
    for dims in DIMS:
        print(f"[vtk] {name} resample to {dims}", flush=True)
        n = int(np.prod(dims))
        out = aw.DeviceArray(ctx, 4 * n)
        gg.resample(lay.d_pressure, dims, out.ptr)
        got = out.download(np.float32, n)
        want = og.resample(p.astype(np.float64), dims).astype(np.float32)
        np.testing.assert_array_equal(got, want)
    f = tmp_path / "p.vtk"
    gg.export_cells_vtk(lay.d_pressure, (37, 11, 53), f)
    _, sec = _read_vtk(f)
    np.testing.assert_array_equal(sec["SCALARS pressure"], og.resample(p.astype(np.float64), (37, 11, 53)))
    sp = np.array(sec["SPACING"])
    np.testing.assert_allclose(sp, [1 / 37, 1 / 11, 1 / 53], rtol=1e-6)
    np.testing.assert_allclose(sec["ORIGIN"], sp / 2, rtol=1e-6)
    b = tmp_path / "b.vtk"
    gg.export_blocks_vtk(b)
    _, sec = _read_vtk(b)
    np.testing.assert_array_equal(sec["SCALARS refineLevel"], gg.levels())


def test_resample_rejects_bad_sizes(ctx):
    gg = gpu_grid(ctx, "dam32")
    pos, vel = particles("dam32")
    gp = GpuParticles(ctx, pos, vel)
    gg.build_topology(gp.pos.ptr, gp.n)
    out = aw.DeviceArray(ctx, 64)
    for dims in [(0, 4, 4), (4, 70000, 4), (4, 4, -1)]:
        with pytest.raises(aw.AWError):
            gg.resample(gg.layout().d_pressure, dims, out.ptr)
