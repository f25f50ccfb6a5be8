"""The particle-radius level rule (aw_grid_build_topology_radius; the
reference's markActiveBlocks + enforceGrading, MSBG.cuh:271-351, with grading
to a fixed point) against the oracle, bit-exact: native levels, cell and face
offsets, per-block counts.  Radii sit at and around the powers of two that
bound each level (fine in a ball, coarser around it, the top of the grid
empty); three are large enough for a support of two or three blocks; some
particles lie outside the grid (binned into the nearest block).  The
particles arrive in generation order, block-sorted with one radius (the
kernel's lanes then repeat their neighbour's marks and skip them), and
shuffled.  A bad radius must raise and leave the topology as it was.  A full
step on a radius topology (rebuild_topology = 0) matches the oracle's FP64
step with the same frozen levels.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
import oracle as O
from anitowave_adaptive_amd import scenes
from helpers import DT, GRAV, GpuParticles, step_checked

pytestmark = pytest.mark.gpu

CASES = [
    # dims in blocks, native levels, particles
    ((6, 5, 4), 4, 300),
    ((32, 32, 32), 4, 200000),
    ((40, 24, 36), 3, 150000),
]


def _radius_cloud(dims, h, m, seed):
    """Particles in the lower 60% of the grid (some just outside it), fine
    radii in a ball, coarser ones around it, the top empty: every level
    occurs on the larger grids."""
    rng = np.random.default_rng(seed)
    ext = np.array([8 * d * h for d in dims])
    pos = (rng.random((m, 3)) * np.array([1.1, 0.65, 1.1]) - 0.05) * ext
    ball = np.linalg.norm((pos - ext * np.array([0.5, 0.3, 0.5])) / ext.min(), axis=1) < 0.12
    k = np.where(ball, rng.choice([-1, 0], m), rng.choice([1, 2, 3], m, p=[0.01, 0.2, 0.79]))
    f = rng.choice([1.0, 0.5, 0.75, 1.5, np.nextafter(1.0, 0.0), np.nextafter(1.0, 2.0)], m)
    r = h * np.ldexp(f, -k)
    if m > 1000:  # on the tiny grid these would cover every block
        r[:3] = [9.0, 8.5, 16.001]
    return pos.astype(np.float32), r.astype(np.float32)


class _Radius:
    def __init__(self, ctx, r):
        self.d = aw.DeviceArray.from_numpy(ctx, np.ascontiguousarray(r, np.float32))


def _check(gg, og, pos, r):
    lv = og.levels_from_radii(pos, r)
    og.set_levels(lv)
    np.testing.assert_array_equal(gg.levels(), lv)
    c, f = gg.offsets()
    oc, ofx, _, _ = og.offsets()
    np.testing.assert_array_equal(c, oc)
    np.testing.assert_array_equal(f, ofx)
    _, cnt = og.bin(pos)
    np.testing.assert_array_equal(gg.particle_counts(), cnt)
    return lv


@pytest.mark.parametrize("dims,levels,m", CASES)
@pytest.mark.parametrize("order", ["generated", "sorted_one_radius", "shuffled"])
def test_radius_topology_bit_exact(ctx, dims, levels, m, order):
    h = 1.0 / 256
    pos, r = _radius_cloud(dims, h, m, seed=sum(dims) + m)
    og = O.OracleGrid(dims, num_levels=levels, h=h)
    if order == "sorted_one_radius":
        key, _ = og.bin(pos)
        pos = pos[np.argsort(key, kind="stable")]
        r = np.full(len(pos), np.float32(h / 2))
    elif order == "shuffled":
        perm = np.random.default_rng(5).permutation(len(pos))
        pos, r = pos[perm], r[perm]
    print(f"[radius] dims {dims}, {levels} levels, {len(pos)} particles, {order}", flush=True)
    gg = aw.MSBGManager(ctx, dims, h, num_levels=levels)
    gp = GpuParticles(ctx, pos, np.zeros_like(pos))
    dr = _Radius(ctx, r)
    assert gg.build_topology_radius(gp.pos.ptr, dr.d.ptr, gp.n)
    lv = _check(gg, og, pos, r)
    if order != "sorted_one_radius" and m > 1000:
        assert len(np.unique(lv)) == levels, "every native level must occur"
    # the same call again: nothing changes
    assert not gg.build_topology_radius(gp.pos.ptr, dr.d.ptr, gp.n)


@pytest.mark.parametrize("bad", [0.0, -2.0, np.nan, np.inf])
def test_bad_radius_leaves_topology(ctx, bad):
    dims, levels, h = (8, 8, 8), 3, 1.0 / 64
    pos, r = _radius_cloud(dims, h, 5000, seed=3)
    og = O.OracleGrid(dims, num_levels=levels, h=h)
    gg = aw.MSBGManager(ctx, dims, h, num_levels=levels)
    gp = GpuParticles(ctx, pos, np.zeros_like(pos))
    good = _Radius(ctx, r)
    gg.build_topology_radius(gp.pos.ptr, good.d.ptr, gp.n)
    before = gg.levels()
    r2 = r.copy()
    r2[4321] = bad
    badr = _Radius(ctx, r2)
    print(f"[radius] bad radius {bad}", flush=True)
    with pytest.raises(aw.AWError):
        gg.build_topology_radius(gp.pos.ptr, badr.d.ptr, gp.n)
    np.testing.assert_array_equal(gg.levels(), before)
    _check(gg, og, pos, r)


def test_step_on_radius_topology(ctx):
    """64^3 pool with a falling sphere: pool particles of radius h/2 (level 1
    around them), sphere particles of radius h (level 0), air at level 2."""
    n, L = 64, 3
    h = 1.0 / n
    pos, vel = scenes.pool_sphere(n=n, seed=1234)
    r = np.where(pos[:, 1] < 0.3, h / 2, h).astype(np.float32)
    og = O.OracleGrid((n // 8,) * 3, num_levels=L, h=h)
    og.set_params()
    lv = og.levels_from_radii(pos, r)
    assert set(np.unique(lv)) == {0, 1, 2}
    ref = og.step64(pos, vel, DT, GRAV, 0.95, tol=1e-10, levels=lv)
    assert ref["status"] == 0
    gg = aw.MSBGManager(ctx, (n // 8,) * 3, h, num_levels=L)
    sv = aw.VCycleSolver(gg)
    gp = GpuParticles(ctx, pos, vel)
    dr = _Radius(ctx, r)
    gg.build_topology_radius(gp.pos.ptr, dr.d.ptr, gp.n)
    np.testing.assert_array_equal(gg.levels(), lv)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6, rebuild_topology=False)
    st = step_checked(step, gp)
    print(f"[radius] step iters gpu {st.solve.iterations} oracle {ref['iters']}", flush=True)
    assert not st.topology_changed
    np.testing.assert_array_equal(gg.levels(), lv)
    P, V = gp.unpermuted()
    assert np.abs(P - ref["pos"]).max() < 1e-5
    vmax = max(np.abs(ref["vel"]).max(), 1e-6)
    assert np.abs(V - ref["vel"]).max() / vmax < 1e-3


def test_steps_rebuilding_with_radius_rule(ctx):
    """rebuild_topology = 2 re-marks the levels from the radii every step,
    finding each particle's radius by its id as the step reorders the
    particles: the levels equal the oracle's rule on the radii in the
    current order, and the rollout equals building the radius topology
    before each step and stepping with the topology kept (particles within
    1e-6 absolute: P2G's LDS float adds are not order-deterministic).  A
    step without ids is rejected."""
    n, L = 64, 3
    h = 1.0 / n
    pos, vel = scenes.pool_sphere(n=n, seed=1234, vy=-8.0)
    r = np.where(pos[:, 1] < 0.3, h / 2, h).astype(np.float32)
    dr = _Radius(ctx, r)
    og = O.OracleGrid((n // 8,) * 3, num_levels=L, h=h)
    runs = []
    for mode in (2, 0):
        gg = aw.MSBGManager(ctx, (n // 8,) * 3, h, num_levels=L)
        sv = aw.VCycleSolver(gg)
        gp = GpuParticles(ctx, pos, vel)
        step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6, rebuild_topology=mode,
                           d_radius=dr.d.ptr if mode == 2 else None)
        runs.append((gg, gp, step))
    (ga, pa, sa), (gb, pb, sb) = runs
    changes = []
    for k in range(6):
        # the oracle sees exactly the input run A's rebuild-2 step sees
        ida = pa.ids.download(np.int32, pa.n)
        pcur = pa.pos.download(np.float32, 3 * pa.n).reshape(-1, 3)
        ids = pb.ids.download(np.int32, pb.n)
        rslot = _Radius(ctx, r[ids])
        gb.build_topology_radius(pb.pos.ptr, rslot.d.ptr, pb.n)
        step_checked(sb, pb)
        st = step_checked(sa, pa)
        changes.append(st.topology_changed)
        np.testing.assert_array_equal(ga.levels(), og.levels_from_radii(pcur, r[ida]), err_msg=f"step {k}")
        np.testing.assert_array_equal(ga.levels(), gb.levels())
    print(f"[radius] rebuild 2: topology changes {changes}", flush=True)
    PA, VA = pa.unpermuted()
    PB, VB = pb.unpermuted()
    # the two runs differ only in the order of float reductions (the P2G
    # deposit's LDS adds, the solver's dot products), which each PCG stop at
    # 1e-6 of |b| carries into the velocities: after 6 steps the positions
    # (world size 1) agree to a few 1e-6, not to the last bit
    assert np.abs(PA - PB).max() < 1e-5
    assert np.abs(VA - VB).max() / max(np.abs(VB).max(), 1e-6) < 1e-4
    with pytest.raises(aw.AWError):
        sa(pa.pos.ptr, pa.vel.ptr, None, pa.n)


def test_failed_radius_rebuild_invalidates_the_adjoint_state(ctx):
    """A rebuild_topology = 2 step that fails on a bad radius rebins the
    particles before it fails: the previous step's sorted state no longer
    matches the per-block counts, so the step adjoint must refuse to run
    (AW_ERR_INVALID) instead of reading the stale particle ranges."""
    n, L = 32, 3
    h = 1.0 / n
    pos, vel = scenes.pool_sphere(n=n, seed=7, radius=0.15)
    r = np.full(len(pos), h, np.float32)
    gg = aw.MSBGManager(ctx, (n // 8,) * 3, h, num_levels=L)
    sv = aw.VCycleSolver(gg)
    gp = GpuParticles(ctx, pos, vel)
    good = _Radius(ctx, r)
    step = aw.FlipStep(gg, sv, dt=DT, gravity=GRAV, alpha=0.95, tol=1e-6, rebuild_topology=2, d_radius=good.d.ptr)
    step_checked(step, gp)
    adj = aw.Adjoint(gg, sv)
    xb = aw.DeviceArray.from_numpy(ctx, np.zeros((gp.n, 3), np.float32))
    vb = aw.DeviceArray.from_numpy(ctx, np.zeros((gp.n, 3), np.float32))
    adj.step(step.params, xb.ptr, vb.ptr)   # valid after a good step
    bad_r = r.copy()
    bad_r[len(r) // 2] = np.nan
    bad = _Radius(ctx, bad_r)
    step.params.d_radius = bad.d.ptr
    with pytest.raises(aw.AWError) as e:
        step(gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n)
    assert e.value.status == aw.AW_ERR_INVALID
    with pytest.raises(aw.AWError) as e:
        adj.step(step.params, xb.ptr, vb.ptr)
    assert e.value.status == aw.AW_ERR_INVALID
