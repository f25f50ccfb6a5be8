"""Topology and binning on grids larger than one scan chunk (4096 blocks) and
with wide distance bands, against the oracle, bit-exact: native levels, cell
and face offsets, per-block particle counts.  Binning is checked on particles
in generation order, on the block-sorted output of a previous binning, on a
random permutation, and on input whose work-groups of consecutive particles
span more blocks than the per-work-group aggregation table holds.

The oracle's band rule scans the whole (2 bmax + 1)^3 block window around each
block; the HIP path computes the same Chebyshev distance one axis at a time.
"""
import numpy as np
import pytest

import anitowave_adaptive_amd as aw
import oracle as O
from helpers import GpuParticles

pytestmark = pytest.mark.gpu

CASES = [
    # dims in blocks, native levels, bands in blocks
    ((32, 32, 32), 4, (2, 4, 7, 0)),
    ((40, 24, 36), 3, (3, 5, 0, 0)),
    ((72, 8, 16), 2, (2, 0, 0, 0)),
]


def _cloud(dims, h, seed):
    """A slab along the floor, a ball and a sparse spray: surface blocks at
    many distances, empty blocks between."""
    rng = np.random.default_rng(seed)
    ext = np.array(dims, np.float64) * 8 * h
    slab = rng.random((120000, 3)) * ext * np.array([1.0, 0.3, 1.0])
    ball = rng.standard_normal((60000, 3))
    ball = ext * np.array([0.6, 0.6, 0.4]) + 0.12 * ext.min() * ball / np.maximum(
        np.linalg.norm(ball, axis=1, keepdims=True), 1e-9) * rng.random((60000, 1)) ** (1 / 3)
    spray = rng.random((12, 3)) * ext
    pos = np.concatenate([slab, ball, spray]).astype(np.float32)
    return np.clip(pos, 0.0, (ext - 1e-4 * h).astype(np.float32))


@pytest.mark.parametrize("dims,levels,band", CASES)
def test_wide_band_topology_bit_exact(ctx, dims, levels, band):
    h = 1.0 / 256
    pos = _cloud(dims, h, seed=sum(dims))
    print(f"[topology] dims {dims}, {levels} levels, bands {band}, {len(pos)} particles", flush=True)
    og = O.OracleGrid(dims, num_levels=levels, h=h, band=band)
    og.build_topology(pos)
    gg = aw.MSBGManager(ctx, dims, h, num_levels=levels, band=band)
    gp = GpuParticles(ctx, pos, np.zeros_like(pos))
    gg.build_topology(gp.pos.ptr, gp.n)
    lv = gg.levels()
    np.testing.assert_array_equal(lv, og.levels())
    assert len(np.unique(lv)) == levels, "every native level must occur"
    c, f = gg.offsets()
    oc, ofx, _, _ = og.offsets()
    np.testing.assert_array_equal(c, oc)
    np.testing.assert_array_equal(f, ofx)
    _, cnt = og.bin(pos)
    np.testing.assert_array_equal(gg.particle_counts(), cnt)


@pytest.mark.parametrize("order", ["generated", "sorted", "shuffled", "strided"])
def test_binning_orders(ctx, order):
    dims, levels, band = CASES[0]
    h = 1.0 / 256
    pos = _cloud(dims, h, seed=7)
    rng = np.random.default_rng(3)
    og = O.OracleGrid(dims, num_levels=levels, h=h, band=band)
    if order == "sorted":
        key, _ = og.bin(pos)
        pos = pos[np.argsort(key, kind="stable")]
    elif order == "shuffled":
        pos = pos[rng.permutation(len(pos))]
    elif order == "strided":
        # consecutive particles in different blocks: every work-group of 2048
        # particles sees far more distinct blocks than its 512 table slots
        key, _ = og.bin(pos)
        srt = pos[np.argsort(key, kind="stable")]
        pos = srt[np.argsort(np.arange(len(srt)) % 997, kind="stable")]
    print(f"[topology] binning, {order} input, {len(pos)} particles", flush=True)
    vel = rng.standard_normal(pos.shape).astype(np.float32)
    gg = aw.MSBGManager(ctx, dims, h, num_levels=levels, band=band)
    gp = GpuParticles(ctx, pos, vel)
    gg.build_topology(gp.pos.ptr, gp.n)
    for rep in range(2):   # the second pass bins the block-sorted output of the first
        aw.P2GTransfer.bin(gg, gp.pos.ptr, gp.vel.ptr, gp.ids.ptr, gp.n)
        p, v, ids = gp.get()
        np.testing.assert_array_equal(np.sort(ids), np.arange(gp.n))
        np.testing.assert_array_equal(p, pos[ids])
        np.testing.assert_array_equal(v, vel[ids])
        key, cnt = og.bin(p)
        assert np.all(np.diff(key) >= 0), ("particles not grouped by block", rep)
        np.testing.assert_array_equal(gg.particle_counts(), cnt)
