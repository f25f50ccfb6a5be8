"""aw_remap_cells_noalloc (the entry point that allocates no device memory)
against aw_remap_cells through the C ABI: bit-identical on two calls in a
row, the second on a smaller grid (the shared scratch keeps the first call's
contents), no change in the library's device memory across the calls, and a
grid beyond the scratch's 2^18 blocks rejected with AW_ERR_INVALID."""
import ctypes

import numpy as np
import pytest

import anitowave_adaptive_amd as aw

pytestmark = pytest.mark.gpu


def _graded(rng, D, L):
    from scipy.ndimage import minimum_filter
    lv = rng.integers(0, L, (D[2], D[1], D[0]))
    while True:
        new = np.minimum(lv, minimum_filter(lv, size=3, mode="nearest") + 1)
        if np.array_equal(new, lv):
            return lv.reshape(-1).astype(np.uint8)
        lv = new


def _case(ctx, D, seed):
    rng = np.random.default_rng(seed)
    prev, cur = _graded(rng, D, 4), _graded(rng, D, 4)
    gg = aw.MSBGManager(ctx, D, 1.0 / 256, num_levels=4)
    gg.set_levels(prev)
    src = rng.standard_normal(int(gg.layout().total_cells)).astype(np.float32)
    dsrc = aw.DeviceArray.from_numpy(ctx, src)
    dlev = aw.DeviceArray.from_numpy(ctx, prev)
    gg.set_levels(cur)
    ncur = int(gg.layout().total_cells)
    want = aw.DeviceArray(ctx, 4 * ncur)
    got = aw.DeviceArray(ctx, 4 * ncur)
    return gg, dlev, dsrc, want, got, ncur


def test_noalloc_remap_matches_and_allocates_nothing(ctx):
    L = aw.lib()
    for D, seed in (((24, 20, 16), 1), ((6, 5, 4), 2)):
        print(f"[noalloc] grid {D} ({np.prod(D)} blocks)", flush=True)
        gg, dlev, dsrc, want, got, ncur = _case(ctx, D, seed)
        gg.remap_cells(dlev.ptr, dsrc.ptr, want.ptr)
        ctx.synchronize()
        before = aw.memory_stats()[0]
        aw.check(L.aw_remap_cells_noalloc(gg.handle, dlev.ptr, dsrc.ptr, got.ptr), "aw_remap_cells_noalloc")
        ctx.synchronize()
        assert aw.memory_stats()[0] == before
        np.testing.assert_array_equal(got.download(np.float32, ncur), want.download(np.float32, ncur))


def test_noalloc_rejects_a_grid_beyond_its_scratch(ctx):
    D = (65, 64, 64)   # 266,240 blocks > 2^18
    gg = aw.MSBGManager(ctx, D, 1.0 / (8 * 64), num_levels=4)
    gg.set_levels(np.full(int(np.prod(D)), 3, np.uint8))
    n = int(gg.layout().total_cells)
    dlev = aw.DeviceArray.from_numpy(ctx, np.full(int(np.prod(D)), 3, np.uint8))
    src = aw.DeviceArray.from_numpy(ctx, np.zeros(n, np.float32))
    dst = aw.DeviceArray(ctx, 4 * n)
    st = aw.lib().aw_remap_cells_noalloc(gg.handle, dlev.ptr, src.ptr, dst.ptr)
    assert st == aw.AW_ERR_INVALID and b"at most 262144" in aw.lib().aw_last_error()
