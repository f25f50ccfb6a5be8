"""Synthetic scenes: counter-hashed jittered particle seeding on a unit cube.

Every particle is keyed by (seed, fine-cell linear index, k) through SplitMix64,
so the same particles come out whatever the chunking, and FP32 positions are
rounded once from FP64: (cell + U) * h.
"""
import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    with np.errstate(over="ignore"):
        z = z + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def _unit(z):
    return (z >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def seed_cells(mask_fn, n, ppc=8, seed=1234, vel_fn=None):
    """Seed `ppc` particles in every fine cell (h = 1/n) whose centre satisfies
    mask_fn(x, y, z) (arrays of cell centres).  Returns (pos, vel) float32 (N,3)."""
    h = 1.0 / n
    pos_chunks, vel_chunks = [], []
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")  # y, x
    ii = ii.ravel().astype(np.int64)
    jj = jj.ravel().astype(np.int64)
    for k in range(n):
        cx = (ii + 0.5) * h
        cy = (jj + 0.5) * h
        cz = np.full_like(cx, (k + 0.5) * h)
        m = mask_fn(cx, cy, cz)
        if not np.any(m):
            continue
        xi, yi = ii[m], jj[m]
        lin = (xi + n * (yi + n * k)).astype(np.uint64)
        base = lin[:, None] * np.uint64(ppc) + np.arange(ppc, dtype=np.uint64)[None, :]
        key = _mix(np.uint64(seed) ^ _mix(base))
        g = np.stack([np.broadcast_to(xi[:, None], key.shape), np.broadcast_to(yi[:, None], key.shape),
                      np.full(key.shape, k, np.int64)], axis=-1).astype(np.float64)
        u = np.stack([_unit(_mix(key + np.uint64(a + 1))) for a in range(3)], axis=-1)
        p = ((g + u) * h).astype(np.float32).reshape(-1, 3)
        pos_chunks.append(p)
        if vel_fn is not None:
            vel_chunks.append(vel_fn(p.astype(np.float64)).astype(np.float32).reshape(-1, 3))
        else:
            vel_chunks.append(np.zeros_like(p))
    if not pos_chunks:
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)
    return np.concatenate(pos_chunks), np.concatenate(vel_chunks)


def dam_break(n=64, ppc=8, seed=1234):
    """Dam-break column x < 0.5, y < 0.5 (all z), at rest."""
    return seed_cells(lambda x, y, z: (x < 0.5) & (y < 0.5), n, ppc, seed)


def pool_sphere(n=256, ppc=8, seed=1234, centre=(0.5, 0.7, 0.5), radius=0.1, vy=-2.0, depth=0.3):
    """Pool y < depth plus a sphere falling at vy."""
    c = np.array(centre)

    def mask(x, y, z):
        return (y < depth) | ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < radius * radius)

    def vel(p):
        v = np.zeros_like(p)
        ins = ((p - c) ** 2).sum(-1) < radius * radius
        ins &= p[:, 1] >= depth
        v[ins, 1] = vy
        return v

    return seed_cells(mask, n, ppc, seed, vel)


def slosh(n=128, ppc=8, seed=1234, depth=0.4, amp=0.05):
    """Pool y < depth + amp sin(2 pi (x + p1)) sin(2 pi (z + p2)), phases from the seed."""
    ph = _unit(_mix(np.array([seed, seed + 1], np.uint64)))

    def mask(x, y, z):
        return y < depth + amp * np.sin(2 * np.pi * (x + ph[0])) * np.sin(2 * np.pi * (z + ph[1]))

    return seed_cells(mask, n, ppc, seed)


def double_dam(n=512, ppc=8, seed=1234):
    """Two columns [0,.3]x[0,.6]x[0,.3] and [.7,1]x[0,.6]x[.7,1]."""
    def mask(x, y, z):
        a = (x < 0.3) & (y < 0.6) & (z < 0.3)
        b = (x > 0.7) & (y < 0.6) & (z > 0.7)
        return a | b

    return seed_cells(mask, n, ppc, seed)


# Workloads (grid dims in 8^3 blocks, native levels, bands in blocks).
WORKLOADS = {
    "dam64": dict(n=64, dims=(8, 8, 8), levels=1, band=(0, 0, 0, 0), scene=dam_break, steps=50),
    "pool256": dict(n=256, dims=(32, 32, 32), levels=3, band=(1, 4, 0, 0), scene=pool_sphere, steps=100),
    "slosh128": dict(n=128, dims=(16, 16, 16), levels=2, band=(1, 0, 0, 0), scene=slosh, steps=16),
    "double512": dict(n=512, dims=(64, 64, 64), levels=4, band=(1, 4, 16, 0), scene=double_dam, steps=8),
}
