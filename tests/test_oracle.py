"""Pins of the FP64 oracle against closed forms, brute force and invariants.

None of these re-types the oracle's own formulas: layouts are checked against
the reference's index definitions, topology against a brute-force Python rule
(the particle-radius rule in exact rational arithmetic),
the operator against a hand-assembled 7-point Laplacian and against the
coarse-fine coefficients fixed by the geometry, transfers against partition of
unity / linear reproduction / mass conservation, the projection against the
hydrostatic column and against D u_new = residual.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from anitowave_adaptive_amd import scenes


def interleave(x, y, z):
    m = 0
    for i in range(10):
        m |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return m


def test_morton_known_values():
    assert O.morton3d(0, 0, 0) == 0
    assert O.morton3d(1, 0, 0) == 1
    assert O.morton3d(0, 1, 0) == 2
    assert O.morton3d(0, 0, 1) == 4
    assert O.morton3d(1, 1, 1) == 7
    assert O.morton3d(1023, 1023, 1023) == (1 << 30) - 1
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(200, 3)):
        assert O.morton3d(int(x), int(y), int(z)) == interleave(int(x), int(y), int(z))


def test_layout_offsets_closed_form():
    g = O.OracleGrid((3, 2, 2), num_levels=4, h=1.0 / 24)
    rng = np.random.default_rng(1)
    lv = np.zeros(g.nb, np.uint8)
    # a graded random map: levels from a smooth ramp along x
    for b in range(g.nb):
        x = b % 3
        lv[b] = min(3, x + rng.integers(0, 2)) if x < 2 else 3
    lv[[0, 3]] = 0
    lv = np.minimum(lv, 3)
    g.set_levels(lv)
    c, fx, fy, fz = g.offsets()
    r = 8 >> lv.astype(np.int64)
    np.testing.assert_array_equal(c, np.concatenate([[0], np.cumsum(r ** 3)]))
    # MSBG.cuh:118-120: (r+1) r^2 slots on every axis
    np.testing.assert_array_equal(fx, np.concatenate([[0], np.cumsum((r + 1) * r * r)]))
    np.testing.assert_array_equal(fx, fy)
    np.testing.assert_array_equal(fx, fz)
    t = g.totals()
    assert t[0] == c[-1] and t[1] == fx[-1]


def test_binning_matches_floor():
    g = O.OracleGrid((4, 4, 4), num_levels=1, h=1.0 / 32)
    rng = np.random.default_rng(2)
    pos = rng.random((5000, 3)).astype(np.float32)
    pos[:10] = np.float32(0.0)
    pos[10:20] = np.nextafter(np.float32(1.0), np.float32(0))
    key, cnt = g.bin(pos)
    gc = np.clip(np.floor(pos.astype(np.float64) * 32).astype(int), 0, 31) // 8
    np.testing.assert_array_equal(key, gc[:, 0] + 4 * (gc[:, 1] + 4 * gc[:, 2]))
    np.testing.assert_array_equal(cnt, np.bincount(key, minlength=64))


def brute_levels(D, L, band, count):
    occ = count.reshape(D[2], D[1], D[0]) > 0
    nb = occ.size
    coords = list(itertools.product(range(D[2]), range(D[1]), range(D[0])))
    surf = np.zeros_like(occ)
    for z, y, x in coords:
        if not occ[z, y, x]:
            continue
        for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
            q = (z + dz, y + dy, x + dx)
            if all(0 <= q[i] < (D[2], D[1], D[0])[i] for i in range(3)) and not occ[q]:
                surf[z, y, x] = True
    S = np.argwhere(surf)
    lv = np.full(occ.shape, L - 1, int)
    for z, y, x in coords:
        d = np.max(np.abs(S - [z, y, x]), axis=1).min() if len(S) else 10 ** 6
        for k in range(L - 1):
            if d <= band[k]:
                lv[z, y, x] = k
                break
    # grading to a fixed point
    while True:
        new = lv.copy()
        for z, y, x in coords:
            for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
                q = (z + dz, y + dy, x + dx)
                if all(0 <= q[i] < (D[2], D[1], D[0])[i] for i in range(3)):
                    new[z, y, x] = min(new[z, y, x], lv[q] + 1)
        if np.array_equal(new, lv):
            break
        lv = new
    return lv.reshape(nb).astype(np.uint8)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_levels_brute_force(seed):
    D = (6, 5, 4)
    rng = np.random.default_rng(seed)
    count = (rng.random(D[0] * D[1] * D[2]) < 0.35).astype(np.uint32) * rng.integers(1, 9, D[0] * D[1] * D[2]).astype(np.uint32)
    for L, band in [(2, (1, 0, 0, 0)), (3, (0, 2, 0, 0)), (4, (0, 1, 2, 0))]:
        g = O.OracleGrid(D, num_levels=L, h=1.0 / 64, band=band)
        got = g.levels_from_counts(count)
        np.testing.assert_array_equal(got, brute_levels(D, L, band, count))


def brute_grade(lv, D):
    """2:1 grading over the 26-neighbourhood, iterated until nothing changes."""
    coords = list(itertools.product(range(D[2]), range(D[1]), range(D[0])))
    while True:
        new = lv.copy()
        for z, y, x in coords:
            for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
                q = (z + dz, y + dy, x + dx)
                if all(0 <= q[i] < (D[2], D[1], D[0])[i] for i in range(3)):
                    new[z, y, x] = min(new[z, y, x], lv[q] + 1)
        if np.array_equal(new, lv):
            return lv
        lv = new


def brute_radius_levels(D, L, h, pos, radius):
    """The reference's particle-radius rule (MSBG.cuh:286-298) in exact
    rational arithmetic: plev = floor(max(0, log2(h / r))) as the largest k
    with 2^k <= h / r, support = ceil(r / res), block = floor(x / h) // 8
    clamped into the domain."""
    from fractions import Fraction
    import math
    lv = np.full((D[2], D[1], D[0]), L - 1, int)
    smax = max(D)
    H = Fraction(h)
    for p, r in zip(pos, radius):
        R = Fraction(float(r))
        k = 0
        while k < L - 1 and R * 2 ** (k + 1) <= H:
            k += 1
        res = 8 >> k
        sup = min(math.ceil(R / res), smax)
        c = [min(max(math.floor(Fraction(float(p[a])) / H), 0), 8 * D[a] - 1) // 8 for a in range(3)]
        for z in range(max(0, c[2] - sup), min(D[2] - 1, c[2] + sup) + 1):
            for y in range(max(0, c[1] - sup), min(D[1] - 1, c[1] + sup) + 1):
                for x in range(max(0, c[0] - sup), min(D[0] - 1, c[0] + sup) + 1):
                    lv[z, y, x] = min(lv[z, y, x], k)
    return brute_grade(lv, D).reshape(-1).astype(np.uint8)


def radius_cases(rng, m, h, D):
    """Positions over the grid (some outside it) and radii at and around the
    powers of two that bound each level, plus a few large enough for a
    support of two blocks."""
    ext = np.array([8 * d * h for d in D])
    pos = (rng.random((m, 3)) * 1.2 - 0.1) * ext
    k = rng.integers(-1, 5, m)
    f = rng.choice([1.0, 0.5, 0.75, 1.5, np.nextafter(1.0, 0.0), np.nextafter(1.0, 2.0), 0.999, 1.001], m)
    r = h * np.ldexp(f, -k)
    big = [9.0, 8.5, 16.0 + 1e-3]
    r[: min(m, 3)] = big[: min(m, 3)]
    return pos.astype(np.float32), r.astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_radius_levels_brute_force(seed):
    D = (6, 5, 4)
    rng = np.random.default_rng(seed)
    for L in (1, 2, 3, 4):
        g = O.OracleGrid(D, num_levels=L, h=1.0 / 64)
        for m in (0, 1, 7, 60):
            pos, r = radius_cases(rng, m, 1.0 / 64, D)
            got = g.levels_from_radii(pos, r)
            np.testing.assert_array_equal(got, brute_radius_levels(D, L, 1.0 / 64, pos, r), err_msg=f"L={L} m={m}")


def test_radius_levels_known_answers():
    """One particle in block (2, 3, 1) of a 6^3-block grid, 4 levels: radius
    h/4 marks its 27 blocks with level 2 and leaves level 3 elsewhere; radius
    h marks them with level 0, and grading then sets every block to
    min(3, max(0, d - 1)) with d its Chebyshev block distance to (2, 3, 1)."""
    h = 1.0 / 64
    D = (6, 6, 6)
    g = O.OracleGrid(D, num_levels=4, h=h)
    pos = np.array([[(2 * 8 + 3) * h, (3 * 8 + 5) * h, (1 * 8 + 7.5) * h]], np.float32)
    z, y, x = np.meshgrid(range(6), range(6), range(6), indexing="ij")
    d = np.maximum(np.maximum(np.abs(x - 2), np.abs(y - 3)), np.abs(z - 1)).reshape(-1)
    got = g.levels_from_radii(pos, np.array([h / 4], np.float32))
    np.testing.assert_array_equal(got, np.where(d <= 1, 2, 3))
    got = g.levels_from_radii(pos, np.array([h], np.float32))
    np.testing.assert_array_equal(got, np.minimum(3, np.maximum(0, d - 1)))
    # just above h / 4 the level is 1, just below h / 8 it is 3
    got = g.levels_from_radii(pos, np.array([np.nextafter(np.float32(h / 4), np.float32(1))], np.float32))
    np.testing.assert_array_equal(got, np.where(d <= 1, 1, np.minimum(3, d)))
    got = g.levels_from_radii(pos, np.array([np.nextafter(np.float32(h / 8), np.float32(0))], np.float32))
    np.testing.assert_array_equal(got, np.full(216, 3))


@pytest.mark.parametrize("bad", [0.0, -1.0, np.nan, np.inf])
def test_radius_levels_reject_bad_radius(bad):
    g = O.OracleGrid((4, 4, 4), num_levels=3, h=1.0 / 32)
    pos = np.full((2, 3), 0.5, np.float32)
    with pytest.raises(ValueError):
        g.levels_from_radii(pos, np.array([1.0 / 64, bad], np.float32))


def test_levels_pool_scene_graded():
    pos, _ = scenes.pool_sphere(n=64)
    g = O.OracleGrid((8, 8, 8), num_levels=3, h=1.0 / 64, band=(1, 2, 0, 0))
    lv = g.build_topology(pos).reshape(8, 8, 8)
    for ax in range(3):
        assert np.abs(np.diff(lv.astype(int), axis=ax)).max() <= 1


def dense_A(g, k=0):
    rp, col, val, dg = g.matrix(k)
    n = len(dg)
    A = sp.csr_matrix((val, col, rp), shape=(n, n)) + sp.diags(dg)
    liq = g.mg_get(k, "liquid") > 0
    return A, liq


def set_liquid(g, liq):
    g.set_cell("liquid", liq.astype(np.float64))


def test_operator_single_level_is_the_7_point_laplacian():
    # one level, 2x2x2 blocks (16^3 cells): hand-assembled 7-point Laplacian
    # with Neumann tank walls and Dirichlet air above y = 10
    g = O.OracleGrid((2, 2, 2), num_levels=1, h=1.0 / 16)
    g.set_levels(np.zeros(8, np.uint8))
    n = 16
    # native cell order: block-major, x fastest inside a block
    idx = np.zeros((n, n, n), int)
    k = 0
    for bz, by, bx in itertools.product(range(2), repeat=3):
        for z, y, x in itertools.product(range(8), repeat=3):
            idx[bz * 8 + z, by * 8 + y, bx * 8 + x] = k
            k += 1
    liq_grid = np.zeros((n, n, n), bool)
    liq_grid[:, :10, :] = True
    liq = np.zeros(n ** 3)
    liq[idx[liq_grid]] = 1
    set_liquid(g, liq)
    g.mg_build()
    A, L = dense_A(g)
    B = sp.lil_matrix((n ** 3, n ** 3))
    for z, y, x in itertools.product(range(n), repeat=3):
        if not liq_grid[z, y, x]:
            continue
        i = idx[z, y, x]
        for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            q = (z + d[0], y + d[1], x + d[2])
            if not all(0 <= c < n for c in q):
                continue  # tank wall: Neumann
            B[i, i] += 1.0
            if liq_grid[q]:
                B[i, idx[q]] -= 1.0
    Lm = np.asarray(L)
    diff = (A - B.tocsr())[Lm][:, Lm]
    assert abs(diff).max() == 0.0


def test_operator_coarse_fine_coefficients():
    # block 0 at level 0 (8^3 cells, s=1), block 1 at level 1 (4^3, s=2): every
    # coarse cell on the interface couples to 4 fine cells with A/d = 1/1.5
    g = O.OracleGrid((2, 1, 1), num_levels=2, h=1.0 / 16)
    g.set_levels(np.array([0, 1], np.uint8))
    nc = int(g.totals()[0])
    set_liquid(g, np.ones(nc))
    g.mg_build()
    A, liq = dense_A(g)
    A = A.toarray()
    assert np.allclose(A, A.T, atol=0, rtol=0)
    fine_x7 = [x + 8 * (y + 8 * z) for z in range(8) for y in range(8) for x in [7]]
    coarse_x0 = [512 + 4 * (y + 4 * z) for z in range(4) for y in range(4)]
    for c in coarse_x0:
        cpl = -A[c, :512]
        nz = np.nonzero(cpl)[0]
        assert len(nz) == 4 and set(nz) <= set(fine_x7)
        assert np.allclose(cpl[nz], 1.0 / 1.5)
    # equal-level couplings: 1 between fine cells, 2 between coarse cells
    assert A[0, 1] == -1.0
    assert A[512, 513] == -2.0
    # all-Neumann box: constants are in the kernel
    assert np.abs(A @ np.ones(nc)).max() < 1e-12


@pytest.mark.parametrize("name", ["dam", "pool"])
def test_operator_spd_and_multigrid_transpose(name):
    if name == "dam":
        pos, vel = scenes.dam_break(n=32)
        g = O.OracleGrid((4, 4, 4), num_levels=1, h=1.0 / 32)
    else:
        pos, vel = scenes.pool_sphere(n=32, radius=0.15)
        g = O.OracleGrid((4, 4, 4), num_levels=3, h=1.0 / 32, band=(0, 1, 0, 0))
    g.build_topology(pos)
    g.p2g(pos, vel)
    g.set_params()
    g.mg_build()
    for k in range(g.num_mg_levels()):
        A, liq = dense_A(g, k)
        A = A[liq][:, liq]
        if A.shape[0] == 0:
            continue
        assert abs(A - A.T).max() == 0.0
        if A.shape[0] < 3000:
            w = np.linalg.eigvalsh(A.toarray())
            assert w.min() > 0
    rng = np.random.default_rng(5)
    for k in range(g.num_mg_levels() - 1):
        nf = g.mg_info(k)["ncells"]
        ncoarse = g.mg_info(k + 1)["ncells"]
        r = rng.standard_normal(nf)
        e = rng.standard_normal(ncoarse)
        par = g.mg_get(k, "parent").astype(int)
        # restriction is the sum over children: <R r, e> = <r, P e> with P injection
        g.mg_set(k, "r", r)
        g.restrict(k)
        Rr = g.mg_get(k + 1, "b")
        np.testing.assert_allclose(Rr, np.bincount(par, weights=r, minlength=ncoarse), atol=1e-12)
        assert abs(Rr @ e - r @ e[par]) < 1e-9 * (1 + abs(r @ e[par]))


def test_dense_solve_single_level():
    pos, vel = scenes.dam_break(n=16)
    g = O.OracleGrid((2, 2, 2), num_levels=1, h=1.0 / 16)
    g.build_topology(pos)
    g.p2g(pos, vel, 1 / 240, (0, -9.81, 0))
    g.divergence()
    g.set_params()
    g.mg_build()
    st, it, rel = g.solve(tol=1e-12, max_iters=100)
    assert st == 0
    A, liq = dense_A(g)
    b = g.cell("rhs")
    x = np.zeros_like(b)
    x[liq] = np.linalg.solve(A[liq][:, liq].toarray(), b[liq])
    np.testing.assert_allclose(g.cell("p"), x, rtol=0, atol=1e-9 * np.abs(x).max())


def test_p2g_mass_and_uniform_velocity():
    # particles away from the +walls: every particle deposits weight 1 per axis
    rng = np.random.default_rng(3)
    n = 32
    pos = (0.1 + 0.75 * rng.random((20000, 3))).astype(np.float32)
    v0 = np.array([0.3, -1.25, 2.0], np.float32)
    vel = np.tile(v0, (len(pos), 1))
    for L, band in [(1, (0, 0, 0, 0)), (3, (0, 1, 0, 0))]:
        g = O.OracleGrid((4, 4, 4), num_levels=L, h=1.0 / n, band=band)
        lv = np.array([min(L - 1, (b % 4)) for b in range(64)], np.uint8) if L > 1 else np.zeros(64, np.uint8)
        g.set_levels(lv)
        g.p2g(pos, vel)
        for a in range(3):
            assert abs(g.face("macc", a).sum() - len(pos)) < 1e-8 * len(pos)
            u = g.face("u_old", a)
            m = g.face("mass", a)
            has = m > 1e-6
            assert has.sum() > 0
            np.testing.assert_allclose(u[has], v0[a], rtol=1e-12, atol=1e-12)


def _linear_face_field(g, coef):
    """u_a at native face centres for u_a(x) = c0 + c1 x + c2 y + c3 z (fine units)."""
    lv = g.levels()
    c, f, _, _ = g.offsets()
    D = g.dims
    out = [np.zeros(int(g.totals()[1 + a])) for a in range(3)]
    for b in range(g.nb):
        B = np.array([b % D[0], (b // D[0]) % D[1], b // (D[0] * D[1])])
        e = int(lv[b])
        s, r = 1 << e, 8 >> e
        for a in range(3):
            dims = [r, r, r]
            dims[a] = r + 1
            for z, y, x in itertools.product(range(dims[2]), range(dims[1]), range(dims[0])):
                sl = np.array([x, y, z], float)
                posf = 8 * B + s * (sl + 0.5)
                posf[a] = 8 * B[a] + s * sl[a]
                if a == 0:
                    k = x + (r + 1) * (y + r * z)
                elif a == 1:
                    k = x + r * (y + (r + 1) * z)
                else:
                    k = x + r * (y + r * z)
                out[a][f[b] + k] = coef[a][0] + coef[a][1:] @ posf
    return out


def test_g2p_linear_field_and_rk3_translation():
    n = 32
    g = O.OracleGrid((4, 4, 4), num_levels=2, h=1.0 / n, band=(0, 0, 0, 0))
    lv = np.zeros(64, np.uint8)
    lv[[b for b in range(64) if b % 4 == 3]] = 1     # one coarse slab at the +x side
    g.set_levels(lv)
    g.p2g(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    coef = [np.array([0.1, 0.0, 0.02, -0.03]), np.array([-0.2, 0.01, 0.0, 0.04]),
            np.array([0.3, 0.05, -0.01, 0.0])]
    fields = _linear_face_field(g, coef)
    for a in range(3):
        g.set_face("u_star", a, fields[a])
        g.set_face("u_old", a, np.zeros_like(fields[a]))
    g.set_cell("p", np.zeros(int(g.totals()[0])))
    g.project()   # p = 0: u_new = u_star except the tank-wall slots
    rng = np.random.default_rng(4)
    # sample points whose stencils stay inside level-0 blocks, off the walls
    pts = np.column_stack([rng.uniform(2.0, 21.0, 400), rng.uniform(2.0, 30.0, 400), rng.uniform(2.0, 30.0, 400)])
    v = g.sample(0, pts)
    for a in range(3):
        np.testing.assert_allclose(v[:, a], coef[a][0] + pts @ coef[a][1:], rtol=0, atol=1e-12)
    # inside the coarse blocks away from their faces Pi is trilinear too
    pts2 = np.column_stack([rng.uniform(26.5, 29.5, 200), rng.uniform(2.0, 30.0, 200), rng.uniform(2.0, 30.0, 200)])
    pts2[:, 1] = np.where(np.abs((pts2[:, 1] % 8) - 4) > 2.5, 4.0, pts2[:, 1])
    pts2[:, 2] = np.where(np.abs((pts2[:, 2] % 8) - 4) > 2.5, 4.0, pts2[:, 2])
    v2 = g.sample(0, pts2)
    for a in range(3):
        np.testing.assert_allclose(v2[:, a], coef[a][0] + pts2 @ coef[a][1:], rtol=0, atol=1e-12)
    # uniform field: RK3 moves particles by exactly dt v
    u0 = [0.5, -0.25, 0.75]
    for a in range(3):
        f = np.full_like(fields[a], u0[a])
        g.set_face("u_star", a, f)
        g.set_face("u_old", a, f)
    g.project()
    p0 = (np.column_stack([rng.uniform(0.3, 0.6, 100)] * 3)).astype(np.float32)
    dt = 1.0 / 240
    v0 = np.zeros_like(p0)
    po, vo = g.g2p_advect(p0, v0, 0.95, dt)
    np.testing.assert_allclose(po, p0 + dt * np.array(u0), atol=1e-12)
    # FLIP: u_new - u_old = 0, so v = 0.95 v_p + 0.05 u_new
    np.testing.assert_allclose(vo, 0.05 * np.array(u0)[None, :].repeat(100, 0), atol=1e-12)


def test_hydrostatic_column_and_divergence_equals_residual():
    n = 32
    pos, vel = scenes.seed_cells(lambda x, y, z: y < 0.5, n)
    g = O.OracleGrid((4, 4, 4), num_levels=1, h=1.0 / n)
    g.build_topology(pos)
    dt, gy = 1 / 240, -9.81
    g.p2g(pos, vel, dt, (0, gy, 0))
    g.divergence()
    g.set_params()
    g.mg_build()
    st, it, rel = g.solve(tol=1e-12, max_iters=200)
    assert st == 0
    p = g.cell("p")
    # native order of a single-level grid: block-major, x fastest
    P = np.zeros((n, n, n))
    k = 0
    for bz, by, bx in itertools.product(range(4), repeat=3):
        for z, y, x in itertools.product(range(8), repeat=3):
            P[bz * 8 + z, by * 8 + y, bx * 8 + x] = p[k]
            k += 1
    # liquid rows y = 0..15; pressure rises by dt |g| per cell with depth
    d = P[:, :14, :] - P[:, 1:15, :]
    np.testing.assert_allclose(d, dt * 9.81, rtol=1e-8)
    assert np.abs(P[:, 16:, :]).max() == 0.0
    g.project()
    # divergence of the projected field equals the solver residual
    b = g.cell("rhs")
    r = b - g.apply(0, p)
    for a in range(3):
        g.set_face("u_star", a, g.face("u_new", a))
    g.divergence()
    np.testing.assert_allclose(-g.cell("rhs"), -r, atol=1e-10)
    # at rest under gravity the projected velocity vanishes in the liquid
    uy = g.face("u_new", 1)
    assert np.abs(uy).max() < 1e-8
