"""Pins of the oracle's variable-density coefficients (beta = 1 / rho, the
channels the reference reserves, MSBG.cuh:21-34; its commented beta deposit,
P2G.cuh:93-97) and of their coarsening through the multigrid hierarchy (the
reference's galerkinCoarsen averages the four fine faces under each coarse
face, MSBG.cuh:401-498).

* Stratified column at rest (heavy layer under a light one, free surface on
  top): the pressure rises by rho_i dt |g| per cell inside each layer and the
  projected velocity vanishes (closed form).
* Constant density rho: every level's operator is the unit-density operator
  divided by rho (linearity of the area average).
* Piecewise-constant beta (c1 for x < X0, c2 for x >= X0, X0 on a face of
  the coarsest cells), set slot by slot with the coarse side of every level
  change holding garbage that must not be read: every coupling on every
  level is c(face) times the unit operator's coupling.
* Random beta: every level's operator is symmetric and positive definite, and
  after the solve and the projection the divergence of u_new equals the
  solver residual (the projection uses the operator's beta).
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from anitowave_adaptive_amd import scenes

DT, G = 1.0 / 240.0, 9.81


def dense_A(g, k):
    rp, col, val, dg = g.matrix(k)
    n = len(dg)
    return sp.csr_matrix((val, col, rp), shape=(n, n)) + sp.diags(dg)


def cell_boxes(g, k):
    """Origin (fine units, x y z) and edge of every cell of MG level k."""
    info = g.mg_info(k)
    D, bs = info["dims"], info["bs"]
    e, coff = g.mg_e(k)
    org = np.zeros((info["ncells"], 3), int)
    edge = np.zeros(info["ncells"], int)
    for b in range(info["nb"]):
        Bc = np.array([b % D[0], (b // D[0]) % D[1], b // (D[0] * D[1])])
        s = 1 << int(e[b])
        r = bs // s
        for li in range(r ** 3):
            l = np.array([li % r, (li // r) % r, li // (r * r)])
            org[coff[b] + li] = Bc * bs + l * s
            edge[coff[b] + li] = s
    return org, edge


def lattice_particles(n, top):
    """2 x 2 x 2 particles per fine cell at the quarter points, in every cell
    below the fine row `top`: the same sub-cell positions everywhere, so each
    row of faces gets the same deposit (a rest state exists; with jittered
    particles the interface faces' densities vary across the columns)."""
    q = (np.arange(2) + 0.5) / 2
    cells = np.stack(np.meshgrid(np.arange(n), np.arange(top), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    sub = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    pos = ((cells[:, None, :] + sub[None, :, :]).reshape(-1, 3) / n).astype(np.float32)
    return pos, np.zeros_like(pos)


def test_stratified_column_at_rest():
    n = 32
    pos, vel = lattice_particles(n, 16)
    rho = np.where(pos[:, 1] < 0.25, 2.0, 0.5)
    g = O.OracleGrid((4, 4, 4), num_levels=1, h=1.0 / n)
    g.build_topology(pos)
    g.p2g(pos, vel, DT, (0, -G, 0), density=rho)
    g.divergence()
    g.set_params()
    g.mg_build()
    st, it, rel = g.solve(tol=1e-12, max_iters=300)
    assert st == 0
    p = g.cell("p")
    P = np.zeros((n, n, n))
    k = 0
    for bz, by, bx in itertools.product(range(4), repeat=3):
        for z, y, x in itertools.product(range(8), repeat=3):
            P[bz * 8 + z, by * 8 + y, bx * 8 + x] = p[k]
            k += 1
    d = P[:, :-1, :] - P[:, 1:, :]   # p(y) - p(y + 1), across the face y + 1
    # faces 1..7 lie in the heavy layer (cells y < 8), faces 9..15 in the light one
    np.testing.assert_allclose(d[:, 0:7, :], 2.0 * DT * G, rtol=1e-8)
    np.testing.assert_allclose(d[:, 8:15, :], 0.5 * DT * G, rtol=1e-8)
    # the interface face: the mean density of its deposit (half its weight from each layer)
    np.testing.assert_allclose(d[:, 7, :], 1.25 * DT * G, rtol=1e-8)
    g.project()
    for a in range(3):
        assert np.abs(g.face("u_new", a)).max() < 1e-9


def _level_grid(seed, liquid_all=True):
    rng = np.random.default_rng(seed)
    g = O.OracleGrid((4, 4, 4), num_levels=3, h=1.0 / 32)
    from scipy.ndimage import minimum_filter
    lv = rng.integers(0, 3, (4, 4, 4))
    while True:
        new = np.minimum(lv, minimum_filter(lv, size=3, mode="nearest") + 1)
        if np.array_equal(new, lv):
            break
        lv = new
    g.set_levels(lv.reshape(-1).astype(np.uint8))
    nc = int(g.totals()[0])
    liq = np.ones(nc) if liquid_all else (rng.random(nc) < 0.9).astype(np.float64)
    g.set_cell("liquid", liq)
    return g, rng


def test_constant_density_scales_every_level():
    pos, vel = scenes.pool_sphere(n=32, radius=0.15)
    mats = []
    for rho in (None, 4.0):
        g = O.OracleGrid((4, 4, 4), num_levels=3, h=1.0 / 32, band=(0, 1, 0, 0))
        g.build_topology(pos)
        g.p2g(pos, vel, density=None if rho is None else np.full(len(pos), rho))
        g.set_params()
        g.mg_build()
        mats.append([dense_A(g, k) for k in range(g.num_mg_levels())])
    assert len(mats[0]) >= 4
    compared = 0
    for A1, A4 in zip(*mats):
        if A1.nnz == 0 or abs(A1).max() == 0.0:   # a coarse level without liquid cells
            assert A4.nnz == 0 or abs(A4).max() == 0.0
            continue
        assert abs(A4 * 4.0 - A1).max() < 1e-12 * abs(A1).max()
        compared += 1
    assert compared >= 3


def _face_slots(g):
    """Per axis: for every native slot its plane coordinate along the axis
    (fine units) and whether it is the coarse side of a level change."""
    lv = g.levels()
    _, f, _, _ = g.offsets()
    D = g.dims
    out = []
    for a in range(3):
        plane = np.zeros(int(g.totals()[1 + a]), int)
        coarse_side = np.zeros(len(plane), bool)
        xcen = np.zeros(len(plane))
        for b in range(g.nb):
            Bc = [b % D[0], (b // D[0]) % D[1], b // (D[0] * D[1])]
            e = int(lv[b])
            s, r = 1 << e, 8 >> e
            dims = [r, r, r]
            dims[a] = r + 1
            for z, y, x in itertools.product(range(dims[2]), range(dims[1]), range(dims[0])):
                sl = (x, y, z)
                k = f[b] + x + dims[0] * (y + dims[1] * z)
                plane[k] = 8 * Bc[a] + sl[a] * s
                xcen[k] = 8 * Bc[0] + (sl[0] + (0 if a == 0 else 0.5)) * s
                if sl[a] in (0, r):
                    nb = list(Bc)
                    nb[a] += -1 if sl[a] == 0 else 1
                    if 0 <= nb[a] < D[a] and lv[nb[0] + D[0] * (nb[1] + D[1] * nb[2])] < e:
                        coarse_side[k] = True
        out.append((plane, coarse_side, xcen))
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_piecewise_constant_beta_through_the_hierarchy(seed):
    X0, c1, c2 = 16, 0.7, 3.0
    mats = []
    for variable in (False, True):
        g, _ = _level_grid(seed)
        if variable:
            for a, (plane, coarse_side, xcen) in enumerate(_face_slots(g)):
                bt = g.face("beta", a)
                wall = bt == 0.0
                val = np.where(xcen < X0, c1, c2)
                val[coarse_side] = 1e3       # not an unknown: never read
                val[wall] = 0.0
                g.set_face("beta", a, val)
        g.set_params()
        g.mg_build()
        mats.append((g, [dense_A(g, k).tocoo() for k in range(g.num_mg_levels())]))
    (g1, A1s), (_, Avs) = mats
    assert g1.num_mg_levels() >= 4
    for k, (A1, Av) in enumerate(zip(A1s, Avs)):
        org, edge = cell_boxes(g1, k)
        assert org[:, 0].max() < 32 and not np.any((org[:, 0] < X0) & (org[:, 0] + edge > X0))
        A1 = A1.tocsr()
        Av = Av.tocsr()
        off = A1.tocoo()
        i, j, v = off.row, off.col, off.data
        m = i != j
        i, j, v = i[m], j[m], v[m]
        # the coupling's face: normal x when the cells touch along x
        touch_x = (org[i, 0] + edge[i] == org[j, 0]) | (org[j, 0] + edge[j] == org[i, 0])
        px = np.where(org[i, 0] + edge[i] == org[j, 0], org[j, 0], org[i, 0])
        c = np.where(touch_x, np.where(px < X0, c1, c2), np.where(org[i, 0] < X0, c1, c2))
        got = np.asarray(Av[i, j]).reshape(-1)
        np.testing.assert_allclose(got, c * v, rtol=1e-13, err_msg=f"level {k}")
        # all-liquid Neumann box: the diagonal is minus the row sum of couplings
        dg = Av.diagonal()
        rows = -np.asarray((Av - sp.diags(dg)).sum(axis=1)).reshape(-1)
        np.testing.assert_allclose(dg, rows, rtol=1e-13, atol=1e-12)


def test_random_density_spd_and_projection_consistent():
    pos, vel = scenes.pool_sphere(n=32, radius=0.15)
    rng = np.random.default_rng(8)
    rho = rng.uniform(0.5, 3.0, len(pos))
    g = O.OracleGrid((4, 4, 4), num_levels=3, h=1.0 / 32, band=(0, 1, 0, 0))
    g.build_topology(pos)
    g.p2g(pos, vel, DT, (0, -G, 0), density=rho)
    bt = np.concatenate([g.face("beta", a) for a in range(3)])
    assert bt.min() >= 0.0 and bt[bt > 0].min() > 1 / 3.0 - 1e-12 and bt.max() <= 2.0 + 1e-12
    g.divergence()
    g.set_params()
    g.mg_build()
    for k in range(g.num_mg_levels()):
        A = dense_A(g, k)
        liq = g.mg_get(k, "liquid") > 0
        A = A[liq][:, liq]
        if A.shape[0] == 0:
            continue
        assert abs(A - A.T).max() <= 1e-14 * abs(A).max()
        if A.shape[0] < 3000:
            assert np.linalg.eigvalsh(A.toarray()).min() > 0
    st, it, rel = g.solve(tol=1e-12, max_iters=300)
    assert st == 0
    p = g.cell("p")
    b = g.cell("rhs")
    r = b - g.apply(0, p)
    g.project()
    for a in range(3):
        g.set_face("u_star", a, g.face("u_new", a))
    g.divergence()
    np.testing.assert_allclose(-g.cell("rhs"), -r, atol=1e-10 * np.abs(b).max())
