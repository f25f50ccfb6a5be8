"""Pins of the oracle's adaptive smoothing (the reference's AdaptiveSmoother:
smoothActiveBlocks, activateBlock and solveLevel, Smoother.cuh:97-291; the
early exit decided per block, A12; the threshold relative to max |b|, A22).

* With theta = 0 and one sweep per block it is the plain block red-black
  smoother with one local sweep.
* A dense numpy restatement from the assembled matrix (the class-ordered
  updates pinned in test_oracle_pins.py, restricted to one block) with the
  rule as Smoother.cuh:105-202 states it: sweeps until the block's largest
  |r| falls below 0.33 times the first sweep's, activation of the block and
  its six face neighbours when the last sweep stays above theta; the oracle's
  x, per-block sweep counts and next-active flags equal it.
* As the smoother of stationary V-cycles (method 1, coarse-correction scale
  0.75) it converges to 1e-8 and agrees with the PCG solution.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from anitowave_adaptive_amd import scenes

OMEGA = float(np.float32(0.85714))


def grid(n=32, levels=3, band=(0, 1, 0, 0)):
    pos, vel = scenes.pool_sphere(n=n, radius=0.15)
    g = O.OracleGrid((n // 8,) * 3, num_levels=levels, h=1.0 / n, band=band)
    g.build_topology(pos)
    g.p2g(pos, vel, 1 / 240, (0, -9.81, 0))
    g.divergence()
    g.set_params()
    g.mg_build()
    return g


def level_geometry(g, k):
    info = g.mg_info(k)
    D, bs = info["dims"], info["bs"]
    e, coff = g.mg_e(k)
    blk = np.zeros(info["ncells"], int)
    cc = np.zeros(info["ncells"], int)
    for b in range(info["nb"]):
        r = bs >> int(e[b])
        for li in range(r ** 3):
            blk[coff[b] + li] = b
            cc[coff[b] + li] = (li % r + (li // r) % r + li // (r * r)) & 1
    Bc = np.array([[b % D[0], (b // D[0]) % D[1], b // (D[0] * D[1])] for b in range(info["nb"])])
    return D, blk, cc, Bc


def numpy_adaptive(g, k, x, b, iters, reverse, theta_rel, max_sweeps, red):
    rp, col, val, dg = g.matrix(k)
    n = len(dg)
    A = (sp.csr_matrix((val, col, rp), shape=(n, n)) + sp.diags(dg)).tocsr()
    liq = g.mg_get(k, "liquid") > 0
    D, blk, cc, Bc = level_geometry(g, k)
    nb = len(Bc)
    bcol = Bc.sum(1) & 1
    cells = [np.nonzero(blk == q)[0] for q in range(nb)]
    x = x.copy()
    act = np.ones(nb, bool)
    sweeps = np.zeros(nb, int)
    theta = theta_rel * np.abs(b[liq]).max()
    order = (1, 0) if reverse else (0, 1)
    for _ in range(iters):
        nxt = np.zeros(nb, bool)
        for bcl in order:
            for q in np.nonzero(bcol == bcl)[0]:
                if not act[q]:
                    sweeps[q] = 0
                    continue
                first = last = 0.0
                for ls in range(max_sweeps):
                    m = 0.0
                    for c in order:
                        S = cells[q][(cc[cells[q]] == c) & liq[cells[q]] & (dg[cells[q]] > 0)]
                        r = b[S] - A[S] @ x
                        x[S] += OMEGA * r / dg[S]
                        if len(S):
                            m = max(m, np.abs(r).max())
                    sweeps[q] = ls + 1
                    if ls == 0:
                        first = m
                    last = m
                    if first > 1e-9 and m < first * red:
                        break
                if last > theta:
                    nxt[q] = True
                    for a in range(3):
                        for sd in (-1, 1):
                            nc = Bc[q].copy()
                            nc[a] += sd
                            if 0 <= nc[a] < D[a]:
                                nxt[nc[0] + D[0] * (nc[1] + D[1] * nc[2])] = True
        act = nxt
    return x, sweeps, act


@pytest.mark.parametrize("k", [0, 1])
def test_theta_zero_one_sweep_is_the_plain_smoother(k):
    g = grid()
    rng = np.random.default_rng(1)
    n = g.mg_info(k)["ncells"]
    x0, b = rng.standard_normal(n), rng.standard_normal(n)
    g.set_params(local_sweeps=1)
    out = []
    for adaptive in (False, True):
        g.set_adaptive(adaptive, theta=0.0, max_block_sweeps=1)
        g.mg_set(k, "x", x0)
        g.mg_set(k, "b", b)
        g.smooth(k, 3, False)
        out.append(g.mg_get(k, "x"))
    np.testing.assert_allclose(out[1], out[0], rtol=0, atol=1e-13 * np.abs(out[0]).max())


@pytest.mark.parametrize("k,reverse,theta", [(0, False, 0.05), (0, True, 0.02), (1, False, 0.03), (2, False, 0.02)])
def test_matches_the_dense_restatement(k, reverse, theta):
    g = grid()
    g.set_adaptive(True, theta=theta, max_block_sweeps=10, block_reduction=0.33)
    rng = np.random.default_rng(10 + k)
    n = g.mg_info(k)["ncells"]
    # a smooth start and a rough right-hand side: blocks need different numbers of sweeps
    x0 = rng.standard_normal(n) * 0.1
    b = rng.standard_normal(n) * np.where(rng.random(n) < 0.3, 5.0, 1.0)
    g.mg_set(k, "x", x0)
    g.mg_set(k, "b", b)
    g.smooth(k, 2, reverse)
    x, sweeps, act = numpy_adaptive(g, k, x0, b, 2, reverse, theta, 10, 0.33)
    sw, ac = g.mg_adaptive(k)
    np.testing.assert_array_equal(sw, sweeps)
    np.testing.assert_array_equal(ac > 0, act)
    np.testing.assert_allclose(g.mg_get(k, "x"), x, rtol=0, atol=1e-11 * np.abs(x).max())
    # the decisions are not trivial: early exits, full runs, inactive blocks
    liq_blocks = sw > 0
    assert len(np.unique(sw[liq_blocks])) >= 2 and (ac == 0).any() and (ac > 0).any()


def test_stationary_vcycles_converge():
    g = grid()
    g.set_params(prolong_scale=0.75)
    n = g.mg_info(0)["ncells"]
    g.set_cell("p", np.zeros(n))
    st, it_pcg, _ = g.solve(tol=1e-10, max_iters=200, method=0)
    assert st == 0
    p_ref = g.cell("p")
    g.set_adaptive(True)
    g.set_cell("p", np.zeros(n))
    st, it, rel = g.solve(tol=1e-8, max_iters=200, method=1)
    assert st == 0 and rel <= 1e-8, (st, it, rel)
    liq = g.cell("liquid") > 0
    err = np.abs(g.cell("p") - p_ref)[liq].max() / np.abs(p_ref[liq]).max()
    print(f"adaptive stationary V-cycles: {it} cycles to 1e-8 (PCG {it_pcg} to 1e-10), p error {err:.1e}")
    assert err < 1e-5
