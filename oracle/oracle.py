"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the FP64 CPU oracle.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libaw_oracle.so")
_lib = None

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_I32 = ctypes.POINTER(ctypes.c_int32)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I64 = ctypes.POINTER(ctypes.c_int64)
_V = ctypes.c_void_p


def build():
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "og_create": (_V, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I]),
            "og_destroy": (None, [_V]),
            "og_morton3d": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
            "og_bin": (ctypes.c_int, [_V, _F, ctypes.c_int64, _I32, _U32]),
            "og_levels_from_counts": (ctypes.c_int, [_V, _U32, _U8]),
            "og_levels_from_radii": (ctypes.c_int, [_V, _F, _F, ctypes.c_int64, _U8]),
            "og_set_levels": (ctypes.c_int, [_V, _U8]),
            "og_resample": (ctypes.c_int, [_V, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]),
            "og_remap_cells": (ctypes.c_int, [_V, _U8, _D, _D]),
            "og_get_levels": (ctypes.c_int, [_V, _U8]),
            "og_get_offsets": (ctypes.c_int, [_V, _U32, _U32, _U32, _U32]),
            "og_totals": (ctypes.c_int, [_V, _I64]),
            "og_p2g": (ctypes.c_int, [_V, _F, _F, ctypes.c_int64] + [ctypes.c_double] * 4 + [_D]),
            "og_divergence": (ctypes.c_int, [_V]),
            "og_set_params": (ctypes.c_int, [_V, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
            "og_mg_build": (ctypes.c_int, [_V]),
            "og_set_adaptive": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
            "og_mg_adaptive": (ctypes.c_int, [_V, ctypes.c_int, _U8, _U8]),
            "og_num_mg_levels": (ctypes.c_int, [_V]),
            "og_mg_level_info": (ctypes.c_int, [_V, ctypes.c_int, _I]),
            "og_mg_level_e": (ctypes.c_int, [_V, ctypes.c_int, _U8, _U32]),
            "og_mg_get": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, _D]),
            "og_mg_set": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, _D]),
            "og_smooth": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "og_residual": (ctypes.c_int, [_V, ctypes.c_int]),
            "og_apply": (ctypes.c_int, [_V, ctypes.c_int, _D, _D]),
            "og_restrict": (ctypes.c_int, [_V, ctypes.c_int]),
            "og_prolong": (ctypes.c_int, [_V, ctypes.c_int]),
            "og_vcycle": (ctypes.c_int, [_V, ctypes.c_int]),
            "og_export_matrix": (ctypes.c_int, [_V, ctypes.c_int, _I, _I, _D, _D]),
            "og_solve": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_double, ctypes.c_int, _I, _D]),
            "og_project": (ctypes.c_int, [_V]),
            "og_g2p_advect": (ctypes.c_int, [_V, _F, _F, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _D, _D]),
            "og_sample": (ctypes.c_int, [_V, ctypes.c_int, _D, ctypes.c_int64, _D]),
            "og_get_face": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, _D]),
            "og_set_face": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, _D]),
            "og_get_cell": (ctypes.c_int, [_V, ctypes.c_int, _D]),
            "og_set_cell": (ctypes.c_int, [_V, ctypes.c_int, _D]),
            "og_step": (ctypes.c_int, [_V, _F, _F, ctypes.c_int64] + [ctypes.c_double] * 5
                        + [ctypes.c_int, ctypes.c_double, ctypes.c_int, _D, _D, _I, _D, _D, _D]),
            "og_num_threads": (ctypes.c_int, []),
            "og_set_threads": (ctypes.c_int, [ctypes.c_int]),
            "og_get_fine": (ctypes.c_int, [_V, ctypes.c_int, _D]),
            "og_set_fine": (ctypes.c_int, [_V, ctypes.c_int, _D]),
            "og_pi": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int]),
            "og_pi_transpose": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int]),
            "og_step64": (ctypes.c_int, [_V, _D, _D, ctypes.c_int64] + [ctypes.c_double] * 5
                          + [ctypes.c_int, ctypes.c_double, ctypes.c_int, _U8, _U8, _U8, _U8, _D, _D, _I, _D]),
            "og_adj_g2p": (ctypes.c_int, [_V, ctypes.c_int, _V, _V, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                          _D, _D, _D, _D]),
            "og_adj_project": (ctypes.c_int, [_V, ctypes.c_double, ctypes.c_int, _I, _D]),
            "og_adj_p2g": (ctypes.c_int, [_V, ctypes.c_int, _V, _V, ctypes.c_int64, _D, _D]),
            "og_step_adjoint": (ctypes.c_int, [_V, ctypes.c_int, _V, _V, ctypes.c_int64, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double, ctypes.c_int, _D, _D, _I, _D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


FACE_FIELDS = {"Macc": 0, "macc": 1, "u_old": 2, "u_star": 3, "u_new": 4, "mass": 5, "beta": 6,
               "ubar_new": 7, "ubar_old": 8, "ubar_star": 9, "Mbar": 10, "mbar": 11}
CELL_FIELDS = {"p": 0, "rhs": 1, "liquid": 2, "pcount": 3, "pbar": 4, "rhsbar": 5}
FINE_FIELDS = {"new": 0, "delta": 1, "bar_nd": 2, "bar_d": 3, "Mbar": 4, "mbar": 5, "M": 6, "m": 7}


def _particles(pos, vel):
    """Contiguous (n, 3) particle arrays and the oracle's precision flag."""
    fp64 = pos.dtype == np.float64
    t = np.float64 if fp64 else np.float32
    return np.ascontiguousarray(pos, t).reshape(-1, 3), np.ascontiguousarray(vel, t).reshape(-1, 3), int(fp64)
MG_FIELDS = {"x": 0, "b": 1, "r": 2, "liquid": 3, "diag": 4, "parent": 5}


class OracleGrid:
    """FP64 MSBG grid: dims in blocks of 8^3 cells, `num_levels` native levels."""

    def __init__(self, dims, num_levels=1, h=None, band=(1, 4, 16, 0)):
        self.dims = tuple(int(d) for d in dims)
        self.h = float(h) if h is not None else 1.0 / (8 * self.dims[0])
        b = (ctypes.c_int * 4)(*[int(x) for x in (list(band) + [0, 0, 0, 0])[:4]])
        self._h = lib().og_create(*self.dims, int(num_levels), self.h, b)
        if not self._h:
            raise ValueError("invalid oracle grid")
        self.nb = self.dims[0] * self.dims[1] * self.dims[2]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().og_destroy(self._h)
            self._h = None

    # topology
    def bin(self, pos):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        key = np.zeros(len(pos), np.int32)
        cnt = np.zeros(self.nb, np.uint32)
        lib().og_bin(self._h, _p(pos, _F), len(pos), _p(key, _I32), _p(cnt, _U32))
        return key, cnt

    def levels_from_counts(self, cnt):
        cnt = np.ascontiguousarray(cnt, np.uint32)
        out = np.zeros(self.nb, np.uint8)
        lib().og_levels_from_counts(self._h, _p(cnt, _U32), _p(out, _U8))
        return out

    def levels_from_radii(self, pos, radius):
        pos = np.ascontiguousarray(pos, np.float32)
        radius = np.ascontiguousarray(radius, np.float32)
        out = np.zeros(self.nb, np.uint8)
        if lib().og_levels_from_radii(self._h, _p(pos, _F), _p(radius, _F), len(radius), _p(out, _U8)) != 0:
            raise ValueError("particle radius not positive and finite")
        return out

    def build_topology_radius(self, pos, radius):
        lv = self.levels_from_radii(pos, radius)
        self.set_levels(lv)
        return lv

    def resample(self, field, out_dims):
        """Nearest-cell resample of a cell field onto out_dims voxel centres (x fastest)."""
        field = np.ascontiguousarray(field, np.float64)
        assert len(field) == int(self.totals()[0])
        out = np.zeros(int(np.prod(out_dims)), np.float64)
        if lib().og_resample(self._h, _p(field, _D), *[int(v) for v in out_dims], _p(out, _D)) != 0:
            raise ValueError("bad output size")
        return out

    def remap_cells(self, level_prev, src):
        """Cell field of the layout with native levels level_prev onto the current layout (each cell
        takes the previous cell containing its centre; a centre on a face takes the cell above)."""
        lp = np.ascontiguousarray(level_prev, np.uint8)
        src = np.ascontiguousarray(src, np.float64)
        r = 8 >> lp.astype(np.int64)
        assert len(src) == int((r ** 3).sum())
        out = np.zeros(int(self.totals()[0]), np.float64)
        if lib().og_remap_cells(self._h, _p(lp, _U8), _p(src, _D), _p(out, _D)) != 0:
            raise ValueError("level out of range")
        return out

    def set_levels(self, lv):
        lv = np.ascontiguousarray(lv, np.uint8)
        if lib().og_set_levels(self._h, _p(lv, _U8)) != 0:
            raise ValueError("level out of range")

    def build_topology(self, pos):
        _, cnt = self.bin(pos)
        lv = self.levels_from_counts(cnt)
        self.set_levels(lv)
        return lv

    def levels(self):
        out = np.zeros(self.nb, np.uint8)
        lib().og_get_levels(self._h, _p(out, _U8))
        return out

    def offsets(self):
        o = [np.zeros(self.nb + 1, np.uint32) for _ in range(4)]
        lib().og_get_offsets(self._h, *[_p(x, _U32) for x in o])
        return o

    def totals(self):
        t = np.zeros(4, np.int64)
        lib().og_totals(self._h, _p(t, _I64))
        return t

    # transfers
    def p2g(self, pos, vel, dt=0.0, gravity=(0.0, 0.0, 0.0), density=None):
        """density: relative density per particle (beta = 1 / rho on faces with mass), or None (beta = 1)."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        dens = None if density is None else np.ascontiguousarray(density, np.float64).reshape(len(pos))
        lib().og_p2g(self._h, _p(pos, _F), _p(vel, _F), len(pos), float(dt), *[float(g) for g in gravity],
                     None if dens is None else _p(dens, _D))

    def divergence(self):
        lib().og_divergence(self._h)

    def project(self):
        lib().og_project(self._h)

    def g2p_advect(self, pos, vel, alpha, dt):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        po = np.zeros(pos.shape, np.float64)
        vo = np.zeros(vel.shape, np.float64)
        lib().og_g2p_advect(self._h, _p(pos, _F), _p(vel, _F), len(pos), float(alpha), float(dt), _p(po, _D), _p(vo, _D))
        return po, vo

    def sample(self, field, units):
        units = np.ascontiguousarray(units, np.float64).reshape(-1, 3)
        out = np.zeros(units.shape, np.float64)
        if lib().og_sample(self._h, int(field), _p(units, _D), len(units), _p(out, _D)) != 0:
            raise RuntimeError("fine lattice not built")
        return out

    def face(self, name, axis):
        n = int(self.totals()[1 + axis])
        out = np.zeros(n, np.float64)
        lib().og_get_face(self._h, FACE_FIELDS[name], axis, _p(out, _D))
        return out

    def set_face(self, name, axis, vals):
        vals = np.ascontiguousarray(vals, np.float64)
        lib().og_set_face(self._h, FACE_FIELDS[name], axis, _p(vals, _D))

    def cell(self, name):
        out = np.zeros(int(self.totals()[0]), np.float64)
        lib().og_get_cell(self._h, CELL_FIELDS[name], _p(out, _D))
        return out

    def set_cell(self, name, vals):
        vals = np.ascontiguousarray(vals, np.float64)
        lib().og_set_cell(self._h, CELL_FIELDS[name], _p(vals, _D))

    # solver
    def set_params(self, omega=float(np.float32(0.85714)), local_sweeps=3, pre=2, post=2, coarse_sweeps=16,
                   prolong_scale=1.0):
        lib().og_set_params(self._h, float(omega), local_sweeps, pre, post, coarse_sweeps, float(prolong_scale))

    def set_adaptive(self, on=True, theta=1e-4, max_block_sweeps=10, block_reduction=0.33):
        """The reference's adaptive smoothing on every smoothing call except the coarsest level's sweeps."""
        if lib().og_set_adaptive(self._h, int(bool(on)), float(theta), int(max_block_sweeps),
                                 float(block_reduction)) != 0:
            raise ValueError("bad adaptive smoothing parameters")

    def mg_adaptive(self, k):
        """(sweeps run per block, flags left active) of the last adaptive smoothing of level k."""
        nb = self.mg_info(k)["nb"]
        sw = np.zeros(nb, np.uint8)
        act = np.zeros(nb, np.uint8)
        if lib().og_mg_adaptive(self._h, int(k), _p(sw, _U8), _p(act, _U8)) != 0:
            raise ValueError("no adaptive smoothing on this level")
        return sw, act

    def mg_build(self):
        lib().og_mg_build(self._h)

    def num_mg_levels(self):
        return lib().og_num_mg_levels(self._h)

    def mg_info(self, k):
        info = np.zeros(7, np.int32)
        lib().og_mg_level_info(self._h, k, _p(info, _I))
        return dict(dims=tuple(info[:3]), nb=int(info[3]), ncells=int(info[4]), bs=int(info[5]), nnz=int(info[6]))

    def mg_e(self, k):
        inf = self.mg_info(k)
        e = np.zeros(inf["nb"], np.uint8)
        c = np.zeros(inf["nb"] + 1, np.uint32)
        lib().og_mg_level_e(self._h, k, _p(e, _U8), _p(c, _U32))
        return e, c

    def mg_get(self, k, name):
        out = np.zeros(self.mg_info(k)["ncells"], np.float64)
        lib().og_mg_get(self._h, k, MG_FIELDS[name], _p(out, _D))
        return out

    def mg_set(self, k, name, vals):
        vals = np.ascontiguousarray(vals, np.float64)
        if lib().og_mg_set(self._h, k, MG_FIELDS[name], _p(vals, _D)) != 0:
            raise ValueError(name)

    def smooth(self, k, iters=1, reverse=False):
        lib().og_smooth(self._h, k, iters, int(reverse))

    def residual(self, k):
        lib().og_residual(self._h, k)

    def apply(self, k, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        lib().og_apply(self._h, k, _p(x, _D), _p(y, _D))
        return y

    def restrict(self, k):
        lib().og_restrict(self._h, k)

    def prolong(self, k):
        lib().og_prolong(self._h, k)

    def vcycle(self, k=0):
        lib().og_vcycle(self._h, k)

    def matrix(self, k):
        """Level-k operator as a dense-able CSR: (rowptr, col, val, diag)."""
        inf = self.mg_info(k)
        rp = np.zeros(inf["ncells"] + 1, np.int32)
        col = np.zeros(max(inf["nnz"], 1), np.int32)
        val = np.zeros(max(inf["nnz"], 1), np.float64)
        dg = np.zeros(inf["ncells"], np.float64)
        lib().og_export_matrix(self._h, k, _p(rp, _I), _p(col, _I), _p(val, _D), _p(dg, _D))
        return rp, col[: inf["nnz"]], val[: inf["nnz"]], dg

    def solve(self, tol=1e-6, max_iters=200, method=0):
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        st = lib().og_solve(self._h, int(method), float(tol), int(max_iters), ctypes.byref(it), ctypes.byref(rel))
        return st, it.value, rel.value

    def step(self, pos, vel, dt, gravity, alpha, method=0, tol=1e-6, max_iters=200, density=None):
        """One step; density: relative density per particle (input order) or None."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        dens = None if density is None else np.ascontiguousarray(density, np.float64).reshape(len(pos))
        po = np.zeros(pos.shape, np.float64)
        vo = np.zeros(vel.shape, np.float64)
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        ms = np.zeros(4, np.float64)
        st = lib().og_step(self._h, _p(pos, _F), _p(vel, _F), len(pos), float(dt), *[float(g) for g in gravity],
                           float(alpha), int(method), float(tol), int(max_iters), _p(po, _D), _p(vo, _D),
                           ctypes.byref(it), ctypes.byref(rel), _p(ms, _D), None if dens is None else _p(dens, _D))
        return dict(status=st, pos=po, vel=vo, iters=it.value, relres=rel.value, stage_ms=ms)


    # FP64 rollouts and the adjoint
    def step64(self, pos, vel, dt, gravity, alpha, tol=1e-10, max_iters=400, levels=None, liquid=None, method=0):
        """One FP64 step on FP64 particles; `levels` / `liquid` freeze the topology and the liquid mask."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        po = np.zeros(pos.shape, np.float64)
        vo = np.zeros(vel.shape, np.float64)
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        lv_in = None if levels is None else np.ascontiguousarray(levels, np.uint8)
        lq_in = None if liquid is None else np.ascontiguousarray(liquid, np.uint8)
        lv_out = np.zeros(self.nb, np.uint8)
        # the liquid mask size follows the topology: size it after the step from the totals
        lq_buf = np.zeros(self.nb * 512, np.uint8)
        st = lib().og_step64(self._h, _p(pos, _D), _p(vel, _D), len(pos), float(dt), *[float(g) for g in gravity],
                             float(alpha), int(method), float(tol), int(max_iters),
                             None if lv_in is None else _p(lv_in, _U8), None if lq_in is None else _p(lq_in, _U8),
                             _p(lv_out, _U8), _p(lq_buf, _U8), _p(po, _D), _p(vo, _D), ctypes.byref(it),
                             ctypes.byref(rel))
        nc = int(self.totals()[0])
        return dict(status=st, pos=po, vel=vo, iters=it.value, relres=rel.value, levels=lv_out,
                    liquid=lq_buf[:nc].copy())

    def adj_g2p(self, pos, vel, xbar_out, vbar_out, alpha, dt):
        """Adjoint of G2P + RK3: particle partials (xbar, vbar); fills ubar_new, ubar_old."""
        pos, vel, f = _particles(pos, vel)
        xo = np.ascontiguousarray(xbar_out, np.float64).reshape(-1, 3)
        vo = np.ascontiguousarray(vbar_out, np.float64).reshape(-1, 3)
        xi = np.zeros_like(xo)
        vi = np.zeros_like(vo)
        if lib().og_adj_g2p(self._h, f, pos.ctypes.data, vel.ctypes.data, len(pos), float(alpha), float(dt),
                            _p(xo, _D), _p(vo, _D), _p(xi, _D), _p(vi, _D)) != 0:
            raise RuntimeError("no forward fine lattice")
        return xi, vi

    def adj_project(self, tol=1e-12, max_iters=400):
        """Adjoint of the projection: ubar_new -> pbar -> A^{-1} -> rhsbar -> ubar_star."""
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        st = lib().og_adj_project(self._h, float(tol), int(max_iters), ctypes.byref(it), ctypes.byref(rel))
        return st, it.value, rel.value

    def adj_p2g(self, pos, vel, xbar, vbar):
        """Adjoint of P2G (fold, Pi, deposit): accumulates into xbar, vbar (modified in place)."""
        pos, vel, f = _particles(pos, vel)
        assert xbar.dtype == np.float64 and vbar.dtype == np.float64 and xbar.flags.c_contiguous
        lib().og_adj_p2g(self._h, f, pos.ctypes.data, vel.ctypes.data, len(pos), _p(xbar, _D), _p(vbar, _D))

    def step_adjoint(self, pos, vel, xbar, vbar, alpha, dt, tol=1e-12, max_iters=400):
        """dL/d(x, v) of the last step run with these particles, from dL/d(x', v')."""
        pos, vel, f = _particles(pos, vel)
        xb = np.array(xbar, np.float64).reshape(-1, 3)
        vb = np.array(vbar, np.float64).reshape(-1, 3)
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        st = lib().og_step_adjoint(self._h, f, pos.ctypes.data, vel.ctypes.data, len(pos), float(alpha), float(dt),
                                   float(tol), int(max_iters), _p(xb, _D), _p(vb, _D), ctypes.byref(it),
                                   ctypes.byref(rel))
        if st not in (0,):
            raise RuntimeError(f"adjoint solve status {st}")
        return xb, vb

    def fine(self, name):
        out = np.zeros(self.nb * 3 * 512, np.float64)
        lib().og_get_fine(self._h, FINE_FIELDS[name], _p(out, _D))
        return out

    def set_fine(self, name, vals):
        vals = np.ascontiguousarray(vals, np.float64)
        lib().og_set_fine(self._h, FINE_FIELDS[name], _p(vals, _D))

    def pi(self, face_name, fine_name):
        lib().og_pi(self._h, FACE_FIELDS[face_name], FINE_FIELDS[fine_name])

    def pi_transpose(self, fine_name, face_name):
        lib().og_pi_transpose(self._h, FINE_FIELDS[fine_name], FACE_FIELDS[face_name])


def rollout64(og, pos, vel, steps, dt, gravity, alpha, tol=1e-10, frozen=None):
    """FP64 rollout; returns the states [(x_k, v_k)] and the (levels, liquid) each step used.
    `frozen` (a list of (levels, liquid)) freezes every step's topology and mask."""
    states = [(np.array(pos, np.float64), np.array(vel, np.float64))]
    masks = []
    for k in range(steps):
        lv, lq = (None, None) if frozen is None else frozen[k]
        r = og.step64(states[-1][0], states[-1][1], dt, gravity, alpha, tol=tol, levels=lv, liquid=lq)
        if r["status"] != 0:
            raise RuntimeError(f"oracle step {k} did not converge")
        masks.append((r["levels"], r["liquid"]))
        states.append((r["pos"], r["vel"]))
    return states, masks


def rollout_grad64(og, states, masks, xbar, vbar, dt, gravity, alpha, tol=1e-10, adj_tol=1e-12):
    """Reverse pass over a stored FP64 rollout: re-runs each step's forward (frozen to its masks) to
    restore the intermediates, then its adjoint.  Returns dL/d(x_0, v_0)."""
    xb, vb = np.array(xbar, np.float64), np.array(vbar, np.float64)
    for k in range(len(masks) - 1, -1, -1):
        x, v = states[k]
        og.step64(x, v, dt, gravity, alpha, tol=tol, levels=masks[k][0], liquid=masks[k][1])
        xb, vb = og.step_adjoint(x, v, xb, vb, alpha, dt, tol=adj_tol)
    return xb, vb


def morton3d(x, y, z):
    return lib().og_morton3d(x, y, z)


def num_threads():
    return lib().og_num_threads()


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops from here on."""
    if lib().og_set_threads(int(n)) != 0:
        raise ValueError(n)
