"""MI355X-native MSBG FLIP step: ctypes host over the C ABI in include/aw.h.

The classes mirror the reference's host surface (MSBGManager, P2GTransfer,
G2PManager, AdaptiveSmoother, MultigridTransfer, VCycleSolver; reference
include/*.cuh) on top of the HIP library libanitowave.so built in-tree by
anitowave_adaptive_amd.build.  There is no CPU fallback: if the library is
missing, loading fails with an error.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libanitowave.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

AW_OK, AW_ERR_INVALID, AW_ERR_OOM, AW_ERR_DEVICE, AW_ERR_NOT_CONVERGED, AW_ERR_NAN, AW_ERR_UNSUPPORTED = range(7)
MAX_LEVELS = 4

_V = ctypes.c_void_p
_P = ctypes.POINTER


class GridDesc(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 3), ("base_res", ctypes.c_int), ("num_levels", ctypes.c_int),
                ("h", ctypes.c_float), ("domain_min", ctypes.c_float * 3), ("band", ctypes.c_int * MAX_LEVELS)]


class Layout(ctypes.Structure):
    _fields_ = [("num_blocks", ctypes.c_uint32), ("total_cells", ctypes.c_uint32),
                ("total_faces", ctypes.c_uint32 * 3), ("num_mg_levels", ctypes.c_int),
                ("d_level", _V), ("d_cell_off", _V), ("d_face_off", _V * 3), ("d_particle_count", _V),
                ("d_particle_off", _V), ("d_pressure", _V), ("d_rhs", _V), ("d_liquid", _V), ("d_beta", _V * 3),
                ("d_u_old", _V * 3), ("d_u_star", _V * 3), ("d_u_new", _V * 3), ("d_mass", _V * 3), ("d_fine", _V)]


class SolverParams(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("pre_sweeps", ctypes.c_int), ("post_sweeps", ctypes.c_int),
                ("coarse_sweeps", ctypes.c_int), ("local_sweeps", ctypes.c_int), ("omega", ctypes.c_float),
                ("prolong_scale", ctypes.c_float), ("precision", ctypes.c_int), ("adaptive", ctypes.c_int),
                ("theta", ctypes.c_float), ("max_block_sweeps", ctypes.c_int), ("block_reduction", ctypes.c_float)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("b_norm", ctypes.c_double), ("recurrence_rel_residual", ctypes.c_double)]


class StepParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_float), ("gravity", ctypes.c_float * 3), ("flip_alpha", ctypes.c_float),
                ("rel_tol", ctypes.c_double), ("max_iters", ctypes.c_int), ("rebuild_topology", ctypes.c_int),
                ("warm_start", ctypes.c_int), ("d_radius", _V), ("d_density", _V)]


class StepStats(ctypes.Structure):
    _fields_ = [("topology_changed", ctypes.c_int), ("solve", SolveStats),
                ("num_blocks_by_level", ctypes.c_uint32 * MAX_LEVELS), ("total_cells", ctypes.c_uint32),
                ("liquid_cells", ctypes.c_uint32)]


class LevelInfo(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 3), ("num_blocks", ctypes.c_uint32), ("total_cells", ctypes.c_uint32),
                ("total_faces", ctypes.c_uint32 * 3), ("aggregated", ctypes.c_int), ("d_e", _V),
                ("d_cell_off", _V), ("d_x", _V), ("d_b", _V), ("d_r", _V), ("d_liquid", _V), ("d_x64", _V),
                ("d_block_sweeps", _V), ("d_active", _V), ("smoothed_blocks", ctypes.c_uint32)]


class AdjointLayout(ctypes.Structure):
    _fields_ = [("d_ubar_new", _V * 3), ("d_ubar_old", _V * 3), ("d_ubar_star", _V * 3), ("d_Mbar", _V * 3),
                ("d_mbar", _V * 3), ("d_pbar", _V), ("d_rhsbar", _V), ("d_fine_bar", _V), ("d_fine_mbar", _V),
                ("d_sorted_ids", _V), ("num_sorted", ctypes.c_int64)]


class RolloutStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("forward_ms", ctypes.c_double), ("recompute_ms", ctypes.c_double),
                ("backward_ms", ctypes.c_double), ("forward_iterations", ctypes.c_int),
                ("adjoint_iterations", ctypes.c_int), ("forward_rel_residual", ctypes.c_double),
                ("adjoint_rel_residual", ctypes.c_double), ("replay_mismatched_steps", ctypes.c_int),
                ("replay_max_state_diff", ctypes.c_double)]


_I = ctypes.c_int
_I64 = ctypes.c_int64
_F = ctypes.c_float
_D = ctypes.c_double

# name -> (restype, argtypes); every entry is declared in include/aw.h
SIGNATURES = {
    "aw_status_str": (ctypes.c_char_p, [_I]),
    "aw_last_error": (ctypes.c_char_p, []),
    "aw_version": (_I, []),
    "aw_ctx_create": (_I, [_I, _V, _P(_V)]),
    "aw_ctx_destroy": (_I, [_V]),
    "aw_ctx_synchronize": (_I, [_V]),
    "aw_malloc": (_I, [_V, ctypes.c_uint64, _P(_V)]),
    "aw_free": (_I, [_V, _V]),
    "aw_copy": (_I, [_V, _V, _V, ctypes.c_uint64]),
    "aw_fill_f32": (_I, [_V, _V, ctypes.c_uint64, _F]),
    "aw_memory_stats": (_I, [_P(ctypes.c_uint64), _P(ctypes.c_uint64), _I]),
    "aw_set_memory_limit": (_I, [ctypes.c_uint64]),
    "aw_profile_enable": (_I, [_V, _I]),
    "aw_profile_report": (_I, [_V, ctypes.c_char_p, ctypes.c_uint64]),
    "aw_grid_create": (_I, [_V, _P(GridDesc), _P(_V)]),
    "aw_grid_destroy": (_I, [_V]),
    "aw_grid_build_topology": (_I, [_V, _V, _I64, _P(_I)]),
    "aw_grid_build_topology_radius": (_I, [_V, _V, _V, _I64, _P(_I)]),
    "aw_grid_set_levels": (_I, [_V, _V]),
    "aw_grid_beta_changed": (_I, [_V]),
    "aw_grid_layout": (_I, [_V, _P(Layout)]),
    "aw_bin_particles": (_I, [_V, _V, _V, _V, _I64]),
    "aw_p2g": (_I, [_V, _V, _V, _I64, _F, _P(_F)]),
    "aw_divergence": (_I, [_V]),
    "aw_project_velocity": (_I, [_V]),
    "aw_g2p_advect": (_I, [_V, _V, _V, _I64, _F, _F]),
    "aw_p2g_particles": (_I, [_V, _V, _V, _I64, _F, _P(_F)]),
    "aw_g2p": (_I, [_V, _V, _V, _I64, _F]),
    "aw_advect": (_I, [_V, _V, _I64, _F]),
    "aw_solver_create": (_I, [_V, _P(SolverParams), _P(_V)]),
    "aw_solver_destroy": (_I, [_V]),
    "aw_solver_set_graph": (_I, [_V, _I]),
    "aw_solver_setup": (_I, [_V]),
    "aw_solver_level_info": (_I, [_V, _I, _P(LevelInfo)]),
    "aw_solver_smooth": (_I, [_V, _I, _I, _I]),
    "aw_solver_residual": (_I, [_V, _I]),
    "aw_solver_apply": (_I, [_V, _I, _V, _V]),
    "aw_solver_restrict": (_I, [_V, _I]),
    "aw_solver_prolong": (_I, [_V, _I]),
    "aw_solver_vcycle": (_I, [_V, _I]),
    "aw_solver_solve": (_I, [_V, _D, _I, _P(SolveStats)]),
    "aw_step": (_I, [_V, _V, _V, _V, _V, _I64, _P(StepParams), _P(StepStats)]),
    "aw_grid_adjoint_layout": (_I, [_V, _P(AdjointLayout)]),
    "aw_g2p_advect_adjoint": (_I, [_V, _F, _F, _V, _V]),
    "aw_project_adjoint": (_I, [_V, _D, _I, _P(SolveStats)]),
    "aw_p2g_adjoint": (_I, [_V, _V, _V]),
    "aw_step_adjoint": (_I, [_V, _V, _P(StepParams), _V, _V, _P(SolveStats)]),
    "aw_tape_create": (_I, [_V, _V, _I64, _I, _I, _I, _P(_V)]),
    "aw_tape_destroy": (_I, [_V]),
    "aw_rollout_forward": (_I, [_V, _V, _V, _V, _P(StepParams), _P(RolloutStats)]),
    "aw_rollout_backward": (_I, [_V, _V, _V, _P(RolloutStats)]),
    "aw_tape_masks": (_I, [_V, _I, _V, _V, ctypes.c_uint32, _P(ctypes.c_uint32)]),
    "aw_resample_cells": (_I, [_V, _V, _P(_I), _V]),
    "aw_remap_cells": (_I, [_V, _V, _V, _V]),
    "aw_remap_cells_noalloc": (_I, [_V, _V, _V, _V]),
    "aw_write_vtk_structured": (_I, [ctypes.c_char_p, ctypes.c_char_p, _V, _P(_I), _P(_F), _P(_F)]),
    "aw_write_vtk_blocks": (_I, [ctypes.c_char_p, _P(_I), _F, _V]),
    "aw_write_vtk_particles": (_I, [ctypes.c_char_p, _V, _V, _V, _I64]),
    "aw_export_cells_vtk": (_I, [_V, _V, _P(_I), ctypes.c_char_p, ctypes.c_char_p]),
    "aw_export_blocks_vtk": (_I, [_V, ctypes.c_char_p]),
}

_lib = None


class AWError(RuntimeError):
    def __init__(self, status, where, detail):
        super().__init__(f"{where}: {detail} (status {status})")
        self.status = status


def lib():
    """Load libanitowave.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m anitowave_adaptive_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != AW_OK:
        raise AWError(status, where, lib().aw_last_error().decode())
    return status


class Context:
    """One device and stream (aw_ctx)."""

    def __init__(self, device=0, stream=None):
        self._h = _V()
        check(lib().aw_ctx_create(device, stream, ctypes.byref(self._h)), "aw_ctx_create")

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        check(lib().aw_ctx_synchronize(self._h), "aw_ctx_synchronize")

    def profile(self, on=True, runs=False):
        """Kernel timing with HIP events: per launch, or (runs=True) per run of
        back-to-back launches of one kernel group."""
        check(lib().aw_profile_enable(self._h, (2 if runs else 1) if on else 0), "aw_profile_enable")

    def profile_report(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib().aw_profile_report(self._h, buf, len(buf)), "aw_profile_report")
        return json.loads(buf.value.decode())

    def close(self):
        if self._h:
            lib().aw_ctx_destroy(self._h)
            self._h = _V()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """Device buffer owned by the library's allocator (aw_malloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = _V()
        check(lib().aw_malloc(ctx.handle, self.nbytes, ctypes.byref(self.ptr)), "aw_malloc")

    @classmethod
    def from_numpy(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(ctx, arr.nbytes)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().aw_copy(self.ctx.handle, self.ptr, arr.ctypes.data_as(_V), arr.nbytes), "aw_copy")

    def download(self, dtype, count=None):
        return download(self.ctx, self.ptr, dtype, count if count is not None else self.nbytes // np.dtype(dtype).itemsize)

    def free(self):
        if self.ptr:
            lib().aw_free(self.ctx.handle, self.ptr)
            self.ptr = _V()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def memory_stats(reset_peak=False):
    """(current, peak) bytes of device memory the library holds in this process."""
    cur, peak = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().aw_memory_stats(ctypes.byref(cur), ctypes.byref(peak), int(bool(reset_peak))), "aw_memory_stats")
    return cur.value, peak.value


def set_memory_limit(nbytes):
    """Cap on the device memory the library holds (0: none); past it, AW_ERR_OOM."""
    check(lib().aw_set_memory_limit(int(nbytes)), "aw_set_memory_limit")


def download(ctx, ptr, dtype, count):
    out = np.empty(int(count), dtype)
    if count:
        check(lib().aw_copy(ctx.handle, out.ctypes.data_as(_V), ptr, out.nbytes), "aw_copy")
    return out


def upload(ctx, ptr, arr):
    arr = np.ascontiguousarray(arr)
    if arr.nbytes:
        check(lib().aw_copy(ctx.handle, ptr, arr.ctypes.data_as(_V), arr.nbytes), "aw_copy")


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def write_vtk_structured(path, name, data, dims, origin, spacing):
    """Binary legacy VTK STRUCTURED_POINTS scalar field; data has dims[0] *
    dims[1] * dims[2] float32 values, x fastest (host only)."""
    data = np.ascontiguousarray(data, np.float32).reshape(-1)
    if data.size != int(dims[0]) * int(dims[1]) * int(dims[2]):
        raise ValueError("data size does not match dims")
    check(lib().aw_write_vtk_structured(os.fsencode(path), name.encode(), data.ctypes.data_as(_V), _i3(dims),
                                        _f3(origin), _f3(spacing)), "aw_write_vtk_structured")


def write_vtk_blocks(path, dims, block_size, levels):
    """One hexahedron per block with its level as refineLevel (host only)."""
    lv = np.ascontiguousarray(levels, np.uint8).reshape(-1)
    if lv.size != int(dims[0]) * int(dims[1]) * int(dims[2]):
        raise ValueError("levels size does not match dims")
    check(lib().aw_write_vtk_blocks(os.fsencode(path), _i3(dims), float(block_size), lv.ctypes.data_as(_V)),
          "aw_write_vtk_blocks")


def write_vtk_particles(path, pos, radius=None, vel=None):
    """Particles as POLYDATA points with one vertex cell each, optional
    radius and velocity (host only)."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    r = None if radius is None else np.ascontiguousarray(radius, np.float32).reshape(n)
    v = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(n, 3)
    check(lib().aw_write_vtk_particles(os.fsencode(path), pos.ctypes.data_as(_V),
                                       None if r is None else r.ctypes.data_as(_V),
                                       None if v is None else v.ctypes.data_as(_V), n), "aw_write_vtk_particles")


class MSBGManager:
    """The block grid (MSBGManager, reference include/MSBG.cuh:502-677).

    dims: blocks per axis (8^3 cells at native level 0); h: finest cell size;
    num_levels native levels; band[l]: level-l band width in blocks."""

    def __init__(self, ctx, dims, h, num_levels=1, band=(0, 0, 0, 0)):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        d = GridDesc()
        d.dims[:] = list(self.dims)
        d.base_res = 8
        d.num_levels = int(num_levels)
        d.h = float(h)
        d.domain_min[:] = [0.0, 0.0, 0.0]
        d.band[:] = [int(x) for x in (list(band) + [0, 0, 0, 0])[:4]]
        self.desc = d
        self._h = _V()
        check(lib().aw_grid_create(ctx.handle, ctypes.byref(d), ctypes.byref(self._h)), "aw_grid_create")
        self.num_blocks = self.dims[0] * self.dims[1] * self.dims[2]

    @property
    def handle(self):
        return self._h

    def build_topology(self, d_pos, n):
        ch = ctypes.c_int(0)
        check(lib().aw_grid_build_topology(self._h, d_pos, int(n), ctypes.byref(ch)), "aw_grid_build_topology")
        return bool(ch.value)

    def build_topology_radius(self, d_pos, d_radius, n):
        """Levels from per-particle radii (the reference's markActiveBlocks
        rule, MSBG.cuh:271-314, with grading to a fixed point); d_radius is
        a device float[n] in d_pos's order."""
        ch = ctypes.c_int(0)
        check(lib().aw_grid_build_topology_radius(self._h, d_pos, d_radius, int(n), ctypes.byref(ch)),
              "aw_grid_build_topology_radius")
        return bool(ch.value)

    def resample(self, d_field, out_dims, d_out):
        """Nearest-cell resample of a device cell field onto out_dims voxel
        centres, x fastest (sampleMSBG, vtx_visualization.cuh:8-62)."""
        check(lib().aw_resample_cells(self._h, d_field, _i3(out_dims), d_out), "aw_resample_cells")

    def remap_cells(self, d_level_prev, d_src, d_dst):
        """Cell field of the layout with native levels d_level_prev (device
        uint8[num_blocks]) onto the current layout: each cell takes the value
        of the previous cell containing its centre."""
        check(lib().aw_remap_cells(self._h, d_level_prev, d_src, d_dst), "aw_remap_cells")

    def export_cells_vtk(self, d_field, out_dims, path, name="pressure"):
        """Resample and write a binary legacy VTK STRUCTURED_POINTS file
        (exportUniformPressure_VTK, vtx_visualization.cuh:64-126)."""
        check(lib().aw_export_cells_vtk(self._h, d_field, _i3(out_dims), name.encode(), os.fsencode(path)),
              "aw_export_cells_vtk")

    def export_blocks_vtk(self, path):
        """Blocks as hexahedra with their level (exportMultigrid_VTK, vtx_visualization.cuh:128-190)."""
        check(lib().aw_export_blocks_vtk(self._h, os.fsencode(path)), "aw_export_blocks_vtk")

    def beta_changed(self):
        """The face coefficients d_beta were written directly: coarsen them at the next solver setup."""
        check(lib().aw_grid_beta_changed(self._h), "aw_grid_beta_changed")

    def set_levels(self, levels):
        lv = np.ascontiguousarray(levels, np.uint8)
        check(lib().aw_grid_set_levels(self._h, lv.ctypes.data_as(_V)), "aw_grid_set_levels")

    def layout(self):
        L = Layout()
        check(lib().aw_grid_layout(self._h, ctypes.byref(L)), "aw_grid_layout")
        return L

    # host copies of grid arrays (for tests and tools)
    def levels(self):
        L = self.layout()
        return download(self.ctx, L.d_level, np.uint8, L.num_blocks)

    def offsets(self):
        L = self.layout()
        c = download(self.ctx, L.d_cell_off, np.uint32, L.num_blocks + 1)
        f = download(self.ctx, L.d_face_off[0], np.uint32, L.num_blocks + 1)
        return c, f

    def particle_counts(self):
        L = self.layout()
        return download(self.ctx, L.d_particle_count, np.uint32, L.num_blocks)

    def cell(self, name):
        L = self.layout()
        ptr = {"p": L.d_pressure, "rhs": L.d_rhs}.get(name)
        if name == "liquid":
            return download(self.ctx, L.d_liquid, np.uint8, L.total_cells)
        return download(self.ctx, ptr, np.float32, L.total_cells)

    def set_cell(self, name, vals):
        L = self.layout()
        ptr = {"p": L.d_pressure, "rhs": L.d_rhs, "liquid": L.d_liquid}[name]
        dt = np.uint8 if name == "liquid" else np.float32
        upload(self.ctx, ptr, np.ascontiguousarray(vals, dt))

    def face(self, name, axis):
        L = self.layout()
        ptr = {"u_old": L.d_u_old, "u_star": L.d_u_star, "u_new": L.d_u_new, "mass": L.d_mass,
               "beta": L.d_beta}[name][axis]
        return download(self.ctx, ptr, np.float32, L.total_faces[axis])

    def set_face(self, name, axis, vals):
        L = self.layout()
        ptr = {"u_old": L.d_u_old, "u_star": L.d_u_star, "u_new": L.d_u_new, "beta": L.d_beta}[name][axis]
        upload(self.ctx, ptr, np.ascontiguousarray(vals, np.float32))

    def fine(self):
        """Padded fine lattice [block][field][axis][900] (field 0 = u_new, 1 = u_new - u_old): each
        block's 8^3 nodes and a one-node halo, filled for the blocks within one block of a block
        holding particles (other blocks' entries are not current)."""
        L = self.layout()
        return download(self.ctx, L.d_fine, np.float32, L.num_blocks * 5400).reshape(L.num_blocks, 5400)

    def close(self):
        if self._h:
            lib().aw_grid_destroy(self._h)
            self._h = _V()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class P2GTransfer:
    """Particle-to-grid (P2GTransfer::run, reference include/P2G.cuh:173-196)."""

    @staticmethod
    def bin(grid, d_pos, d_vel, d_ids, n):
        check(lib().aw_bin_particles(grid.handle, d_pos, d_vel, d_ids, int(n)), "aw_bin_particles")

    @staticmethod
    def run(grid, d_pos, d_vel, n, dt=0.0, gravity=(0.0, 0.0, 0.0)):
        """P2G of particles binned by `bin` for the current topology."""
        g = (ctypes.c_float * 3)(*gravity)
        check(lib().aw_p2g(grid.handle, d_pos, d_vel, int(n), float(dt), g), "aw_p2g")

    @staticmethod
    def run_unsorted(grid, d_pos, d_vel, n, dt=0.0, gravity=(0.0, 0.0, 0.0)):
        """P2G of particles in any order; the caller's arrays are not reordered."""
        g = (ctypes.c_float * 3)(*gravity)
        check(lib().aw_p2g_particles(grid.handle, d_pos, d_vel, int(n), float(dt), g), "aw_p2g_particles")


class G2PManager:
    """FLIP/PIC update and RK3 advection (G2PManager, reference include/G2P.cuh:175-228)."""

    @staticmethod
    def divergence(grid):
        check(lib().aw_divergence(grid.handle), "aw_divergence")

    @staticmethod
    def project(grid):
        check(lib().aw_project_velocity(grid.handle), "aw_project_velocity")

    @staticmethod
    def g2p_advect(grid, d_pos, d_vel, n, alpha=0.95, dt=1.0 / 240.0):
        check(lib().aw_g2p_advect(grid.handle, d_pos, d_vel, int(n), float(alpha), float(dt)), "aw_g2p_advect")

    @staticmethod
    def perform_g2p(grid, d_vel, d_pos, n, alpha=0.97):
        """FLIP/PIC velocity update in place (performG2P, G2P.cuh:192-214)."""
        check(lib().aw_g2p(grid.handle, d_pos, d_vel, int(n), float(alpha)), "aw_g2p")

    @staticmethod
    def advect(grid, d_pos, n, dt):
        """RK3 advection through u_new in place (G2PManager::advect, G2P.cuh:216-227)."""
        check(lib().aw_advect(grid.handle, d_pos, int(n), float(dt)), "aw_advect")


def solver_params(method=0, pre=2, post=2, coarse=16, local=3, omega=0.85714, prolong_scale=1.0, precision=0,
                  adaptive=0, theta=1e-4, max_block_sweeps=10, block_reduction=0.33):
    """adaptive=1: the reference's adaptive smoothing (method 1 only): early exit
    per block at block_reduction, activation above theta * max|b| (Smoother.cuh)."""
    p = SolverParams()
    p.method, p.pre_sweeps, p.post_sweeps, p.coarse_sweeps, p.local_sweeps = method, pre, post, coarse, local
    p.omega, p.prolong_scale, p.precision = omega, prolong_scale, precision
    p.adaptive, p.theta, p.max_block_sweeps, p.block_reduction = adaptive, theta, max_block_sweeps, block_reduction
    return p


class VCycleSolver:
    """Multigrid pressure solver (VCycleSolver + AdaptiveSmoother +
    MultigridTransfer, reference include/VCycle.cuh, Smoother.cuh,
    MultigridTransfer.cuh)."""

    def __init__(self, grid, **kw):
        self.grid = grid
        self.params = solver_params(**kw)
        self._h = _V()
        check(lib().aw_solver_create(grid.handle, ctypes.byref(self.params), ctypes.byref(self._h)),
              "aw_solver_create")

    @property
    def handle(self):
        return self._h

    def use_graph(self, on=True):
        """Replay each PCG iteration from a captured HIP graph (default off)."""
        check(lib().aw_solver_set_graph(self._h, int(bool(on))), "aw_solver_set_graph")

    def setup(self):
        check(lib().aw_solver_setup(self._h), "aw_solver_setup")

    def level_info(self, k):
        L = LevelInfo()
        check(lib().aw_solver_level_info(self._h, int(k), ctypes.byref(L)), "aw_solver_level_info")
        return L

    def num_levels(self):
        k = 0
        while True:
            L = LevelInfo()
            if lib().aw_solver_level_info(self._h, k, ctypes.byref(L)) != AW_OK:
                return k
            k += 1

    def level_get(self, k, name):
        L = self.level_info(k)
        if name == "liquid":
            return download(self.grid.ctx, L.d_liquid, np.uint8, L.total_cells)
        if name == "e":
            return download(self.grid.ctx, L.d_e, np.uint8, L.num_blocks)
        if name in ("block_sweeps", "active"):
            ptr = L.d_block_sweeps if name == "block_sweeps" else L.d_active
            if not ptr:
                raise ValueError("no adaptive smoothing call on this level yet")
            return download(self.grid.ctx, ptr, np.uint8, L.num_blocks)
        if name == "coff":
            return download(self.grid.ctx, L.d_cell_off, np.uint32, L.num_blocks + 1)
        if name == "x64":
            if not L.d_x64:
                raise ValueError("no FP64 solution (precision 0 or no solve yet)")
            return download(self.grid.ctx, L.d_x64, np.float64, L.total_cells)
        return download(self.grid.ctx, {"x": L.d_x, "b": L.d_b, "r": L.d_r}[name], np.float32, L.total_cells)

    def level_set(self, k, name, vals):
        L = self.level_info(k)
        upload(self.grid.ctx, {"x": L.d_x, "b": L.d_b, "r": L.d_r}[name], np.ascontiguousarray(vals, np.float32))

    def smooth(self, k, iters=1, reverse=False):
        check(lib().aw_solver_smooth(self._h, int(k), int(iters), int(bool(reverse))), "aw_solver_smooth")

    def residual(self, k):
        check(lib().aw_solver_residual(self._h, int(k)), "aw_solver_residual")

    def apply(self, k, d_x, d_y):
        check(lib().aw_solver_apply(self._h, int(k), d_x, d_y), "aw_solver_apply")

    def restrict(self, k):
        check(lib().aw_solver_restrict(self._h, int(k)), "aw_solver_restrict")

    def prolong(self, k):
        check(lib().aw_solver_prolong(self._h, int(k)), "aw_solver_prolong")

    def run_cycle(self, k=0):
        check(lib().aw_solver_vcycle(self._h, int(k)), "aw_solver_vcycle")

    def solve(self, tol=1e-6, max_iters=200, allow_not_converged=False):
        """Solve to true relative residual `tol`; AW_ERR_NOT_CONVERGED raises
        unless allowed (then `self.status` says which it was)."""
        st = SolveStats()
        s = lib().aw_solver_solve(self._h, float(tol), int(max_iters), ctypes.byref(st))
        self.status = s
        if not (allow_not_converged and s == AW_ERR_NOT_CONVERGED):
            check(s, "aw_solver_solve")
        return st

    def close(self):
        if self._h:
            lib().aw_solver_destroy(self._h)
            self._h = _V()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FlipStep:
    """The whole step (aw_step) on device particle buffers."""

    def __init__(self, grid, solver, dt=1.0 / 240.0, gravity=(0.0, -9.81, 0.0), alpha=0.95, tol=1e-6,
                 max_iters=200, rebuild_topology=True, warm_start=False, d_radius=None, d_density=None):
        """rebuild_topology: False keeps the topology, True (1) rebuilds it
        from the distance bands every step, 2 from the particle radii
        d_radius (device float[n] indexed by particle id; the step then needs
        d_ids).  warm_start: start each pressure solve from the previous
        step's pressure (remapped across topology changes) instead of zero.
        d_density: device float[n] relative density per particle id (two-phase
        beta = 1 / rho; the step then needs d_ids), or None for beta = 1."""
        self.grid, self.solver = grid, solver
        p = StepParams()
        p.dt = dt
        p.gravity[:] = list(gravity)
        p.flip_alpha = alpha
        p.rel_tol = tol
        p.max_iters = max_iters
        p.rebuild_topology = int(rebuild_topology)
        p.warm_start = int(warm_start)
        p.d_radius = d_radius
        p.d_density = d_density
        self.params = p

    def __call__(self, d_pos, d_vel, d_ids, n, allow_not_converged=False):
        """One step; a pressure solve above its tolerance (AW_ERR_NOT_CONVERGED,
        e.g. an FP32 solve at its floor) raises unless allowed, and
        `self.status` keeps the step's status either way."""
        st = StepStats()
        s = lib().aw_step(self.grid.handle, self.solver.handle, d_pos, d_vel, d_ids, int(n), ctypes.byref(self.params),
                          ctypes.byref(st))
        self.status = s
        if not (allow_not_converged and s == AW_ERR_NOT_CONVERGED):
            check(s, "aw_step")
        return st


class Adjoint:
    """Reverse-mode derivative of the step on a grid and solver (no reference
    counterpart: the reference has no gradient code).  Particle adjoints are
    float32 (n, 3) device buffers indexed by particle id."""

    def __init__(self, grid, solver):
        self.grid, self.solver = grid, solver

    def layout(self):
        L = AdjointLayout()
        check(lib().aw_grid_adjoint_layout(self.grid.handle, ctypes.byref(L)), "aw_grid_adjoint_layout")
        return L

    def face(self, name, axis):
        L = self.layout()
        n = self.grid.layout().total_faces[axis]
        ptr = {"ubar_new": L.d_ubar_new, "ubar_old": L.d_ubar_old, "ubar_star": L.d_ubar_star, "Mbar": L.d_Mbar,
               "mbar": L.d_mbar}[name][axis]
        return download(self.grid.ctx, ptr, np.float32, n)

    def cell(self, name):
        L = self.layout()
        n = self.grid.layout().total_cells
        return download(self.grid.ctx, {"pbar": L.d_pbar, "rhsbar": L.d_rhsbar}[name], np.float32, n)

    def g2p_advect(self, d_xbar, d_vbar, alpha=0.95, dt=1.0 / 240.0):
        check(lib().aw_g2p_advect_adjoint(self.grid.handle, float(alpha), float(dt), d_xbar, d_vbar),
              "aw_g2p_advect_adjoint")

    def project(self, tol=1e-6, max_iters=200):
        st = SolveStats()
        check(lib().aw_project_adjoint(self.solver.handle, float(tol), int(max_iters), ctypes.byref(st)),
              "aw_project_adjoint")
        return st

    def p2g(self, d_xbar, d_vbar):
        check(lib().aw_p2g_adjoint(self.grid.handle, d_xbar, d_vbar), "aw_p2g_adjoint")

    def step(self, params, d_xbar, d_vbar):
        st = SolveStats()
        check(lib().aw_step_adjoint(self.grid.handle, self.solver.handle, ctypes.byref(params), d_xbar, d_vbar,
                                    ctypes.byref(st)), "aw_step_adjoint")
        return st


class Rollout:
    """Checkpointed rollout of `steps` FLIP steps with its reverse pass
    (aw_tape_*, aw_rollout_*)."""

    def __init__(self, grid, solver, n, steps, checkpoint_every=4, keep_masks=False):
        self.grid, self.solver, self.n, self.steps = grid, solver, int(n), int(steps)
        self._h = _V()
        check(lib().aw_tape_create(grid.handle, solver.handle, self.n, self.steps, int(checkpoint_every),
                                   int(bool(keep_masks)), ctypes.byref(self._h)), "aw_tape_create")
        self.stats = RolloutStats()

    def forward(self, step_params, d_pos, d_vel, d_ids):
        check(lib().aw_rollout_forward(self._h, d_pos, d_vel, d_ids, ctypes.byref(step_params),
                                       ctypes.byref(self.stats)), "aw_rollout_forward")
        return self.stats

    def backward(self, d_xbar, d_vbar):
        check(lib().aw_rollout_backward(self._h, d_xbar, d_vbar, ctypes.byref(self.stats)), "aw_rollout_backward")
        return self.stats

    def masks(self, step):
        lv = np.zeros(self.grid.num_blocks, np.uint8)
        nc = ctypes.c_uint32(0)
        check(lib().aw_tape_masks(self._h, int(step), lv.ctypes.data_as(_V), None, 0, ctypes.byref(nc)),
              "aw_tape_masks")
        lq = np.zeros(nc.value, np.uint8)
        check(lib().aw_tape_masks(self._h, int(step), lv.ctypes.data_as(_V), lq.ctypes.data_as(_V), nc.value,
                                  ctypes.byref(nc)), "aw_tape_masks")
        return lv, lq

    def close(self):
        if self._h:
            lib().aw_tape_destroy(self._h)
            self._h = _V()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
