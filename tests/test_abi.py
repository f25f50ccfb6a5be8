"""The C ABI library loads on the CPU and exports every entry point include/aw.h declares."""
import ctypes
import os
import re

import pytest

import anitowave_adaptive_amd as aw


def declared_functions():
    text = open(os.path.join(aw.INCLUDE_DIR, "aw.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aw_[a-z0-9_]+)\s*\(", text)))


def test_header_and_bindings_agree():
    decl = declared_functions()
    assert len(decl) >= 30
    assert sorted(aw.SIGNATURES) == decl


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(aw.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_status_strings_and_errors_without_device():
    L = aw.lib()
    assert L.aw_status_str(aw.AW_ERR_NOT_CONVERGED) == b"not converged"
    assert L.aw_status_str(99) == b"unknown status"
    # NULL handles are rejected before any device call
    assert L.aw_grid_destroy(None) == aw.AW_ERR_INVALID
    assert b"NULL" in L.aw_last_error() or b"grid" in L.aw_last_error()
    assert L.aw_solver_solve(None, 1e-6, 10, None) == aw.AW_ERR_INVALID
    assert L.aw_step(None, None, None, None, None, 0, None, None) == aw.AW_ERR_INVALID
    assert L.aw_ctx_create(0, None, None) == aw.AW_ERR_INVALID


def test_struct_sizes_match_c_layout():
    # sizes fixed by the C declarations (x86-64 SysV)
    assert ctypes.sizeof(aw.GridDesc) == 4 * 3 + 4 + 4 + 4 + 4 * 3 + 4 * 4
    assert ctypes.sizeof(aw.SolverParams) == 12 * 4
    assert ctypes.sizeof(aw.SolveStats) == 32
    # dt, gravity, alpha, pad, rel_tol, max_iters, rebuild, warm_start, pad, d_radius, d_density
    assert ctypes.sizeof(aw.StepParams) == 4 + 12 + 4 + 4 + 8 + 4 + 4 + 4 + 4 + 8 + 8


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof and field offsets of every aw.h struct, from gcc, against the ctypes mirrors."""
    import subprocess
    structs = {"aw_grid_desc": aw.GridDesc, "aw_layout": aw.Layout, "aw_solver_params": aw.SolverParams,
               "aw_solve_stats": aw.SolveStats, "aw_step_params": aw.StepParams, "aw_step_stats": aw.StepStats,
               "aw_level_info": aw.LevelInfo, "aw_adjoint_layout": aw.AdjointLayout,
               "aw_rollout_stats": aw.RolloutStats}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "aw.h"', "int main(void) {"]
    for cname, py in structs.items():
        src.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'  printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src.append("  return 0;\n}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", aw.INCLUDE_DIR, str(c), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.split("\n") if line)
    for cname, py in structs.items():
        assert int(out[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)
