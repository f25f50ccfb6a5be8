"""Uniform resample (the oracle's restatement of sampleMSBG,
vtx_visualization.cuh:8-62) and the binary legacy VTK writers of the C ABI
(host-only calls; no device needed).

The oracle's resample is pinned by a brute force in exact rational arithmetic
that finds each voxel centre's fine cell as floor of a Fraction and the cell's
slot from block offsets summed independently; output sizes include powers of
two (centres on cell faces), sizes equal to the fine grid, and odd sizes.
The writers are read back by a small parser of the legacy format.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import anitowave_adaptive_amd as aw
import oracle as O


def brute_resample(D, levels, field, out):
    r = 8 >> levels.astype(int)
    off = np.concatenate([[0], np.cumsum(r ** 3)])
    N = [8 * d for d in D]
    res = np.zeros(int(np.prod(out)))
    for z in range(out[2]):
        for y in range(out[1]):
            for x in range(out[0]):
                q = (x, y, z)
                f = [math.floor((q[a] + Fraction(1, 2)) * N[a] / out[a]) for a in range(3)]
                b = f[0] // 8 + D[0] * (f[1] // 8 + D[1] * (f[2] // 8))
                s = 1 << int(levels[b])
                lc = [(f[a] % 8) // s for a in range(3)]
                rr = int(r[b])
                res[x + out[0] * (y + out[1] * z)] = field[off[b] + lc[0] + rr * (lc[1] + rr * lc[2])]
    return res


@pytest.mark.parametrize("out", [(8, 8, 8), (16, 8, 4), (24, 16, 32), (7, 13, 5), (12, 9, 20)])
def test_oracle_resample_brute_force(out):
    D = (3, 2, 4)
    rng = np.random.default_rng(sum(out))
    g = O.OracleGrid(D, num_levels=4, h=1.0 / 32)
    lv = rng.integers(0, 4, D[0] * D[1] * D[2]).astype(np.uint8)
    g.set_levels(lv)
    field = rng.standard_normal(int(g.totals()[0]))
    np.testing.assert_array_equal(g.resample(field, out), brute_resample(D, lv, field, out))


def test_oracle_resample_fine_grid_is_the_field():
    """One native level and one voxel per fine cell: the output is the field
    in x-fastest order over the whole grid."""
    D = (2, 3, 2)
    g = O.OracleGrid(D, num_levels=1, h=1.0 / 16)
    n = int(g.totals()[0])
    field = np.arange(n, dtype=np.float64)
    out = g.resample(field, (16, 24, 16)).reshape(16, 24, 16)
    for (z, y, x) in [(0, 0, 0), (15, 23, 15), (9, 17, 3), (8, 8, 8)]:
        b = x // 8 + 2 * (y // 8 + 3 * (z // 8))
        assert out[z, y, x] == b * 512 + (x % 8) + 8 * ((y % 8) + 8 * (z % 8))


# ---- legacy VTK reader (binary sections big-endian) ----

def _read_vtk(path):
    data = open(path, "rb").read()
    pos = 0
    sections = {}

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    head = [line() for _ in range(4)]
    while pos < len(data):
        s = line()
        if not s:
            continue
        tok = s.split()
        key = tok[0]
        if key in ("DIMENSIONS", "ORIGIN", "SPACING"):
            sections[key] = [float(v) for v in tok[1:]]
        elif key in ("POINT_DATA", "CELL_DATA"):
            sections[key] = int(tok[1])
        elif key == "LOOKUP_TABLE":
            continue
        else:
            if key == "POINTS":
                n, dt = int(tok[1]) * 3, ">f4"
            elif key in ("CELLS", "VERTICES"):
                n, dt = int(tok[2]), ">i4"
            elif key == "CELL_TYPES":
                n, dt = int(tok[1]), ">i4"
            elif key == "SCALARS":
                s2 = line()
                assert s2 == "LOOKUP_TABLE default"
                n = sections.get("POINT_DATA", sections.get("CELL_DATA"))
                dt = ">f4" if tok[2] == "float" else ">i4"
                key = "SCALARS " + tok[1]
            elif key == "VECTORS":
                n, dt = 3 * sections["POINT_DATA"], ">f4"
                key = "VECTORS " + tok[1]
            else:
                raise AssertionError(f"unexpected line {s!r}")
            sections[key] = np.frombuffer(data, dt, n, pos).copy()
            pos += 4 * n
    return head, sections


def test_write_structured_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    dims = (5, 3, 4)
    vals = rng.standard_normal(60).astype(np.float32)
    p = tmp_path / "f.vtk"
    aw.write_vtk_structured(p, "pressure", vals, dims, (0.1, 0.2, 0.3), (0.5, 0.25, 0.125))
    head, sec = _read_vtk(p)
    assert head[0] == "# vtk DataFile Version 3.0" and head[2] == "BINARY" and head[3] == "DATASET STRUCTURED_POINTS"
    assert sec["DIMENSIONS"] == [5, 3, 4] and sec["POINT_DATA"] == 60
    np.testing.assert_allclose(sec["ORIGIN"], [0.1, 0.2, 0.3], rtol=1e-7)
    assert sec["SPACING"] == [0.5, 0.25, 0.125]
    np.testing.assert_array_equal(sec["SCALARS pressure"], vals)
    with pytest.raises(aw.AWError):
        aw.write_vtk_structured(p, "two words", vals, dims, (0, 0, 0), (1, 1, 1))
    with pytest.raises(aw.AWError):
        aw.write_vtk_structured(tmp_path / "missing" / "f.vtk", "p", vals, dims, (0, 0, 0), (1, 1, 1))


def test_write_blocks_round_trip(tmp_path):
    dims = (3, 2, 2)
    lv = np.array([0, 1, 2, 3, 1, 0, 2, 2, 1, 0, 3, 1], np.uint8)
    p = tmp_path / "b.vtk"
    aw.write_vtk_blocks(p, dims, 0.25, lv)
    head, sec = _read_vtk(p)
    assert head[3] == "DATASET UNSTRUCTURED_GRID"
    pts = sec["POINTS"].reshape(12, 8, 3)
    # block 10 = 1 + 3 (1 + 2 * 1) = (1, 1, 1): corners in hexahedron order
    c = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    np.testing.assert_array_equal(pts[10], (np.array([1, 1, 1]) + c) * 0.25)
    np.testing.assert_array_equal(pts[7], (np.array([1, 0, 1]) + c) * 0.25)
    cells = sec["CELLS"].reshape(12, 9)
    assert np.all(cells[:, 0] == 8)
    np.testing.assert_array_equal(cells[:, 1:], np.arange(96).reshape(12, 8))
    assert np.all(sec["CELL_TYPES"] == 12)
    np.testing.assert_array_equal(sec["SCALARS refineLevel"], lv)


@pytest.mark.parametrize("with_data", [False, True])
def test_write_particles_round_trip(tmp_path, with_data):
    rng = np.random.default_rng(2)
    pos = rng.random((1000, 3)).astype(np.float32)
    r = rng.random(1000).astype(np.float32) if with_data else None
    v = rng.standard_normal((1000, 3)).astype(np.float32) if with_data else None
    p = tmp_path / "p.vtk"
    aw.write_vtk_particles(p, pos, r, v)
    head, sec = _read_vtk(p)
    assert head[3] == "DATASET POLYDATA"
    np.testing.assert_array_equal(sec["POINTS"].reshape(-1, 3), pos)
    vc = sec["VERTICES"].reshape(-1, 2)
    assert np.all(vc[:, 0] == 1)
    np.testing.assert_array_equal(vc[:, 1], np.arange(1000))
    if with_data:
        np.testing.assert_array_equal(sec["SCALARS radius"], r)
        np.testing.assert_array_equal(sec["VECTORS velocity"].reshape(-1, 3), v)
    else:
        assert "POINT_DATA" not in sec
