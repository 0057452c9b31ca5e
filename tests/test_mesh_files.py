"""Mesh files: the reference's own mesh assets (tests/golden/mesh_assets.npz,
written by tests/golden/make_mesh_fixtures.py) read by the engine's loader.
CPU only; the GPU steps on these meshes are in test_gpu_mesh_files.py."""
from pathlib import Path

import numpy as np
import pytest

import octopusengine_amd as oe

ASSETS = np.load(Path(__file__).resolve().parent / "golden" / "mesh_assets.npz")


def asset(tmp_path, name):
    p = tmp_path / name
    p.write_bytes(ASSETS[name].tobytes())
    return p


def _tokens_after(lines, tag, count, conv):
    i = next(k for k, l in enumerate(lines) if l.startswith(tag))
    out = []
    for l in lines[i + 1:]:
        out += [conv(v) for v in l.split()]
        if len(out) >= count:
            break
    return out[:count]


def _parse_vtk(text):
    """Plain token reader of the legacy-VTK sections (independent of the
    engine's loader)."""
    lines = text.split("\n")
    np_ = int(next(l for l in lines if l.startswith("POINTS")).split()[1])
    pts = np.array(_tokens_after(lines, "POINTS", 3 * np_, float), np.float32).reshape(-1, 3)
    ncell1, nids = (int(v) for v in next(l for l in lines if l.startswith("CELLS")).split()[1:3])
    conn = np.array(_tokens_after(lines, "CONNECTIVITY", nids, int))
    nct = int(next(l for l in lines if l.startswith("CELL_TYPES")).split()[1])
    types = np.array(_tokens_after(lines, "CELL_TYPES", nct, int))
    return pts, conn, types


def test_bunny_p2_roundtrip(tmp_path):
    """bunny_P2.vtk was written by the reference's writer: reading it and
    writing it back gives the same mesh sections byte for byte."""
    p = asset(tmp_path, "bunny_P2.vtk")
    xyz, topos = oe.load_mesh(p)
    assert xyz.shape == (2918, 3) and list(topos) == [oe.TETRA10] and topos[oe.TETRA10].shape == (1603, 10)
    pts, conn, types = _parse_vtk(p.read_text())
    assert np.array_equal(xyz, pts) and np.array_equal(topos[oe.TETRA10].reshape(-1), conn)
    assert set(types.tolist()) == {24}
    q = tmp_path / "w.vtk"
    oe.VtkWriter(q).mesh(oe.TETRA10, xyz, topos[oe.TETRA10]).close()
    a, b = p.read_text(), q.read_text()
    assert a[a.index("DATASET"):a.index("CELL_TYPES")] == b[b.index("DATASET"):b.index("CELL_TYPES")]
    # a real curved P2 mesh: per-point element records, positive volumes
    assert oe.host_affine_records(oe.TETRA10, xyz, topos[oe.TETRA10]) is None
    assert oe.host_fem_constants(oe.TETRA10, xyz, topos[oe.TETRA10])[..., 9].min() > 0


def test_sphere_p3_double_points(tmp_path):
    """POINTS stored as double are read as float (as the reference does)."""
    p = asset(tmp_path, "sphere_P3.vtk")
    xyz, topos = oe.load_mesh(p)
    assert xyz.shape == (860, 3) and topos[oe.TETRA20].shape == (142, 20)
    pts, conn, types = _parse_vtk(p.read_text())
    assert np.array_equal(xyz, pts) and np.array_equal(topos[oe.TETRA20].reshape(-1), conn)
    assert set(types.tolist()) == {71}
    assert oe.host_fem_constants(oe.TETRA20, xyz, topos[oe.TETRA20])[..., 9].min() > 0


def test_armadillo_q1_wrapped_lines(tmp_path):
    """A ParaView-written file (several values per line, field metadata)."""
    p = asset(tmp_path, "armadillo3.vtk")
    xyz, topos = oe.load_mesh(p)
    assert xyz.shape == (540, 3) and topos[oe.HEXA].shape == (270, 8)
    pts, conn, _ = _parse_vtk(p.read_text())
    assert np.array_equal(xyz, pts) and np.array_equal(topos[oe.HEXA].reshape(-1), conn)


def test_gmsh_bar(tmp_path):
    """bar_tetra_864.msh: 250 nodes numbered from 0, 864 linear tetrahedra
    filling a 4 x 4 x 9 bar (total volume 144)."""
    p = asset(tmp_path, "bar_tetra_864.msh")
    xyz, topos = oe.load_mesh(p)
    assert xyz.shape == (250, 3) and topos[oe.TETRA].shape == (864, 4)
    assert topos[oe.TETRA].min() == 0 and topos[oe.TETRA].max() == 249
    V = oe.host_fem_constants(oe.TETRA, xyz, topos[oe.TETRA])[..., 9]
    assert V.min() > 0 and abs(float(V.sum()) - 144.0) < 1e-3
    assert np.allclose(xyz.min(0), 0) and np.allclose(xyz.max(0), [4, 4, 9])


def test_hexa27_roundtrip_through_vtk_order(tmp_path):
    xyz, topo = oe.beam_mesh(oe.HEXA27, (2, 1, 1))
    q = tmp_path / "q2.vtk"
    oe.VtkWriter(q).mesh(oe.HEXA27, xyz, topo).close()
    x2, t2 = oe.load_mesh(q)
    assert np.array_equal(t2[oe.HEXA27], topo)
    assert np.abs(x2 - xyz).max() <= 1e-6


def test_loader_errors(tmp_path):
    with pytest.raises(oe.OEError):
        oe.load_mesh(tmp_path / "missing.vtk")
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\n")
    with pytest.raises(oe.OEError):
        oe.load_mesh(bad)
    trunc = tmp_path / "t.vtk"
    text = ASSETS["bunny_P2.vtk"].tobytes().decode()
    trunc.write_text(text[: text.index("CONNECTIVITY") + 200])
    with pytest.raises(oe.OEError):
        oe.load_mesh(trunc)
