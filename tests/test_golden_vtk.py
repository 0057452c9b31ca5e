"""Golden vectors from the reference's own result files (CPU only).

tests/golden/vtk_vbd_scaling.npz holds, for eight beams of size 4 x 1 x 1
(P1, Q1, P2, Q2 at two resolutions) written by the reference's VTK_Formater,
the exact text of the mesh sections (POINTS, CELLS/OFFSETS/CONNECTIVITY,
CELL_TYPES) and the per-vertex "weights" (masses, density 1000, FemShape).
They pin the beam generator, the P2/Q2 converters' node numbering, the VTK
writer (Hexa27 node order, number formatting) and the mass distribution, for
both the engine and the oracle.  tests/golden/make_vtk_fixtures.py wrote them.
"""
from pathlib import Path

import numpy as np
import pytest

import octopusengine_amd as oe
import pyoracle as O

GOLDEN = np.load(Path(__file__).resolve().parent / "golden" / "vtk_vbd_scaling.npz")
NAMES = [str(n) for n in GOLDEN["names"]]
SIZE = (4.0, 1.0, 1.0)


def _case(name):
    el = int(GOLDEN[f"{name}__element"])
    cells = tuple(int(c) for c in GOLDEN[f"{name}__cells"])
    mesh = GOLDEN[f"{name}__mesh"].tobytes().decode()
    return el, cells, mesh, GOLDEN[f"{name}__weights"]


def _sections(text):
    lines = text.split("\n")
    a = next(i for i, l in enumerate(lines) if l.startswith("DATASET"))
    b = next(i for i, l in enumerate(lines) if l.startswith("CELL_TYPES"))
    return "\n".join(lines[a:b + 2]) + "\n"


def _parse_mesh(mesh):
    lines = mesh.split("\n")
    n = int(lines[1].split()[1])
    pts = np.array([[float(v) for v in l.split()] for l in lines[2:2 + n]])
    i = lines.index("CONNECTIVITY vtktypeint64")
    j = next(k for k, l in enumerate(lines) if l.startswith("CELL_TYPES"))
    conn = np.array([[int(v) for v in l.split()] for l in lines[i + 1:j]])
    return pts, conn


def _vtk_order(el, topo):
    if el != oe.HEXA27:
        return topo
    perm = {12: 16, 13: 17, 14: 18, 15: 19, 16: 12, 17: 13, 18: 14, 19: 15, 20: 24, 21: 22, 22: 21, 24: 20}
    return topo[:, [perm.get(j, j) for j in range(27)]]


def test_fixture_covers_the_element_types():
    els = {int(GOLDEN[f"{n}__element"]) for n in NAMES}
    assert els == {oe.TETRA, oe.HEXA, oe.TETRA10, oe.HEXA27}
    assert len(NAMES) == 8


@pytest.mark.parametrize("name", NAMES)
def test_engine_mesh_and_writer_reproduce_reference_file(name, tmp_path):
    """Beam generation, converter numbering and the writer give the
    reference's mesh sections byte for byte."""
    el, cells, mesh, _ = _case(name)
    xyz, topo = oe.beam_mesh(el, cells, SIZE)
    # the reference scenes place successive beams 1.1 apart along z
    pts, _ = _parse_mesh(mesh)
    xyz = (xyz + np.float32(pts[0, 2] - xyz[0, 2]) * np.array([0, 0, 1], np.float32)).astype(np.float32)
    path = tmp_path / "beam.vtk"
    oe.VtkWriter(path).mesh(el, xyz, topo).close()
    text = path.read_text()
    assert text.startswith("# vtk DataFile Version 5.1\nvtk output\nASCII\n\n")
    assert _sections(text) == mesh


@pytest.mark.parametrize("name", NAMES)
def test_oracle_mesh_matches_reference_file(name):
    el, cells, mesh, _ = _case(name)
    pts, conn = _parse_mesh(mesh)
    xyz, topo = O.beam(el, cells, SIZE)
    assert np.array_equal(_vtk_order(el, topo), conn)
    # the file keeps 6 significant digits; beams are placed 1.1 apart along z
    xyz = xyz + np.array([0.0, 0.0, pts[0, 2] - xyz[0, 2]])
    assert np.abs(xyz - pts).max() <= 5e-6 * max(1.0, np.abs(pts).max())


@pytest.mark.parametrize("name", NAMES)
def test_masses_match_reference_weights(name):
    el, cells, _, weights = _case(name)
    xyz, topo = oe.beam_mesh(el, cells, SIZE)
    m = oe.host_masses(el, xyz, topo, 1000.0, oe.MASS_FEMSHAPE)
    assert np.abs(m - weights).max() <= 5e-6 * weights.max()
    mo = O.masses(el, xyz, topo, 1000.0, True)
    assert np.abs(mo - weights).max() <= 5e-6 * weights.max()


def test_writer_data_blocks_and_errors(tmp_path):
    xyz, topo = oe.beam_mesh(oe.TETRA, (2, 1, 1))
    path = tmp_path / "d.vtk"
    w = oe.VtkWriter(path).mesh(oe.TETRA, xyz, topo)
    with pytest.raises(oe.OEError):
        w.scalars("early", np.zeros(xyz.shape[0]))       # before POINT_DATA
    w.point_data().scalars("s", np.arange(xyz.shape[0], dtype=np.float32))
    with pytest.raises(oe.OEError):
        w.vectors("bad", np.zeros((xyz.shape[0] - 1, 3)))
    w.vectors("u", np.full((xyz.shape[0], 3), 0.5, np.float32))
    w.cell_data().scalars("c", np.full(topo.shape[0], 2.5e-7, np.float32))
    w.close()
    lines = path.read_text().split("\n")
    i = lines.index(f"POINT_DATA {xyz.shape[0]}")
    assert lines[i + 1] == "SCALARS s float 1" and lines[i + 2] == "LOOKUP_TABLE default"
    assert lines[i + 3].split() == [str(k) for k in range(xyz.shape[0])]
    assert lines[i + 4] == "VECTORS u float"
    assert lines[i + 5] == "0.5 0.5 0.5  " * xyz.shape[0]
    j = lines.index(f"CELL_DATA {topo.shape[0]}")
    assert lines[j + 3].split() == ["2.5e-07"] * topo.shape[0]


# The reference's larger result files (3_*, 4_* and the ref_* solutions):
# SHA-256 of the mesh sections, the beam's z offset in its scene, and the
# weights (float32), written by tests/golden/make_vtk_fixtures.py.  These pin
# the converters' numbering and the masses at the 32 x 8 x 8 Q2 / P2 and
# 64 x 16 x 16 Q1 sizes of the workloads.
LARGE = np.load(Path(__file__).resolve().parent / "golden" / "vtk_vbd_scaling_large.npz")
LARGE_NAMES = [str(n) for n in LARGE["names"]]


def test_large_fixture_covers_the_workload_meshes():
    assert {"4_Hexa27_32_8_8", "4_Tetra10_32_8_8", "4_Hexa_64_16_16"} <= set(LARGE_NAMES)
    assert len(LARGE_NAMES) == 10


@pytest.mark.parametrize("name", LARGE_NAMES)
def test_engine_reproduces_large_reference_mesh_digest(name, tmp_path):
    import hashlib

    el = int(LARGE[f"{name}__element"])
    cells = tuple(int(c) for c in LARGE[f"{name}__cells"])
    xyz, topo = oe.beam_mesh(el, cells, SIZE)
    xyz = (xyz + np.float32(float(LARGE[f"{name}__zoff"]) - xyz[0, 2]) * np.array([0, 0, 1], np.float32))
    path = tmp_path / "beam.vtk"
    oe.VtkWriter(path).mesh(el, xyz.astype(np.float32), topo).close()
    digest = hashlib.sha256(_sections(path.read_text()).encode()).hexdigest()
    assert digest == str(LARGE[f"{name}__mesh_sha256"])


@pytest.mark.parametrize("name", LARGE_NAMES)
def test_masses_match_large_reference_weights(name):
    el = int(LARGE[f"{name}__element"])
    cells = tuple(int(c) for c in LARGE[f"{name}__cells"])
    weights = LARGE[f"{name}__weights"].astype(np.float64)
    xyz, topo = oe.beam_mesh(el, cells, SIZE)
    m = oe.host_masses(el, xyz, topo, 1000.0, oe.MASS_FEMSHAPE)
    assert m.shape == weights.shape
    # The reference builds J = sum_k X_k (x) dN_k in float from coordinates of
    # magnitude up to 4 + |z offset| at element size h, so det J (and the
    # weights) carry a relative rounding error up to about
    # 3 (4 + |z|) 2^-24 / h; the files also keep 6 significant digits.
    h = SIZE[0] / cells[0]
    rel = max(5e-6, 3.0 * (4.0 + abs(float(LARGE[f"{name}__zoff"]))) * 2.0 ** -24 / h)
    assert np.abs(m - weights).max() <= rel * weights.max()
    mo = O.masses(el, xyz, topo, 1000.0, True)
    assert np.abs(mo - weights).max() <= rel * weights.max()
