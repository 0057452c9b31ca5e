"""Writes tests/golden/vtk_vbd_scaling.npz from the reference's own result
files (data only, nothing of the reference is run).

Source: assets/result/vtk/VBD Scaling/{1,2}_<Element>_<nx>_<ny>_<nz>_4x1x1_*.vtk
in the reference repository: beams of size 4 x 1 x 1 written by its
VTK_Formater (src/Mesh/Converter/VTK_Formater.cpp).  For each file the fixture
keeps the element, the cell counts, the exact text of the mesh sections
(DATASET .. CELL_TYPES values) and the SCALARS "weights" point data (the
per-vertex masses, density 1000, FemShape distribution).

The larger files (3_*, 4_* and the ref_* solutions, up to the 18,785-vertex
Q2 / P2 32 x 8 x 8 and Q1 64 x 16 x 16 beams) go to
vtk_vbd_scaling_large.npz in a compact form: the SHA-256 of the mesh sections,
the z offset of the beam in its scene (first point's z), and the weights as
float32.

Run in a container that has the reference at /root/reference:
    python tests/golden/make_vtk_fixtures.py
"""
import hashlib
import re
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/assets/result/vtk/VBD Scaling")
OUT = Path(__file__).resolve().parent / "vtk_vbd_scaling.npz"
OUT_LARGE = Path(__file__).resolve().parent / "vtk_vbd_scaling_large.npz"
NAME_RE = r"(\d|ref)_(Tetra10|Tetra20|Tetra|Hexa27|Hexa)_(\d+)_(\d+)_(\d+)_4x1x1"
ELEMENTS = {"Tetra": 3, "Hexa": 6, "Tetra10": 7, "Tetra20": 8, "Hexa27": 9}


def parse(path: Path):
    txt = path.read_text()
    lines = txt.split("\n")
    start = next(i for i, l in enumerate(lines) if l.startswith("DATASET"))
    ct = next(i for i, l in enumerate(lines) if l.startswith("CELL_TYPES"))
    mesh = "\n".join(lines[start:ct + 2]) + "\n"
    w0 = next(i for i, l in enumerate(lines) if l.startswith("SCALARS weights"))
    weights = np.array([float(v) for v in lines[w0 + 2].split()], np.float64)
    return mesh, weights


def main():
    data = {}
    names = []
    for f in sorted(SRC.glob("[12]_*_4x1x1_*.vtk")):
        m = re.match(r"(\d)_(Tetra10|Tetra20|Tetra|Hexa27|Hexa)_(\d+)_(\d+)_(\d+)_4x1x1_", f.name)
        if not m:
            continue
        el = ELEMENTS[m.group(2)]
        cells = np.array([int(m.group(3)), int(m.group(4)), int(m.group(5))], np.int32)
        mesh, weights = parse(f)
        key = f"{m.group(2)}_{cells[0]}_{cells[1]}_{cells[2]}"
        names.append(key)
        data[f"{key}__element"] = np.int32(el)
        data[f"{key}__cells"] = cells
        data[f"{key}__mesh"] = np.frombuffer(mesh.encode(), np.uint8)
        data[f"{key}__weights"] = weights
    data["names"] = np.array(names)
    np.savez_compressed(OUT, **data)
    print(OUT, names)

    large = {}
    lnames = []
    for f in sorted(SRC.glob("*_4x1x1*.vtk")):
        m = re.match(NAME_RE, f.name)
        if not m or m.group(1) in ("1", "2"):
            continue
        el = ELEMENTS[m.group(2)]
        cells = np.array([int(m.group(3)), int(m.group(4)), int(m.group(5))], np.int32)
        mesh, weights = parse(f)
        key = f"{m.group(1)}_{m.group(2)}_{cells[0]}_{cells[1]}_{cells[2]}"
        lnames.append(key)
        first_point = mesh.split("\n")[2].split()
        large[f"{key}__element"] = np.int32(el)
        large[f"{key}__cells"] = cells
        large[f"{key}__zoff"] = np.float64(float(first_point[2]))
        large[f"{key}__mesh_sha256"] = np.array(hashlib.sha256(mesh.encode()).hexdigest())
        large[f"{key}__weights"] = weights.astype(np.float32)
    large["names"] = np.array(lnames)
    np.savez_compressed(OUT_LARGE, **large)
    print(OUT_LARGE, lnames)


if __name__ == "__main__":
    main()
