"""Writes tests/golden/mesh_assets.npz: four of the reference's mesh data files
(assets/mesh/vtk/bunny_P2.vtk, sphere_P3.vtk, armadillo3.vtk and
assets/mesh/msh/bar_tetra_864.msh), stored byte for byte (compressed) so that
the loader and writer tests read exactly what the reference reads.

Run in a container that has the reference at /root/reference:
    python tests/golden/make_mesh_fixtures.py
"""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/assets/mesh")
FILES = ["vtk/bunny_P2.vtk", "vtk/sphere_P3.vtk", "vtk/armadillo3.vtk", "msh/bar_tetra_864.msh"]
OUT = Path(__file__).resolve().parent / "mesh_assets.npz"


def main():
    data = {Path(f).name: np.frombuffer((SRC / f).read_bytes(), np.uint8) for f in FILES}
    np.savez_compressed(OUT, **data)
    print(OUT, {k: v.size for k, v in data.items()})


if __name__ == "__main__":
    main()
