"""The solvers on the reference's own (non-beam) meshes, read from files:
curved P2 bunny, curved P3 sphere, ParaView-written Q1 armadillo, Gmsh P1 bar.
Each GPU result is compared with the FP64 oracle on the same mesh.  Tolerance:
max vertex-position error <= 1e-5 of the mesh's bounding-box diagonal."""
from pathlib import Path

import numpy as np
import pytest

import octopusengine_amd as oe
import pyoracle as O
from scenes import max_err

pytestmark = pytest.mark.gpu

ASSETS = np.load(Path(__file__).resolve().parent / "golden" / "mesh_assets.npz")


def load(tmp_path, name):
    p = tmp_path / name
    p.write_bytes(ASSETS[name].tobytes())
    xyz, topos = oe.load_mesh(p)
    (el, topo), = topos.items()
    return el, xyz, topo


def pair(engine, el, xyz, topo, **kw):
    """Engine and oracle scenes on one mesh; the lowest 5% (in y) fixed."""
    lo, hi = xyz.min(0), xyz.max(0)
    ids = np.nonzero(xyz[:, 1] < lo[1] + 0.05 * (hi[1] - lo[1]))[0].astype(np.int32)
    tile = kw.pop("tile", None)
    g = oe.Scene(engine, el, xyz, topo, material=kw.get("material", 3), young=kw.get("young", 1e5),
                 poisson=kw.get("poisson", 0.45), solver=kw.get("solver", 0), iterations=kw.get("iterations", 10),
                 xpbd_version=kw.get("xpbd_version", 2))
    o = O.Scene(el, xyz, topo, material=kw.get("material", 3), young=kw.get("young", 1e5),
                poisson=kw.get("poisson", 0.45), solver=kw.get("solver", 0), iterations=kw.get("iterations", 10),
                xpbd_version=kw.get("xpbd_version", 2))
    g.fix_vertices(ids)
    o.fix(ids)
    g.box_limit()
    o.box()
    if tile:
        g.set_tile(*tile)
    diag = float(np.linalg.norm(hi - lo))
    return g, o, diag


@pytest.mark.parametrize("cache", [False, True])
def test_vbd_bunny_p2(engine, tmp_path, cache):
    el, xyz, topo = load(tmp_path, "bunny_P2.vtk")
    g, o, diag = pair(engine, el, xyz, topo, tile=(16, 1, cache))
    assert not g.affine()
    for _ in range(2):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    print(f"bunny P2 cache={cache} 2 steps err={err:.3e} (diag {diag:.3f})", flush=True)
    assert err <= 1e-5 * diag


def test_vbd_armadillo_q1(engine, tmp_path):
    el, xyz, topo = load(tmp_path, "armadillo3.vtk")
    g, o, diag = pair(engine, el, xyz, topo)
    for _ in range(2):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    print(f"armadillo Q1 err={err:.3e}", flush=True)
    assert err <= 1e-5 * diag


def test_xpbd_sphere_p3(engine, tmp_path):
    el, xyz, topo = load(tmp_path, "sphere_P3.vtk")
    g, o, diag = pair(engine, el, xyz, topo, solver=1, material=2, young=1e6, xpbd_version=2)
    g.step(0.01, 20)
    o.step(0.01, 20)
    err = max_err(g.positions(), o.positions())
    print(f"sphere P3 XPBD err={err:.3e}", flush=True)
    assert err <= 1e-5 * diag


def test_explicit_gmsh_bar_p1(engine, tmp_path):
    el, xyz, topo = load(tmp_path, "bar_tetra_864.msh")
    g, o, diag = pair(engine, el, xyz, topo, solver=2, material=2, young=1e6, iterations=1)
    g.step(1e-3, 50)
    o.step(1e-3, 50)
    err = max_err(g.positions(), o.positions())
    print(f"bar P1 explicit err={err:.3e}", flush=True)
    assert err <= 1e-5 * diag
