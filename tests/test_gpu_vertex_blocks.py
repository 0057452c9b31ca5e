"""Element-task vertex blocks at their limits: a vertex with more incident
elements than a task block holds (256 for tetrahedra, 32 for 27-node
hexahedra).  Explicit task tiles are refused; the default pass falls back to
the gather kernel; a vertex just under the limit runs in one block.  Results
are compared with the FP64 oracle (max vertex error <= 1e-5 of the bbox
diagonal)."""
import numpy as np
import pytest

import octopusengine_amd as oe
import pyoracle as O
from meshes import hex_book, p1_to_p2, tet_fan
from scenes import max_err

pytestmark = pytest.mark.gpu


def pair(engine, el, xyz, topo, fixed):
    g = oe.Scene(engine, el, xyz, topo, material=3, young=1e5, poisson=0.45, iterations=10)
    o = O.Scene(el, xyz, topo, material=3, young=1e5, poisson=0.45, iterations=10)
    g.fix_vertices(fixed)
    o.fix(fixed)
    return g, o


def check(g, o, xyz, steps=2):
    for _ in range(steps):
        g.step(0.005, 1)
        o.step(0.005, 1)
    diag = float(np.linalg.norm(xyz.max(0) - xyz.min(0)))
    err = max_err(g.positions(), o.positions())
    assert err <= 1e-5 * diag, err
    return err


def test_p1_fan_task_tile_refused(engine):
    xyz, topo = tet_fan(300)                      # 600 tetrahedra at the origin
    g = oe.Scene(engine, oe.TETRA, xyz, topo, iterations=2)
    before = g.get_tile()
    with pytest.raises(oe.OEError):
        g.set_tile(64, 1, oe.PASS_TASKS)
    assert g.get_tile() == before
    g.step(0.005, 1)                               # still steps with the gather
    assert np.isfinite(g.positions()).all()


@pytest.mark.parametrize("m,tasks", [(300, False), (100, True)])
def test_p2_fan(engine, m, tasks):
    xyz, topo = p1_to_p2(*tet_fan(m))             # 2 m quadratic tetrahedra at the origin
    fixed = np.nonzero(xyz[:, 2] < -0.4)[0].astype(np.int32)
    g, o = pair(engine, oe.TETRA10, xyz, topo, fixed)
    assert (g.get_tile()[2] == oe.PASS_TASKS) == tasks
    err = check(g, o, xyz)
    print(f"P2 fan of {2 * m}: tile {g.get_tile()} err {err:.2e}", flush=True)


@pytest.mark.parametrize("m,tasks", [(40, False), (16, True)])
def test_q2_book(engine, m, tasks):
    xyz, topo = hex_book(m)                       # m hexahedra around the z axis
    fixed = np.nonzero(xyz[:, 2] < 1e-6)[0].astype(np.int32)
    g, o = pair(engine, oe.HEXA27, xyz, topo, fixed)
    assert (g.get_tile()[2] == oe.PASS_TASKS) == tasks
    err = check(g, o, xyz)
    print(f"Q2 book of {m}: tile {g.get_tile()} err {err:.2e}", flush=True)
